"""Multi-start SE(3) seeds of the alignment optimiser, batched over pairs (host, numpy FP64).

Restates ``_initialize_se3_params`` (shepherd_score/alignment/_torch.py:253-326):
  * num_repeats == 1: identity rotation, zero translation;
  * 1 < num_repeats < 5 is raised to 5;
  * seed 0: identity rotation, translation ref_com - fit_com;
  * seeds 1-4: the four principal-axis alignments
    (alignment/utils/pca_np.py:111-166, moments :12-33);
  * seeds 5+: rotations taking +x to Fibonacci-sphere points (_torch.py:173-196,
    generate_point_cloud.py:234-261; the 50-restart table of _torch.py:199-250 is this same
    construction for 45 points), translation ref_com - R fit_com.
Seeds are returned as float32 (q, t) arrays, the precision the reference stores them in.
"""
from __future__ import annotations

import numpy as np


def fibonacci_sphere(n: int) -> np.ndarray:
    """n points on the unit sphere by the golden-spiral construction."""
    offset = 2.0 / n
    inc = np.pi * (3.0 - np.sqrt(5.0))
    i = np.arange(n)
    y = (i * offset - 1.0) + offset / 2.0
    r = np.sqrt(1.0 - y * y)
    phi = np.mod(i + 1, n) * inc
    return np.stack([np.cos(phi) * r, y, np.sin(phi) * r], axis=1)


def quats_from_fibo(n: int) -> np.ndarray:
    """Quaternions (w,x,y,z) rotating +x onto each of n Fibonacci-sphere points."""
    pts = fibonacci_sphere(n)
    u = np.array([1.0, 0.0, 0.0])
    ang = np.arccos(np.clip(pts @ u, -1.0, 1.0))
    same = np.isclose(pts, u[None, :]).sum(1) == 3
    ax = np.cross(np.broadcast_to(u, pts.shape), pts)
    ax[same] = u
    ax = ax / np.maximum(np.linalg.norm(ax, axis=1, keepdims=True), 1e-12)
    return np.concatenate([np.cos(ang / 2)[:, None], ax * np.sin(ang / 2)[:, None]], axis=1)


def rotmat(q: np.ndarray) -> np.ndarray:
    """(..., 4) quaternion (normalised here) -> (..., 3, 3) rotation."""
    q = q / np.linalg.norm(q, axis=-1, keepdims=True)
    w, x, y, z = q[..., 0], q[..., 1], q[..., 2], q[..., 3]
    R = np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w),
                  2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w),
                  2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)], axis=-1)
    return R.reshape(q.shape[:-1] + (3, 3))


def _principal_axes(P: np.ndarray, counts: np.ndarray) -> np.ndarray:
    """Principal axes (rows, longest first) of the inertia tensor of each padded cloud."""
    n, L = P.shape[:2]
    mask = (np.arange(L)[None, :] < counts[:, None]).astype(np.float64)[..., None]
    com = (P * mask).sum(1) / counts[:, None]
    X = (P - com[:, None, :]) * mask
    A = (X ** 2).sum((1, 2))
    Bm = np.einsum("npi,npj->nij", X, X)
    I = (A[:, None, None] * np.eye(3)[None] - Bm) / counts[:, None, None]
    w, v = np.linalg.eigh(I)
    order = np.argsort(-w, axis=1)
    v = np.take_along_axis(v, order[:, None, :], axis=2)
    return np.transpose(v, (0, 2, 1))


def _axis_angle_quat(v1, v2):
    """Quaternion rotating v1 onto v2 (batched rows), the axis/angle recipe of pca_np.py:36-88."""
    v1u = v1 / np.linalg.norm(v1, axis=1, keepdims=True)
    v2u = v2 / np.linalg.norm(v2, axis=1, keepdims=True)
    ang = np.arccos(np.clip((v1u * v2u).sum(1), -1.0, 1.0))
    ax = np.cross(v1, v2)
    close = np.all(np.isclose(v1, v2), axis=1)
    ax[close] = np.array([1.0, 0.0, 0.0])
    ax = ax / np.linalg.norm(ax, axis=1, keepdims=True)
    return np.concatenate([np.cos(ang / 2)[:, None], ax * np.sin(ang / 2)[:, None]], axis=1)


def _qmul(p, q):
    """Hamilton product p*q of row quaternions."""
    w1, x1, y1, z1 = p.T
    w2, x2, y2, z2 = q.T
    return np.stack([w1 * w2 - x1 * x2 - y1 * y2 - z1 * z2,
                     x1 * w2 + w1 * x2 - z1 * y2 + y1 * z2,
                     y1 * w2 + z1 * x2 + w1 * y2 - x1 * z2,
                     z1 * w2 - y1 * x2 + x1 * y2 + w1 * z2], axis=1)


def pca_quaternions(ref: np.ndarray, ref_n: np.ndarray, fit: np.ndarray, fit_n: np.ndarray) -> np.ndarray:
    """(n, 4, 4): the four principal-component alignment quaternions of every pair."""
    n = ref.shape[0]
    pmi_ref = _principal_axes(ref, ref_n)
    fit_mask = (np.arange(fit.shape[1])[None, :] < fit_n[:, None]).astype(np.float64)[..., None]
    fit_com = (fit * fit_mask).sum(1) / fit_n[:, None]
    out = np.zeros((n, 4, 4))
    flips = [(1, 1), (-1, 1), (1, -1), (-1, -1)]
    for qi, (f0, f1) in enumerate(flips):
        target = pmi_ref.copy()
        target[:, 0] *= f0
        target[:, 1] *= f1
        X = (fit - fit_com[:, None, :]) * fit_mask
        quats = []
        for ax in range(2):
            pmi_fit = _principal_axes(X, fit_n)
            qa = _axis_angle_quat(pmi_fit[:, ax], target[:, ax])
            quats.append(qa)
            R = rotmat(qa)
            X = np.einsum("nij,npj->npi", R, X) * fit_mask
        out[:, qi] = _qmul(quats[1], quats[0])
    return out


def initial_seeds(ref: np.ndarray, ref_n: np.ndarray, fit: np.ndarray, fit_n: np.ndarray,
                  num_repeats: int):
    """Seeds for n pairs given padded clouds (n, L, 3) and real counts.

    Returns (q (n, S, 4) float32, t (n, S, 3) float32) with S the effective restart count.
    """
    ref = np.asarray(ref, np.float64)
    fit = np.asarray(fit, np.float64)
    ref_n = np.asarray(ref_n, np.int64)
    fit_n = np.asarray(fit_n, np.int64)
    n = ref.shape[0]
    S = int(num_repeats)
    if S < 1:
        raise ValueError("num_repeats must be >= 1")
    if 1 < S < 5:
        S = 5
    q = np.zeros((n, S, 4))
    t = np.zeros((n, S, 3))
    q[:, :, 0] = 1.0
    if S == 1:
        return q.astype(np.float32), t.astype(np.float32)
    rm = (np.arange(ref.shape[1])[None, :] < ref_n[:, None])[..., None]
    fm = (np.arange(fit.shape[1])[None, :] < fit_n[:, None])[..., None]
    ref_com = (ref * rm).sum(1) / ref_n[:, None]
    fit_com = (fit * fm).sum(1) / fit_n[:, None]
    q[:, 1:5] = pca_quaternions(ref, ref_n, fit, fit_n)
    if S > 5:
        q[:, 5:] = quats_from_fibo(S - 5)[None]
    # The reference keeps the seed quaternions in float32 and builds R from them.
    q = q.astype(np.float32).astype(np.float64)
    R = rotmat(q)
    t[:] = ref_com[:, None, :] - np.einsum("nsij,nj->nsi", R, fit_com)
    t[:, 0] = ref_com - fit_com
    return q.astype(np.float32), t.astype(np.float32)


def translation_seeds(ref: np.ndarray, fit: np.ndarray, centers: np.ndarray, per_trans: int = 10):
    """Seeds of one pair with translation centres (trans_init=True), restating
    ``_initialize_se3_params_with_translations`` (alignment/_torch.py:329-398):
      * seed 0: identity rotation, translation ref_com - fit_com;
      * seeds 1-4: the four principal-axis alignments, translation ref_com - R fit_com;
      * then per_trans Fibonacci-sphere rotations (the 45-point table when per_trans == 45),
        repeated for every centre c (centre-major), translation c - R fit_com: the rotated fit's
        centroid lands on the centre.
    Returns (q (5 + per_trans * C, 4), t (same, 3)) float32; the restart count ignores
    num_repeats, as in the reference."""
    ref = np.asarray(ref, np.float64).reshape(-1, 3)
    fit = np.asarray(fit, np.float64).reshape(-1, 3)
    centers = np.asarray(centers, np.float64).reshape(-1, 3)
    C = len(centers)
    S = 5 + int(per_trans) * C
    q = np.zeros((S, 4))
    q[:, 0] = 1.0
    q[1:5] = pca_quaternions(ref[None], np.array([len(ref)]), fit[None], np.array([len(fit)]))[0]
    if C:
        q[5:] = np.tile(quats_from_fibo(int(per_trans)), (C, 1))
    q = q.astype(np.float32).astype(np.float64)      # stored in float32, R built from those
    R = rotmat(q)
    rc, fc = ref.mean(0), fit.mean(0)
    t = np.empty((S, 3))
    t[0] = rc - fc
    t[1:5] = rc[None] - R[1:5] @ fc
    t[5:] = np.repeat(centers, int(per_trans), axis=0) - R[5:] @ fc
    return q.astype(np.float32), t.astype(np.float32)


def effective_restarts(num_repeats: int) -> int:
    """The reference's restart count rule: 1 stays 1, 2-4 become 5 (_torch.py:253-326)."""
    S = int(num_repeats)
    if S < 1:
        raise ValueError("num_repeats must be >= 1")
    return 5 if 1 < S < 5 else S


def reference_table(S: int) -> np.ndarray:
    """(S, 4) float32 rotation table of the reference rule: identity rows 0-4 (row 0 stays the
    identity; the device fills rows 1-4 with the principal-axis seeds) and the Fibonacci-sphere
    rotations (the 45-entry table at S == 50) from row 5 on."""
    q = np.zeros((S, 4), np.float32)
    q[:, 0] = 1.0
    if S > 5:
        q[5:] = quats_from_fibo(S - 5).astype(np.float32)
    return q


def device_seeds(pb, num_repeats: int, use_pharm: bool = False, stream=None, use_surface: bool = False):
    """Seeds of every pair of a PairBatch on its device (fss_seed_poses with the reference rule):
    (q (n, S, 4), t (n, S, 3)) float32 tensors, S = effective_restarts(num_repeats).  The
    centroids and principal axes come from the pharmacophore anchors when use_pharm, from the
    ShaEP channel's surface points when use_surface."""
    import torch
    from . import _lib
    S = effective_restarts(num_repeats)
    dev = pb.device
    table = torch.from_numpy(reference_table(S)).to(dev)
    q = torch.empty((pb.n_pairs, S, 4), dtype=torch.float32, device=dev)
    t = torch.empty((pb.n_pairs, S, 3), dtype=torch.float32, device=dev)
    if pb.n_pairs:
        st = stream if stream is not None else torch.cuda.current_stream(dev)
        flags = _lib.SEED_REF | (_lib.SEED_PHARM if use_pharm else 0) | (_lib.SEED_SURF if use_surface else 0)
        bs = pb.struct()
        _lib.call("fss_seed_poses", _lib.byref(bs), flags, S, _lib.ptr(table), _lib.ptr(q),
                  _lib.ptr(t), int(st.cuda_stream))
    return q, t


def pad(clouds, width: int = 3):
    """List of (k, width) arrays -> padded (n, L, width) float64 array and counts."""
    counts = np.array([len(c) for c in clouds], dtype=np.int64)
    L = max(1, int(counts.max()) if len(counts) else 1)
    out = np.zeros((len(clouds), L, width))
    for i, c in enumerate(clouds):
        if counts[i]:
            out[i, :counts[i]] = np.asarray(c, np.float64).reshape(-1, width)
    return out, counts
