"""The reference's accelerated kernel table on the HIP kernels.

Drop-in for the names ``shepherd_score.accel.kernels.dispatch`` exports
(accel/kernels/dispatch.py:96-112), with the same signatures, argument meaning and return values
as its Triton / numba implementations (shape_triton.py, esp_triton.py, pharm_triton.py, cpu.py):

  * ``overlap_score_grad_se3_batch``, ``overlap_score_grad_esp_se3_batch``,
    ``pharm_grad_dq_se3_batch``, ``pharm_color_score_grad_se3_batch`` -> (V, dQ, dT)
  * ``pharm_score_grad_se3_batch`` (takes rotation matrices) -> (O, dR, dT)
  * ``esp_comparison_batch`` -> the ShaEP ESP term of one direction (K,)
  * ``_batch_self_overlap``, ``_batch_self_overlap_esp`` -> self-overlaps (K,)
  * ``fused_adam_qt``, ``fused_adam_qt_with_tangent_proj`` -> in-place Adam on (q, t)

Inputs are padded CUDA tensors with optional real counts ``N_real`` / ``M_real`` (rows past them
are ignored), as in the reference.  The padded rows are handed to the kernels in place as
(row, count) ranges, so no packing copy is made beyond a float4 view.  dQ is the gradient
through the unit-quaternion rotation at q / |q| (the drivers keep q unit).  ``BLOCK``,
``num_warps`` and ``num_stages`` are accepted and ignored, as the reference's CPU kernels do.
Results are float32 tensors on the input's device (cast to the input's dtype).  There is no
CPU path: the HIP library must be present.
"""
from __future__ import annotations

from typing import Optional

import torch

from . import _lib
from .align import ScoreSpec
from .constants import LAM_SCALING

__all__ = ["overlap_score_grad_se3_batch", "overlap_score_grad_esp_se3_batch",
           "pharm_grad_dq_se3_batch", "pharm_color_score_grad_se3_batch",
           "pharm_score_grad_se3_batch", "esp_comparison_batch", "_batch_self_overlap",
           "_batch_self_overlap_esp", "fused_adam_qt", "fused_adam_qt_with_tangent_proj"]


def _cuda(x: torch.Tensor) -> torch.Tensor:
    if not x.is_cuda:
        raise RuntimeError("the HIP kernel table takes device tensors")
    return x


def _rows(X: torch.Tensor, w: Optional[torch.Tensor] = None) -> torch.Tensor:
    """(K, N, 3) [+ (K, N)] -> (K * N, 4) float32 rows (x, y, z, w)."""
    K, N = X.shape[0], X.shape[1]
    col = (w.reshape(K, N, 1) if w is not None else torch.zeros((K, N, 1), device=X.device))
    return torch.cat([X.to(torch.float32), col.to(torch.float32)], dim=-1).reshape(K * N, 4).contiguous()


def _ranges(K: int, N: int, real, dev) -> torch.Tensor:
    """(K, 2) int32 [first row, count] of K padded clouds of N rows."""
    first = torch.arange(K, dtype=torch.int32, device=dev) * N
    cnt = (torch.full((K,), N, dtype=torch.int32, device=dev) if real is None
           else real.to(device=dev, dtype=torch.int32).clamp(0, N))
    return torch.stack([first, cnt], dim=1).contiguous()


def _max(real, N: int) -> int:
    return N if real is None else int(real.max().item()) if real.numel() else 0


def _params(mode: int, alpha: float = 0.81, lam: float = 1.0, alphas=None, Ks=None, cats=None,
            directionless: bool = False) -> _lib.FssScoreParams:
    sp = ScoreSpec(mode=mode, alpha=float(alpha), lam=float(lam)).struct()
    if alphas is not None:
        a = [float(x) for x in alphas.detach().cpu().reshape(-1)]
        k = [float(x) for x in Ks.detach().cpu().reshape(-1)]
        c = [int(x) for x in cats.detach().cpu().reshape(-1)]
        if not (len(a) == len(k) == len(c) <= _lib.N_PTYPES):
            raise ValueError(f"per-type tables hold at most {_lib.N_PTYPES} types")
        for i in range(_lib.N_PTYPES):
            if i < len(a):
                sp.ph_alpha[i], sp.ph_k[i], sp.ph_cat[i] = a[i], k[i], c[i]
            else:
                sp.ph_alpha[i], sp.ph_k[i], sp.ph_cat[i] = 1.0, 1.0, 3
            if directionless and sp.ph_cat[i] != 3:
                sp.ph_cat[i] = 0
    return sp


def _run(pb: _lib.FssPairBatch, sp, q, t, need_grad: bool, keep):
    K = pb.n_pairs
    dev = q.device
    V = torch.zeros((K, 2), dtype=torch.float32, device=dev)
    gR = torch.zeros((K, 2, 3, 3), dtype=torch.float32, device=dev) if need_grad else None
    gT = torch.zeros((K, 2, 3), dtype=torch.float32, device=dev) if need_grad else None
    gQ = torch.zeros((K, 2, 4), dtype=torch.float32, device=dev) if need_grad else None
    pp = torch.arange(K, dtype=torch.int32, device=dev)
    qq = q.to(torch.float32).contiguous()
    tt = t.to(torch.float32).contiguous()
    keep += [pp, qq, tt]
    _lib.call("fss_overlap_grad", _lib.byref(pb), _lib.byref(sp), K, _lib.ptr(pp), _lib.ptr(qq),
              _lib.ptr(tt), int(need_grad), _lib.ptr(V), _lib.ptr(gR), _lib.ptr(gT), _lib.ptr(gQ),
              int(torch.cuda.current_stream().cuda_stream))
    return V, gR, gT, gQ


def _zeros_like_grad(K, dev, dt):
    return torch.zeros((K, 4), device=dev, dtype=dt), torch.zeros((K, 3), device=dev, dtype=dt)


def _shape_batch(A, B, wA, wB, N_real, M_real, keep) -> _lib.FssPairBatch:
    K, N, M = A.shape[0], A.shape[1], B.shape[1]
    ra, fb = _rows(A, wA), _rows(B, wB)
    rr, fr = _ranges(K, N, N_real, A.device), _ranges(K, M, M_real, A.device)
    keep += [ra, fb, rr, fr]
    pb = _lib.FssPairBatch()
    pb.n_pairs = K
    pb.ref_pts, pb.fit_pts, pb.ref_rng, pb.fit_rng = _lib.ptr(ra), _lib.ptr(fb), _lib.ptr(rr), _lib.ptr(fr)
    pb.max_ref, pb.max_fit = max(_max(N_real, N), 1), max(_max(M_real, M), 1)
    return pb


# ------------------------------------------------------------------------------ shape / ESP
def overlap_score_grad_se3_batch(A, B, q, t, *, alpha: float = 0.81, N_real=None, M_real=None,
                                 NEED_GRAD: bool = True, BLOCK=None, num_warps=None, num_stages=None):
    """V = K sum exp(-alpha/2 |A - (R(q) B + t)|^2) per pose, dV/dq (K, 4), dV/dt (K, 3)
    (shape_triton.py:191-249, cpu.py:38-110)."""
    A, B = _cuda(A), _cuda(B)
    keep = []
    pb = _shape_batch(A, B, None, None, N_real, M_real, keep)
    V, _, gT, gQ = _run(pb, _params(_lib.MODE_SHAPE, alpha), q, t, NEED_GRAD, keep)
    dt = A.dtype
    if not NEED_GRAD:
        return (V[:, 0].to(dt),) + _zeros_like_grad(A.shape[0], A.device, dt)
    return V[:, 0].to(dt), gQ[:, 0].to(dt), gT[:, 0].to(dt)


def overlap_score_grad_esp_se3_batch(A, B, charges_A, charges_B, q, t, *, alpha: float = 0.81,
                                     lam: float = 0.3, N_real=None, M_real=None, NEED_GRAD: bool = True,
                                     BLOCK=None, num_warps=None, num_stages=None):
    """The overlap with weights exp(-(c_A - c_B)^2 / lam), ``lam`` as given
    (esp_triton.py:175-246, cpu.py:159-234)."""
    A, B = _cuda(A), _cuda(B)
    keep = []
    pb = _shape_batch(A, B, charges_A, charges_B, N_real, M_real, keep)
    V, _, gT, gQ = _run(pb, _params(_lib.MODE_SHAPE_ESP, alpha, lam), q, t, NEED_GRAD, keep)
    dt = A.dtype
    if not NEED_GRAD:
        return (V[:, 0].to(dt),) + _zeros_like_grad(A.shape[0], A.device, dt)
    return V[:, 0].to(dt), gQ[:, 0].to(dt), gT[:, 0].to(dt)


def _identity(K, dev):
    q = torch.zeros((K, 4), dtype=torch.float32, device=dev)
    q[:, 0] = 1.0
    return q, torch.zeros((K, 3), dtype=torch.float32, device=dev)


@torch.no_grad()
def _batch_self_overlap(P_pad, N_real, alpha: float = 0.81):
    """Self-overlap of each padded cloud at the identity pose (shape_triton.py, cpu.py:141-156)."""
    q, t = _identity(P_pad.shape[0], P_pad.device)
    return overlap_score_grad_se3_batch(P_pad, P_pad, q, t, alpha=alpha, N_real=N_real, M_real=N_real,
                                        NEED_GRAD=False)[0]


@torch.no_grad()
def _batch_self_overlap_esp(P_pad, charges, N_real, alpha: float = 0.81, lam: float = 0.3):
    """ESP-weighted self-overlap at the identity pose (esp_triton.py, cpu.py:616-626)."""
    q, t = _identity(P_pad.shape[0], P_pad.device)
    return overlap_score_grad_esp_se3_batch(P_pad, P_pad, charges, charges, q, t, alpha=alpha, lam=lam,
                                            N_real=N_real, M_real=N_real, NEED_GRAD=False)[0]


# ------------------------------------------------------------------------------ pharmacophores
def _pharm_batch(ref_types, fit_types, ref_anchors, fit_anchors, ref_vectors, fit_vectors,
                 N_real, M_real, keep) -> _lib.FssPairBatch:
    K, N, M = ref_anchors.shape[0], ref_anchors.shape[1], fit_anchors.shape[1]
    dev = ref_anchors.device
    ra, fa = _rows(ref_anchors), _rows(fit_anchors)
    rv = _rows(ref_vectors) if ref_vectors is not None else torch.zeros_like(ra)
    fv = _rows(fit_vectors) if fit_vectors is not None else torch.zeros_like(fa)
    rt = ref_types.to(device=dev, dtype=torch.int32).reshape(-1).contiguous()
    ft = fit_types.to(device=dev, dtype=torch.int32).reshape(-1).contiguous()
    rr, fr = _ranges(K, N, N_real, dev), _ranges(K, M, M_real, dev)
    keep += [ra, fa, rv, fv, rt, ft, rr, fr]
    pb = _lib.FssPairBatch()
    pb.n_pairs = K
    pb.ref_anc, pb.ref_vec, pb.ref_typ, pb.ref_prng = _lib.ptr(ra), _lib.ptr(rv), _lib.ptr(rt), _lib.ptr(rr)
    pb.fit_anc, pb.fit_vec, pb.fit_typ, pb.fit_prng = _lib.ptr(fa), _lib.ptr(fv), _lib.ptr(ft), _lib.ptr(fr)
    pb.max_ref_ph, pb.max_fit_ph = max(_max(N_real, N), 1), max(_max(M_real, M), 1)
    return pb


def pharm_grad_dq_se3_batch(q, t, ref_types, fit_types, ref_anchors, fit_anchors, ref_vectors,
                            fit_vectors, alphas, Ks, cats, *, N_real=None, M_real=None,
                            NEED_GRAD: bool = True, BLOCK=None, num_warps=None, num_stages=None):
    """Typed directional pharmacophore overlap, dO/dq and dO/dt; per-type alpha, K and category
    from the fit feature (pharm_triton.py:296-430, cpu.py:501-613)."""
    _cuda(ref_anchors)
    keep = []
    pb = _pharm_batch(ref_types, fit_types, ref_anchors, fit_anchors, ref_vectors, fit_vectors,
                      N_real, M_real, keep)
    V, _, gT, gQ = _run(pb, _params(_lib.MODE_PHARM, alphas=alphas, Ks=Ks, cats=cats), q, t, NEED_GRAD, keep)
    dt = ref_anchors.dtype
    if not NEED_GRAD:
        return (V[:, 1].to(dt),) + _zeros_like_grad(ref_anchors.shape[0], ref_anchors.device, dt)
    return V[:, 1].to(dt), gQ[:, 1].to(dt), gT[:, 1].to(dt)


def pharm_color_score_grad_se3_batch(A, B, q, t, ref_types, fit_types, alphas, Ks, cats, *,
                                     N_real=None, M_real=None, NEED_GRAD: bool = True, BLOCK=None,
                                     num_warps=None, num_stages=None):
    """Directionless colour overlap (every real type isotropic), dO/dq, dO/dt
    (pharm_triton.py:153-286, cpu.py:413-498)."""
    _cuda(A)
    keep = []
    pb = _pharm_batch(ref_types, fit_types, A, B, None, None, N_real, M_real, keep)
    sp = _params(_lib.MODE_PHARM, alphas=alphas, Ks=Ks, cats=cats, directionless=True)
    V, _, gT, gQ = _run(pb, sp, q, t, NEED_GRAD, keep)
    dt = A.dtype
    if not NEED_GRAD:
        return (V[:, 1].to(dt),) + _zeros_like_grad(A.shape[0], A.device, dt)
    return V[:, 1].to(dt), gQ[:, 1].to(dt), gT[:, 1].to(dt)


def _quat_from_rot(R: torch.Tensor) -> torch.Tensor:
    """Unit quaternions (w, x, y, z) of rotation matrices (K, 3, 3), by the largest of the four
    squared components (Shepperd's method), in float64."""
    R = R.to(torch.float64)
    m00, m11, m22 = R[:, 0, 0], R[:, 1, 1], R[:, 2, 2]
    tr = m00 + m11 + m22
    cand = torch.stack([1 + tr, 1 + m00 - m11 - m22, 1 - m00 + m11 - m22, 1 - m00 - m11 + m22], dim=1)
    k = cand.argmax(dim=1)
    s = torch.sqrt(cand.gather(1, k[:, None]).clamp_min(1e-300))[:, 0] * 2.0   # 4 |q_k|
    q = torch.empty((R.shape[0], 4), dtype=torch.float64, device=R.device)
    w = torch.stack([s / 4, (R[:, 2, 1] - R[:, 1, 2]) / s, (R[:, 0, 2] - R[:, 2, 0]) / s,
                     (R[:, 1, 0] - R[:, 0, 1]) / s], dim=1)
    x = torch.stack([(R[:, 2, 1] - R[:, 1, 2]) / s, s / 4, (R[:, 0, 1] + R[:, 1, 0]) / s,
                     (R[:, 0, 2] + R[:, 2, 0]) / s], dim=1)
    y = torch.stack([(R[:, 0, 2] - R[:, 2, 0]) / s, (R[:, 0, 1] + R[:, 1, 0]) / s, s / 4,
                     (R[:, 1, 2] + R[:, 2, 1]) / s], dim=1)
    z = torch.stack([(R[:, 1, 0] - R[:, 0, 1]) / s, (R[:, 0, 2] + R[:, 2, 0]) / s,
                     (R[:, 1, 2] + R[:, 2, 1]) / s, s / 4], dim=1)
    for idx, v in enumerate((w, x, y, z)):
        sel = k == idx
        q[sel] = v[sel]
    return q / q.norm(dim=1, keepdim=True)


def pharm_score_grad_se3_batch(R, t, ref_types, fit_types, ref_anchors, fit_anchors, ref_vectors,
                               fit_vectors, alphas, Ks, cats, *, N_real=None, M_real=None,
                               NEED_GRAD: bool = True, BLOCK=None, num_warps=1):
    """The directional overlap at rotation matrices R (K, 3, 3): (O, dO/dR (K, 3, 3), dO/dt)
    (pharm_triton.py:18-140, 433-467, cpu.py:296-410).  R is taken as the rotation of the
    quaternion nearest to it."""
    _cuda(ref_anchors)
    keep = []
    pb = _pharm_batch(ref_types, fit_types, ref_anchors, fit_anchors, ref_vectors, fit_vectors,
                      N_real, M_real, keep)
    q = _quat_from_rot(R).to(torch.float32)
    V, gR, gT, _ = _run(pb, _params(_lib.MODE_PHARM, alphas=alphas, Ks=Ks, cats=cats), q, t, NEED_GRAD, keep)
    dt = ref_anchors.dtype
    K = ref_anchors.shape[0]
    if not NEED_GRAD:
        return (V[:, 1].to(dt), torch.zeros((K, 3, 3), device=ref_anchors.device, dtype=dt),
                torch.zeros((K, 3), device=ref_anchors.device, dtype=dt))
    return V[:, 1].to(dt), gR[:, 1].to(dt), gT[:, 1].to(dt)


# ------------------------------------------------------------------------------ ShaEP ESP term
@torch.no_grad()
def esp_comparison_batch(points, atoms, charges, point_esp, radii, *, N_real=None, M_real=None,
                         probe_radius: float = 1.0, lam: float = 0.001, BLOCK=None, num_warps=None,
                         num_stages=None):
    """sum_i mask_i exp(-(pe_i - C sum_m q_m / d_im)^2 / (LAM_SCALING lam)) over the real
    observer points of each row against the other molecule's atoms; mask_i = 0 when point i
    lies within radius + probe of an atom (esp_triton.py:265-333, cpu.py:237-290)."""
    P = _cuda(points)
    K, N, M = P.shape[0], P.shape[1], atoms.shape[1]
    dev = P.device
    srf = _rows(P, point_esp)
    atm = _rows(atoms, charges)
    rad = radii.to(device=dev, dtype=torch.float32).reshape(-1).contiguous()
    sr, ar = _ranges(K, N, N_real, dev), _ranges(K, M, M_real, dev)
    none = torch.zeros((K, 2), dtype=torch.int32, device=dev)       # empty ranges
    keep = [srf, atm, rad, sr, ar, none]
    pb = _lib.FssPairBatch()
    pb.n_pairs = K
    # the shape rows (the observer points) only set the frame; the shape term is not used
    pb.ref_pts, pb.fit_pts, pb.ref_rng, pb.fit_rng = _lib.ptr(srf), _lib.ptr(srf), _lib.ptr(sr), _lib.ptr(none)
    pb.max_ref, pb.max_fit = max(_max(N_real, N), 1), 1
    pb.ref_srf, pb.ref_srng, pb.fit_srf, pb.fit_srng = _lib.ptr(srf), _lib.ptr(sr), _lib.ptr(srf), _lib.ptr(none)
    pb.ref_atm, pb.ref_rad, pb.ref_arng = _lib.ptr(atm), _lib.ptr(rad), _lib.ptr(none)
    pb.fit_atm, pb.fit_rad, pb.fit_arng = _lib.ptr(atm), _lib.ptr(rad), _lib.ptr(ar)
    pb.max_ref_srf, pb.max_fit_srf = max(_max(N_real, N), 1), 1
    pb.max_ref_atm, pb.max_fit_atm = 1, max(_max(M_real, M), 1)
    sp = ScoreSpec(mode=_lib.MODE_ESP_COMBO, alpha=0.81, lam=LAM_SCALING * float(lam),
                   probe_radius=float(probe_radius)).struct()
    q, t = _identity(K, dev)
    V, _, _, _ = _run(pb, sp, q, t, False, keep)
    return V[:, 1].to(P.dtype)


# ------------------------------------------------------------------------------ Adam
def _adam(q, t, dQ, dT, m_q, v_q, m_t, v_t, lr, beta1, beta2, eps, proj: bool):
    for x in (q, t, dQ, dT, m_q, v_q, m_t, v_t):
        if not (x.is_cuda and x.dtype == torch.float32 and x.is_contiguous()):
            raise TypeError("fused Adam takes contiguous float32 device tensors")
    n = int(q.shape[0])
    _lib.call("fss_adam_qt", n, _lib.ptr(q), _lib.ptr(t), _lib.ptr(dQ), _lib.ptr(dT), _lib.ptr(m_q),
              _lib.ptr(v_q), _lib.ptr(m_t), _lib.ptr(v_t), float(lr), float(beta1), float(beta2),
              float(eps), int(proj), int(torch.cuda.current_stream().cuda_stream))


@torch.no_grad()
def fused_adam_qt(q, t, dQ, dT, m_q, v_q, m_t, v_t, lr, beta1=0.9, beta2=0.999, eps=1e-8):
    """In-place Adam on (q, t): no bias correction, eps inside the square root, q renormalised
    (shape_triton.py:253-330, cpu.py:129-138)."""
    _adam(q, t, dQ, dT, m_q, v_q, m_t, v_t, lr, beta1, beta2, eps, False)


@torch.no_grad()
def fused_adam_qt_with_tangent_proj(q, t, dQ, dT, m_q, v_q, m_t, v_t, lr, beta1=0.9, beta2=0.999,
                                    eps=1e-8):
    """As :func:`fused_adam_qt` after projecting dQ onto the tangent space of the unit sphere,
    dQ - q (dQ . q) (shape_triton.py:333-406, cpu.py:112-126)."""
    _adam(q, t, dQ, dT, m_q, v_q, m_t, v_t, lr, beta1, beta2, eps, True)
