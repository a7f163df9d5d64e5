"""Differentiable Gaussian (ESP-weighted) shape similarity: score as a loss.

``gaussian_similarity(ref, fit, ...)`` returns the Tanimoto similarity
S = VAB / (VAA + VBB - VAB) of each pair of fixed clouds (score/gaussian_overlap.py:16-106,
score/electrostatic_scoring.py:17-125) with a hand-written backward: the HIP kernel
``fss_score_bwd`` evaluates

    dS/dx_b = K alpha / D^2 * (U sum_a w E (a - x_b) - 2 VAB sum_b' w E (x_b' - x_b)),
    U = VAA + VBB, D = U - VAB,

i.e. the full coordinate gradient including dVBB/dx (VBB is invariant only under rigid
motions), and the symmetric expression for the ref cloud.  FP32 inputs run the FP32 kernels,
FP64 inputs the FP64 instances (for gradcheck).  There is no CPU path.
"""
from __future__ import annotations

from typing import Optional

import torch

from . import _lib
from .constants import LAM_SCALING


def _pack(x: torch.Tensor, w: Optional[torch.Tensor]) -> torch.Tensor:
    P, N, _ = x.shape
    col = w.unsqueeze(-1) if w is not None else torch.zeros((P, N, 1), dtype=x.dtype, device=x.device)
    return torch.cat([x, col.to(x.dtype)], dim=-1).reshape(P * N, 4).contiguous()


_RANGES: dict = {}


def _ranges(P: int, N: int, dev) -> torch.Tensor:
    """(P, 2) int32 [first row, count] of P clouds of N rows each, built once per shape and device."""
    key = (P, N, str(dev))
    r = _RANGES.get(key)
    if r is None:
        if len(_RANGES) > 64:
            _RANGES.clear()
        st = torch.arange(P, dtype=torch.int32, device=dev) * N
        r = _RANGES[key] = torch.stack([st, torch.full_like(st, N)], dim=1).contiguous()
    return r


def _check_shapes(ref, fit, ref_w, fit_w):
    """Raise ValueError unless ref is (P, N, 3), fit (P, M, 3) and the ESP values (P, N), (P, M).
    The kernels index both clouds by the pair count and the row counts taken from these shapes,
    so a mismatch would read past the packed buffers on the device."""
    if ref.dim() != 3 or fit.dim() != 3 or ref.shape[-1] != 3 or fit.shape[-1] != 3:
        raise ValueError(f"ref and fit must be (P, N, 3) and (P, M, 3), got {tuple(ref.shape)} and "
                         f"{tuple(fit.shape)}")
    if ref.shape[0] != fit.shape[0]:
        raise ValueError(f"ref and fit hold {ref.shape[0]} and {fit.shape[0]} clouds: one fit cloud "
                         "per ref cloud")
    P, N, M = ref.shape[0], ref.shape[1], fit.shape[1]
    if ref_w is not None and tuple(ref_w.shape) != (P, N):
        raise ValueError(f"ref_esp must be ({P}, {N}), got {tuple(ref_w.shape)}")
    if fit_w is not None and tuple(fit_w.shape) != (P, M):
        raise ValueError(f"fit_esp must be ({P}, {M}), got {tuple(fit_w.shape)}")


def _forward(ref, fit, ref_w, fit_w, alpha, lam):
    """fss_score_fwd[_f64] on (P, N, 3) / (P, M, 3) clouds: S (P,), ov (P, 4) = [VAB, VAA, VBB, 0]
    and what the backward needs."""
    _check_shapes(ref, fit, ref_w, fit_w)
    if not (ref.is_cuda and fit.is_cuda):
        raise RuntimeError("gaussian_similarity runs on the GPU only (HIP kernels)")
    if ref.dtype != fit.dtype or ref.dtype not in (torch.float32, torch.float64):
        raise TypeError("ref and fit must both be float32 or both float64")
    P, N, M = ref.shape[0], ref.shape[1], fit.shape[1]
    esp = ref_w is not None
    mode = _lib.MODE_SHAPE_ESP if esp else _lib.MODE_SHAPE
    rp, fp = _pack(ref.detach(), ref_w), _pack(fit.detach(), fit_w)
    rr, fr = _ranges(P, N, ref.device), _ranges(P, M, ref.device)
    pb = _lib.FssPairBatch()
    pb.n_pairs, pb.ref_rng, pb.fit_rng = P, _lib.ptr(rr), _lib.ptr(fr)
    pb.max_ref, pb.max_fit = N, M
    sp = _lib.FssScoreParams()
    sp.mode, sp.alpha, sp.lam = mode, float(alpha), float(lam if esp else 1.0)
    S = torch.empty(P, dtype=ref.dtype, device=ref.device)
    ov = torch.empty((P, 4), dtype=ref.dtype, device=ref.device)
    stream = int(torch.cuda.current_stream().cuda_stream)
    if ref.dtype == torch.float32:
        pb.ref_pts, pb.fit_pts = _lib.ptr(rp), _lib.ptr(fp)
        _lib.call("fss_score_fwd", _lib.byref(pb), _lib.byref(sp), _lib.ptr(S), _lib.ptr(ov), stream)
    else:
        _lib.call("fss_score_fwd_f64", _lib.byref(pb), _lib.byref(sp), _lib.ptr(rp), _lib.ptr(fp),
                  _lib.ptr(S), _lib.ptr(ov), stream)
    return S, ov, (rp, fp, rr, fr), (P, N, M, mode, float(alpha), float(lam if esp else 1.0))


def overlaps(ref: torch.Tensor, fit: torch.Tensor, ref_esp: Optional[torch.Tensor] = None,
             fit_esp: Optional[torch.Tensor] = None, alpha: float = 0.81,
             lam: float = 0.3 * LAM_SCALING) -> torch.Tensor:
    """(P, 3) [VAB, VAA, VBB] of each cloud pair, no gradient (the forward kernel's by-product)."""
    if (ref_esp is None) != (fit_esp is None):
        raise ValueError("pass ESP values for both clouds or for neither")
    with torch.no_grad():
        return _forward(ref, fit, ref_esp, fit_esp, alpha, lam)[1][:, :3]


class _GaussianSimilarity(torch.autograd.Function):
    @staticmethod
    def forward(ctx, ref, fit, ref_w, fit_w, alpha, lam):
        S, ov, (rp, fp, rr, fr), meta = _forward(ref, fit, ref_w, fit_w, alpha, lam)
        ctx.save_for_backward(rp, fp, rr, fr, ov)
        ctx.meta = meta
        return S

    @staticmethod
    def backward(ctx, gS):
        rp, fp, rr, fr, ov = ctx.saved_tensors
        P, N, M, mode, alpha, lam = ctx.meta
        pb = _lib.FssPairBatch()
        pb.n_pairs, pb.ref_rng, pb.fit_rng = P, _lib.ptr(rr), _lib.ptr(fr)
        pb.max_ref, pb.max_fit = N, M
        sp = _lib.FssScoreParams()
        sp.mode, sp.alpha, sp.lam = mode, alpha, lam
        need_ref = ctx.needs_input_grad[0]
        d_fit = torch.empty((P * M, 4), dtype=fp.dtype, device=fp.device)
        d_ref = torch.empty((P * N, 4), dtype=fp.dtype, device=fp.device) if need_ref else None
        g = gS.contiguous().to(fp.dtype)
        stream = int(torch.cuda.current_stream().cuda_stream)
        if fp.dtype == torch.float32:
            pb.ref_pts, pb.fit_pts = _lib.ptr(rp), _lib.ptr(fp)
            _lib.call("fss_score_bwd", _lib.byref(pb), _lib.byref(sp), _lib.ptr(ov), _lib.ptr(g),
                      _lib.ptr(d_fit), _lib.ptr(d_ref), stream)
        else:
            _lib.call("fss_score_bwd_f64", _lib.byref(pb), _lib.byref(sp), _lib.ptr(rp), _lib.ptr(fp),
                      _lib.ptr(ov), _lib.ptr(g), _lib.ptr(d_fit), _lib.ptr(d_ref), stream)
        gf = d_fit.view(P, M, 4)[..., :3] if ctx.needs_input_grad[1] else None
        gr = d_ref.view(P, N, 4)[..., :3] if need_ref else None
        return gr, gf, None, None, None, None


def gaussian_similarity(ref: torch.Tensor, fit: torch.Tensor, ref_esp: Optional[torch.Tensor] = None,
                        fit_esp: Optional[torch.Tensor] = None, alpha: float = 0.81,
                        lam: float = 0.3 * LAM_SCALING) -> torch.Tensor:
    """Tanimoto similarity of each (ref, fit) cloud pair, differentiable w.r.t. both clouds.

    ref (P, N, 3), fit (P, M, 3) on the GPU; optional ESP values (P, N), (P, M) switch on the
    ESP-weighted overlap with the (already scaled) ``lam``.  Returns S (P,).
    """
    if (ref_esp is None) != (fit_esp is None):
        raise ValueError("pass ESP values for both clouds or for neither")
    _check_shapes(ref, fit, ref_esp, fit_esp)
    return _GaussianSimilarity.apply(ref, fit, ref_esp, fit_esp, alpha, lam)


def similarity_loss(ref, fit, ref_esp=None, fit_esp=None, alpha=0.81, lam=0.3 * LAM_SCALING):
    """Mean of 1 - S over the batch."""
    return (1.0 - gaussian_similarity(ref, fit, ref_esp, fit_esp, alpha, lam)).mean()
