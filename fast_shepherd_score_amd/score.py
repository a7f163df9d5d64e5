"""Scoring functions with the reference's names and arguments, on the HIP kernels.

Drop-in for the public entry points of ``shepherd_score.score``:
  * ``VAB_2nd_order``, ``shape_tanimoto``, ``get_overlap``           (gaussian_overlap.py:16-106)
  * ``VAB_2nd_order_esp``, ``shape_tanimoto_esp``, ``get_overlap_esp`` (electrostatic_scoring.py:17-125)
  * ``esp_combo_score``                                             (electrostatic_scoring.py:230-320)
  * ``tanimoto_func``, ``tversky_func``, ``get_overlap_pharm``,
    ``get_pharm_combo_score``                                        (pharmacophore_scoring.py:123-974)

Inputs are torch tensors or numpy arrays, single instance ((N, 3) clouds) or batched ((B, N, 3)),
as in the reference; results are torch tensors on the input's device (a 0-d tensor for a single
instance, (B,) for a batch).  Computation runs on the GPU whatever the input device, and there is
no CPU path.

* Shape and ESP similarities are differentiable with respect to both clouds (``fss_score_fwd`` /
  ``fss_score_bwd``), in the input's precision (float32 or float64).
* ``VAB_2nd_order[_esp]`` is the forward kernel's overlap, without a gradient.
* Pharmacophore, combination and ShaEP scores come from the FP64 re-scoring kernel
  (``fss_pose_score`` at the identity pose), without a gradient; the alignment path is the
  differentiable route for them.
* Pharmacophore vectors are normalised, as in the reference's batched scoring
  (pharmacophore_scoring.py:562-563); unit vectors score the same either way.
"""
from __future__ import annotations

from typing import Optional, Sequence

import numpy as np
import torch

from . import _lib
from .align import ScoreSpec, pose_score, self_overlaps
from .autograd import gaussian_similarity, overlaps
from .constants import LAM_SCALING, SIGMA_MAP
from .packing import CloudSet, PairBatch, PharmSet, ShaepSet

__all__ = ["VAB_2nd_order", "shape_tanimoto", "get_overlap", "VAB_2nd_order_esp",
           "shape_tanimoto_esp", "get_overlap_esp", "esp_combo_score", "tanimoto_func",
           "tversky_func", "get_overlap_pharm", "get_pharm_combo_score"]


def _tensor(x, dtype=None) -> torch.Tensor:
    t = x if torch.is_tensor(x) else torch.as_tensor(np.asarray(x))
    if dtype is not None:
        t = t.to(dtype)
    elif not t.is_floating_point():
        t = t.to(torch.float32)
    return t


def _gpu(t: torch.Tensor) -> torch.Tensor:
    return t if t.is_cuda else t.to("cuda")


def _clouds(c1, c2, w1=None, w2=None):
    """Batched GPU copies (B, N, 3) of two clouds (+ (B, N) weights), whether the input was
    batched, and the device to return results on."""
    a, b = _tensor(c1), _tensor(c2)
    if a.dtype != b.dtype:
        b = b.to(a.dtype)
    if a.dim() not in (2, 3) or b.dim() != a.dim():
        raise ValueError(f"centers must both be (N, 3) or both (B, N, 3); got {tuple(a.shape)}, {tuple(b.shape)}")
    single = a.dim() == 2
    out_dev = a.device
    if single:
        a, b = a.unsqueeze(0), b.unsqueeze(0)
    ws = []
    for w in (w1, w2):
        if w is None:
            ws.append(None)
            continue
        w = _tensor(w, a.dtype)
        ws.append(_gpu(w.unsqueeze(0) if single else w))
    return _gpu(a), _gpu(b), ws[0], ws[1], single, out_dev


def _out(x: torch.Tensor, single: bool, dev, dtype=None) -> torch.Tensor:
    x = x.to(dev)
    if dtype is not None:
        x = x.to(dtype)
    return x[0] if single else x


def _no_grad_inputs(*xs):
    if torch.is_grad_enabled() and any(torch.is_tensor(x) and x.requires_grad for x in xs):
        raise NotImplementedError("this score has no gradient on the HIP backend; gradients are "
                                  "provided for get_overlap / get_overlap_esp and by the alignment path")


# ------------------------------------------------------------------------------- shape / ESP
def VAB_2nd_order(centers_1, centers_2, alpha: float) -> torch.Tensor:
    """Overlap K sum exp(-alpha/2 |a - b|^2) of two clouds (gaussian_overlap.py:16-59)."""
    _no_grad_inputs(centers_1, centers_2)
    a, b, _, _, single, dev = _clouds(centers_1, centers_2)
    return _out(overlaps(a, b, alpha=alpha)[:, 0], single, dev)


def shape_tanimoto(centers_1, centers_2, alpha: float) -> torch.Tensor:
    """Tanimoto shape similarity (gaussian_overlap.py:62-70); differentiable."""
    a, b, _, _, single, dev = _clouds(centers_1, centers_2)
    return _out(gaussian_similarity(a, b, alpha=alpha), single, dev)


def get_overlap(centers_1, centers_2, alpha: float = 0.81) -> torch.Tensor:
    """Tanimoto shape similarity of two clouds (gaussian_overlap.py:73-106); differentiable."""
    return shape_tanimoto(centers_1, centers_2, alpha)


def VAB_2nd_order_esp(centers_1, centers_2, charges_1, charges_2, alpha: float, lam: float) -> torch.Tensor:
    """ESP-weighted overlap, w = exp(-(c_1 - c_2)^2 / lam) with ``lam`` as given
    (electrostatic_scoring.py:17-52)."""
    _no_grad_inputs(centers_1, centers_2)
    a, b, wa, wb, single, dev = _clouds(centers_1, centers_2, charges_1, charges_2)
    return _out(overlaps(a, b, wa, wb, alpha=alpha, lam=lam)[:, 0], single, dev)


def shape_tanimoto_esp(centers_1, centers_2, charges_1, charges_2, alpha: float, lam: float) -> torch.Tensor:
    """ESP-weighted Tanimoto (electrostatic_scoring.py:55-66); differentiable w.r.t. the clouds."""
    a, b, wa, wb, single, dev = _clouds(centers_1, centers_2, charges_1, charges_2)
    return _out(gaussian_similarity(a, b, wa, wb, alpha=alpha, lam=lam), single, dev)


def get_overlap_esp(centers_1, centers_2, charges_1, charges_2, alpha: float = 0.81,
                    lam: float = 0.3 * LAM_SCALING) -> torch.Tensor:
    """ESP-weighted Tanimoto with an already scaled ``lam`` (electrostatic_scoring.py:69-125)."""
    return shape_tanimoto_esp(centers_1, centers_2, charges_1, charges_2, alpha, lam)


# ------------------------------------------------------------------------------- similarity
def tanimoto_func(VAB, VAA, VBB):
    """VAB / (VAA + VBB - VAB) (pharmacophore_scoring.py:123-131)."""
    return VAB / (VAA + VBB - VAB)


def tversky_func(VAB, VAA, VBB, sigma: float):
    """min(VAB / (sigma VAA + (1 - sigma) VBB), 1) (pharmacophore_scoring.py:134-145)."""
    return torch.clamp_max(VAB / (sigma * VAA + (1 - sigma) * VBB), max=1.0)


def _sigma(similarity: str) -> Optional[float]:
    sim = similarity.lower()
    if sim == "tanimoto":
        return None
    if sim in SIGMA_MAP:
        return SIGMA_MAP[sim]
    raise ValueError("Argument `similarity` must be one of (tanimoto, tversky, tversky_ref, tversky_fit).")


# ------------------------------------------------------------------------------- FP64 re-scoring
def _rows(x, single: bool) -> list:
    """Per-instance numpy rows of a single or batched array."""
    a = x.detach().cpu().numpy() if torch.is_tensor(x) else np.asarray(x)
    return [a] if single else [a[i] for i in range(a.shape[0])]


def _identity_scores(pb: PairBatch, spec: ScoreSpec, self_ov=None) -> torch.Tensor:
    n = pb.n_pairs
    q = torch.zeros((n, 4), dtype=torch.float32)
    q[:, 0] = 1.0
    so = self_overlaps(pb, spec) if self_ov is None else self_ov
    return pose_score(pb, spec, torch.arange(n, dtype=torch.int32), q, torch.zeros((n, 3)), so)


def _pharm_sets(ptype_1, ptype_2, anchors_1, anchors_2, vectors_1, vectors_2):
    t1 = _tensor(ptype_1, torch.int64)
    single = t1.dim() == 1
    dev = _tensor(anchors_1).device
    sets = []
    for t, a, v in ((ptype_1, anchors_1, vectors_1), (ptype_2, anchors_2, vectors_2)):
        sets.append(PharmSet.from_arrays(_rows(t, single), _rows(a, single), _rows(v, single), device="cuda"))
    return sets[0], sets[1], single, dev


def get_overlap_pharm(ptype_1, ptype_2, anchors_1, anchors_2, vectors_1, vectors_2,
                      similarity: str = "tanimoto", extended_points: bool = False,
                      only_extended: bool = False, precomputed_self_overlaps=None,
                      directionless: bool = False) -> torch.Tensor:
    """Pharmacophore similarity of the poses as given (pharmacophore_scoring.py:422-912), FP64.
    ``precomputed_self_overlaps`` = (VAA, VBB) replaces the self terms unless extended points are
    on, as in the reference (:555, :908); directionless with a precomputed pair is an error (:524-529)."""
    _no_grad_inputs(anchors_1, anchors_2, vectors_1, vectors_2)
    _sigma(similarity)
    if directionless:
        extended_points = only_extended = False
        if precomputed_self_overlaps is not None:
            raise ValueError("Directionless pharmacophore scoring (`directionless=True`) is incompatible "
                             "with `precomputed_self_overlaps`; recompute self-overlaps directionless.")
    r, f, single, dev = _pharm_sets(ptype_1, ptype_2, anchors_1, anchors_2, vectors_1, vectors_2)
    spec = ScoreSpec(mode=_lib.MODE_PHARM, similarity=similarity, directionless=directionless,
                     extended_points=extended_points, only_extended=only_extended)
    pb = PairBatch(None, None, r, f)
    so = self_overlaps(pb, spec)
    if precomputed_self_overlaps is not None and not extended_points:
        vaa, vbb = (torch.as_tensor(np.asarray(x.detach().cpu() if torch.is_tensor(x) else x, np.float64))
                    .reshape(-1).to(so.device) for x in precomputed_self_overlaps)
        so[:, 2], so[:, 3] = vaa, vbb
    return _out(_identity_scores(pb, spec, so), single, dev, torch.float32)


def get_pharm_combo_score(centers_1, centers_2, ptype_1, ptype_2, anchors_1, anchors_2, vectors_1,
                          vectors_2, alpha: float = 0.81, similarity: str = "tanimoto",
                          extended_points: bool = False, only_extended: bool = False,
                          color_weight: float = 0.5, directionless: bool = False) -> torch.Tensor:
    """(1 - color_weight) shape + color_weight pharm, both channels under ``similarity``
    (pharmacophore_scoring.py:915-974)."""
    _no_grad_inputs(centers_1, centers_2)
    sigma = _sigma(similarity)
    pharm = get_overlap_pharm(ptype_1, ptype_2, anchors_1, anchors_2, vectors_1, vectors_2,
                              similarity=similarity, extended_points=extended_points,
                              only_extended=only_extended, directionless=directionless)
    a, b, _, _, single, dev = _clouds(centers_1, centers_2)
    ov = overlaps(a, b, alpha=alpha).to(torch.float64)
    vab, vaa, vbb = ov[:, 0], ov[:, 1], ov[:, 2]
    shape = tanimoto_func(vab, vaa, vbb) if sigma is None else tversky_func(vab, vaa, vbb, sigma)
    shape = _out(shape, single, dev, torch.float32)
    return (1 - color_weight) * shape + color_weight * pharm


def esp_combo_score(centers_w_H_1, centers_w_H_2, centers_1, centers_2, points_1, points_2,
                    partial_charges_1, partial_charges_2, point_charges_1, point_charges_2,
                    radii_1, radii_2, alpha: float, lam: float = 0.001, probe_radius: float = 1.0,
                    esp_weight: float = 0.5) -> torch.Tensor:
    """ShaEP score: esp_weight (esp_1 + esp_2) / (n_1 + n_2) + (1 - esp_weight) Tanimoto of
    ``centers`` at ``alpha``, ``lam`` raw (scaled by LAM_SCALING here, as the reference's
    comparison does) (electrostatic_scoring.py:128-320), FP64."""
    _no_grad_inputs(centers_1, centers_2, points_1, points_2)
    single = _tensor(points_1).dim() == 2
    dev = _tensor(points_1).device
    R = lambda x: _rows(x, single)
    shape_r = CloudSet.from_arrays(R(centers_1), device="cuda")
    shape_f = CloudSet.from_arrays(R(centers_2), device="cuda")
    se_r = ShaepSet.from_arrays(R(points_1), R(point_charges_1), R(centers_w_H_1), R(partial_charges_1),
                                R(radii_1), device="cuda")
    se_f = ShaepSet.from_arrays(R(points_2), R(point_charges_2), R(centers_w_H_2), R(partial_charges_2),
                                R(radii_2), device="cuda")
    spec = ScoreSpec(mode=_lib.MODE_ESP_COMBO, alpha=alpha, lam=LAM_SCALING * lam,
                     esp_weight=esp_weight, probe_radius=probe_radius)
    pb = PairBatch(shape_r, shape_f, None, None, ref_se=se_r, fit_se=se_f)
    return _out(_identity_scores(pb, spec), single, dev, torch.float32)
