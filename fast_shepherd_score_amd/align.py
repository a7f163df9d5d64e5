"""Batched multi-restart SE(3) alignment on libfss_hip.so.

Three execution paths over the same kernels and the same optimiser semantics:

* :func:`align` -- the persistent path: one ``fss_align`` launch runs every optimiser step of
  every (pair, restart-block) workgroup, then FP64 re-scoring and best-restart selection.
* :func:`align_stepped` -- all pairs advance one step per launch pair (pose loss + gradient,
  Adam, on-device early-stop bookkeeping); the step sequence can be captured once in a HIP
  graph and replayed, so there is no host synchronisation inside the loop.
* :func:`align_plain` -- one pair at a time, one launch per step, with a host read of the mean
  loss every step (the control flow of the reference's per-pair optimisers on HIP kernels).

Optimiser semantics (all paths) restate alignment/_torch_analytical.py:23-478:
loss L_s = 1 - S_s per restart, gradient [dL/dq_raw, dL/dt] / B handed to torch-style Adam
(bias corrected, eps outside the sqrt) on raw quaternions, per-pair early stop once
|mean L_k - mean L_{k-1}| <= 1e-5 has held for more than 10 consecutive steps, then the
similarity of every final pose and the argmax over restarts.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Optional, Tuple

import numpy as np
import torch

from . import _lib
from .constants import LAM_SCALING, SIGMA_MAP, pharm_tables
from .packing import PairBatch

MODES = {"shape": _lib.MODE_SHAPE, "esp": _lib.MODE_SHAPE_ESP, "pharm": _lib.MODE_PHARM,
         "shape_pharm": _lib.MODE_SHAPE_PHARM}


@dataclass
class ScoreSpec:
    """What is scored: objective mode and its constants."""
    mode: int = _lib.MODE_SHAPE
    alpha: float = 0.81
    lam: float = 0.3 * LAM_SCALING          # already scaled (ESP_COMBO: LAM_SCALING * lam)
    color_weight: float = 0.5
    similarity: str = "tanimoto"
    directionless: bool = False
    extended_points: bool = False           # pharmacophore extended points (categories 4, 5)
    only_extended: bool = False
    esp_weight: float = 0.5                 # ESP_COMBO (vol_and_surf_esp) constants
    probe_radius: float = 1.0
    coulomb: float = _lib.COULOMB_SCALING

    def struct(self) -> _lib.FssScoreParams:
        s = _lib.FssScoreParams()
        s.mode = int(self.mode)
        s.alpha = float(self.alpha)
        s.lam = float(self.lam)
        s.color_weight = float(self.color_weight)
        s.esp_weight = float(self.esp_weight)
        s.probe_radius = float(self.probe_radius)
        s.coulomb = float(self.coulomb)
        sim = self.similarity.lower()
        if sim == "tanimoto":
            s.similarity, s.sigma = _lib.SIM_TANIMOTO, 0.5
        elif sim in SIGMA_MAP:
            s.similarity, s.sigma = _lib.SIM_TVERSKY, SIGMA_MAP[sim]
        else:
            raise ValueError("similarity must be one of (tanimoto, tversky, tversky_ref, tversky_fit)")
        a, k, c = pharm_tables(self.directionless, self.extended_points, self.only_extended)
        for i in range(_lib.N_PTYPES):
            s.ph_alpha[i], s.ph_k[i], s.ph_cat[i] = a[i], k[i], c[i]
        return s


@dataclass
class OptimSpec:
    """How the pose is optimised."""
    max_steps: int = 200
    lr: float = 0.1
    beta1: float = 0.9
    beta2: float = 0.999
    eps: float = 1e-8
    early_stop: bool = True
    es_tol: float = 1e-5
    es_patience: int = 10


@dataclass
class AlignResult:
    score: torch.Tensor         # (P,) float64 best similarity per pair
    best_idx: torch.Tensor      # (P,) int32
    transform: torch.Tensor     # (P, 4, 4) float32
    q: torch.Tensor             # (P, B, 4) float32 final raw quaternions
    t: torch.Tensor             # (P, B, 3) float32 final translations
    pose_score: torch.Tensor    # (P, B) float64
    n_steps: torch.Tensor       # (P,) int32
    extra: dict = field(default_factory=dict)


def es_start(mode: int) -> float:
    """The early-stop reference loss before the first step: 1.0 for the analytic optimisers
    (alignment/_torch_analytical.py:110, 283, 420), +inf for the ShaEP optimiser
    (alignment/_torch.py:922), so its first step never counts towards the patience."""
    return float("inf") if mode == _lib.MODE_ESP_COMBO else 1.0


def _stream(stream=None) -> int:
    s = stream if stream is not None else torch.cuda.current_stream()
    return int(s.cuda_stream)


def self_overlaps(pb: PairBatch, spec: ScoreSpec, stream=None) -> torch.Tensor:
    """(P, 4) float64 [VAA_shape, VBB_shape, VAA_pharm, VBB_pharm] at the identity pose."""
    out = torch.empty((pb.n_pairs, 4), dtype=torch.float64, device=pb.device)
    bs, sp = pb.struct(), spec.struct()
    _lib.call("fss_self_overlap", _lib.byref(bs), _lib.byref(sp), _lib.ptr(out), _stream(stream))
    return out


def _next_pow2(x: int) -> int:
    return 1 << max(0, int(np.ceil(np.log2(max(1, x)))))


def tiling_efficiency(counts: np.ndarray, B: int, lpp: int, rpw: int) -> float:
    """Fraction of issued lane-work that is useful for a (lanes_per_pose, restarts_per_wg) tiling:
    active lanes of the wavefronts that run x fit rows of the ceil(M / lpp) row slots each lane
    group runs (averaged over the pairs' fit counts)."""
    m = np.maximum(np.asarray(counts, np.float64), 1.0)
    slot_util = float(np.mean(m / (lpp * np.ceil(m / lpp))))
    return float(B * lpp / (64.0 * _waves_per_pair(B, lpp, rpw)) * slot_util)


def _waves_per_pair(B: int, lpp: int, rpw: int) -> int:
    """Wavefronts that run for one pair: a wave whose lane groups are all past restart B exits."""
    full, last = divmod(B, rpw)
    per_wg = lambda r: int(np.ceil(r * lpp / 64.0))
    return full * per_wg(rpw) + (per_wg(last) if last else 0)


def candidate_tilings(B: int, early_stop: bool):
    # one wavefront holding every restart with the widest lane group that fits: 64 // B lanes,
    # a non-power-of-two width when B does not divide 64 (10 restarts -> 6 lanes each)
    if 2 <= B <= 32 and (64 // B) & (64 // B - 1):
        yield 64 // B, B
    divisors = [d for d in range(1, B + 1) if B % d == 0]
    for lpp in (1, 2, 4, 8, 16, 32, 64):
        one_wave = max(1, 64 // lpp)
        rpws = [B] if early_stop else sorted({r for r in divisors + [(B + 1) // 2, (B + 2) // 3,
                                                                    (B + 3) // 4, one_wave,
                                                                    2 * one_wave, 4 * one_wave,
                                                                    8, 16, 32]
                                              if 1 <= r <= B})
        for rpw in rpws:
            if rpw * lpp <= 1024 and rpw * lpp >= 64 or (rpw * lpp < 64 and rpw == B):
                yield lpp, min(rpw, B)


N_SIMD = 256 * 4          # MI355X: 256 CUs x 4 SIMDs


def tiling_cost(ref_counts, fit_counts, B: int, lpp: int, rpw: int, waves_per_simd: int = 4) -> float:
    """Modelled optimiser-step time of a tiling (arbitrary units).

    A wave's serial work U is its lanes' ceil(M / lpp) row slots over the ref cloud (about 10
    issue slots per pair term) plus the lane-group reduction and the per-pose tail.  A SIMD
    running w of the launch's waves takes max(LATENCY, w) x U: alone, a wave is latency-bound
    (about 2.2 times its issue time, measured on small MI355X launches); from about two waves per
    SIMD on, the SIMD is issue-bound.  The waves spread over the CUs the launch's workgroups
    reach (at most 256); a CU holds as many workgroups as fit in ``4 x waves_per_simd`` wave
    slots, and a workgroup whose wave count does not divide evenly over the CU's 4 SIMDs leaves
    SIMDs idle (the measured MI355X behaviour that favours workgroups of 1, 2, 4 or 8 waves)."""
    n = np.maximum(np.asarray(ref_counts, np.float64), 1.0)
    m = np.maximum(np.asarray(fit_counts, np.float64), 1.0)
    # lane-group reduction: DPP levels for power-of-two widths, a shuffle tree plus a broadcast
    # (about 1.5 times the cost per level) otherwise
    levels = np.log2(max(lpp, 1)) if lpp & (lpp - 1) == 0 else 1.5 * (np.ceil(np.log2(lpp)) + 1)
    per_wave = np.ceil(m / lpp) * n * 10.0 + 26.0 * levels + 150.0
    wpg = int(np.ceil(rpw * lpp / 64.0))
    slots = 4 * waves_per_simd
    wgs_per_cu = max(1, slots // wpg)
    resident = wgs_per_cu * wpg
    simd_eff = resident / (4.0 * np.ceil(resident / 4.0))
    waves = float(_waves_per_pair(B, lpp, rpw) * len(m))
    busy_cu = min(N_SIMD // 4, int(np.ceil(B / rpw)) * len(m))
    w = max(waves / (4.0 * busy_cu * simd_eff), float(np.ceil(wpg / 4.0)))
    return float(np.mean(per_wave)) * max(LATENCY_WAVES, w)


LATENCY_WAVES = 2.2       # waves per SIMD below which a wave's latency, not issue, sets the step


def default_tiling(pb: PairBatch, mode: int, B: int, early_stop: bool) -> Tuple[int, int]:
    """(lanes_per_pose, restarts_per_wg) minimising :func:`tiling_cost` over the legal tilings."""
    if mode == _lib.MODE_PHARM:
        rc, fc = pb.ref_pcounts, pb.fit_pcounts
    elif mode == _lib.MODE_ESP_COMBO:    # lanes walk the surface observers, inner loop the atoms
        rc, fc = pb.fit_acounts, pb.ref_scounts
    else:
        rc, fc = pb.ref_counts, pb.fit_counts
    rc = rc if rc is not None and len(rc) else np.ones(1)
    fc = fc if fc is not None and len(fc) else np.ones(1)
    cands = [(tiling_cost(rc, fc, B, l, r), -tiling_efficiency(fc, B, l, r), l, r)
             for l, r in candidate_tilings(B, early_stop)]
    cands.sort()
    return cands[0][2], cands[0][3]


def mfma_by_default(pb: PairBatch, mode: int, early_stop: bool) -> bool:
    """Whether the untuned default is the split-bf16 MFMA loop (shape modes): from clouds of about
    100 points per side with early stop, 150 without (measured on surface + ESP batches of 2048
    pairs, 50 x 200: 60-90 points 30 ms VALU / 37 ms MFMA; 100-150 81 / 59 with early stop, 81 /
    83 without; 150-200 161 / 93 and 153 / 138; 200-400 points 1075 / 466 and 956 / 672 ms on
    4096 pairs)."""
    if mode not in (_lib.MODE_SHAPE, _lib.MODE_SHAPE_ESP) or pb.ref_counts is None or not len(pb.ref_counts):
        return False
    size = float(np.sqrt(np.mean(pb.ref_counts) * np.mean(pb.fit_counts)))
    return size >= (100.0 if early_stop else 150.0)


def is_streamed(pb: PairBatch, spec: ScoreSpec, B: int) -> bool:
    """True when the batch's clouds do not fit one workgroup's LDS in the expanded form, so
    fss_align streams them through LDS tiles (shape modes; other modes then fail to plan)."""
    ap = _lib.FssAlignParams()
    ap.n_restarts, ap.max_steps = B, 1
    ap.lanes_per_pose, ap.restarts_per_wg, ap.variant = 1, 1, 2
    info = _lib.FssAlignPlanInfo()
    bs, sp = pb.struct(), spec.struct()
    lib = _lib.load()
    if lib.fss_align_plan(_lib.byref(bs), _lib.byref(sp), _lib.byref(ap), _lib.byref(info)) != 0:
        return False          # the launch reports its own error
    return info.variant == 4


def streamed_tiling(P: int, B: int, early_stop: bool, max_fit: int, waves_wanted: int = 8 * N_SIMD):
    """(lanes_per_pose, restarts_per_wg, groups_per_wg) of the streamed path.  A lane owns fit
    rows, so a pose's lanes span more wavefronts (up to 1024 lanes) only while the launch holds
    fewer than about ``waves_wanted`` wavefronts and every lane keeps at least four fit rows.
    Workgroups of one-wavefront groups hold two restarts (measured fastest on 12000 x 300-row
    surface pairs); with early stop a workgroup holds every restart of the pair, its lane groups
    running the rest in turn."""
    lpp = 64
    while lpp < 1024 and P * B * lpp // 64 < waves_wanted and 8 * lpp <= max_fit:
        lpp *= 2
    if early_stop:
        return lpp, B, max(1, min(B, 1024 // lpp))
    groups = max(1, min(B, 128 // lpp))
    return lpp, groups, groups


def check_tiling(B: int, lpp: int, rpw: int, early_stop: bool, groups: int = 0, pairs: int = 1) -> None:
    g = groups or rpw
    ppw = max(int(pairs or 1), 1)
    if g > rpw:
        raise ValueError("groups_per_wg exceeds restarts_per_wg")
    if ppw > 1 and rpw < B:
        raise ValueError("pairs_per_wg > 1 needs every restart of each pair in the workgroup")
    if lpp > 64 and lpp in (128, 256, 512, 1024) and ppw == 1 and g * lpp <= 1024:
        pass                  # lane groups of several wavefronts (the streamed form checks the rest)
    elif lpp < 1 or lpp > 64 or (lpp & (lpp - 1) and ppw * g * lpp > 64):
        raise ValueError("lanes_per_pose must be in [1, 64] (or 128 .. 1024 on the streamed form), and a power "
                         "of two unless the workgroup's pairs x lane groups x lanes_per_pose fit one wavefront (64)")
    if early_stop and rpw < B:
        raise ValueError("per-pair early stop needs every restart of a pair in one workgroup")
    if ppw * g * lpp > 1024:
        raise ValueError("pairs x lane groups x lanes_per_pose exceed 1024 threads")


def _prep_seeds(pb: PairBatch, q0: torch.Tensor, t0: torch.Tensor):
    P = pb.n_pairs
    if q0.dim() != 3 or q0.shape[0] != P or q0.shape[2] != 4:
        raise ValueError(f"q0 must be (n_pairs={P}, B, 4), got {tuple(q0.shape)}")
    B = int(q0.shape[1])
    if tuple(t0.shape) != (P, B, 3):
        raise ValueError(f"t0 must be ({P}, {B}, 3), got {tuple(t0.shape)}")
    q = q0.to(device=pb.device, dtype=torch.float32).contiguous().clone()
    t = t0.to(device=pb.device, dtype=torch.float32).contiguous().clone()
    return q, t, B


def align(pb: PairBatch, spec: ScoreSpec, q0: torch.Tensor, t0: torch.Tensor,
          opt: OptimSpec = OptimSpec(), *, lanes_per_pose: Optional[int] = None,
          restarts_per_wg: Optional[int] = None, order: str = "cost",
          self_ov: Optional[torch.Tensor] = None, stream=None, events: Optional[list] = None,
          tune: bool = False, variant: int = 0, groups_per_wg: Optional[int] = None,
          pairs_per_wg: Optional[int] = None, tile_rows: int = 0) -> AlignResult:
    """Persistent path: every step of every pair in one launch.

    Tiling: explicit ``lanes_per_pose`` / ``restarts_per_wg``, else an autotuned choice
    (``tune=True``, or one cached for this batch profile), else the cost model's choice.
    ``variant`` picks the shape inner loop: 1 direct form, 2 expanded form, 3 split-bf16 MFMA
    (one wavefront per pose, so 64 lanes per pose), 0 the tuned or default choice (untuned: the
    MFMA loop for shape-mode clouds of about 100 points per side with early stop, 150 without,
    see :func:`mfma_by_default`; the expanded form otherwise).
    ``groups_per_wg``: lane groups per workgroup, each running restarts_per_wg / groups_per_wg
    restarts in turn every step (None: the tuned choice, else one restart per group).
    ``pairs_per_wg``: pairs per workgroup (each with all its restarts): consecutive pairs of the
    launch order share a workgroup, so small clouds share one wavefront's per-step pose work.
    Clouds that do not fit LDS in the requested layout run the expanded form, streamed through
    LDS tiles in the shape modes (variant 4; ``tile_rows`` rows of each side per tile, 0 for
    1024); ``variant=4`` asks for the streamed form at any size.  ``extra`` reports the launch
    (``variant``, ``tile_ref``, ``tile_fit``).
    ``events``: if a list is given, HIP events bracketing the optimiser kernel and the finishing
    kernels are appended to it as (start, optimiser_end, finish_end).
    """
    q, t, B = _prep_seeds(pb, q0, t0)
    P = pb.n_pairs
    dev = pb.device
    if self_ov is None:
        self_ov = self_overlaps(pb, spec, stream)
    streamed = lanes_per_pose is None and restarts_per_wg is None and is_streamed(pb, spec, B)
    if streamed:
        # clouds beyond LDS: lane groups of several wavefronts when the launch is short of them
        max_fit = int(pb.fit_counts.max()) if pb.fit_counts is not None and len(pb.fit_counts) else 1
        lanes_per_pose, restarts_per_wg, groups_per_wg = streamed_tiling(P, B, opt.early_stop, max_fit)
        variant = 4
    if lanes_per_pose is None and restarts_per_wg is None:
        from . import autotune as _at
        tuned = _at.autotune(pb, spec, q0, t0, opt.early_stop, max_steps=opt.max_steps,
                             self_ov=self_ov) if tune else \
            _at.cached(pb, spec.mode, B, opt.early_stop, opt.max_steps)
        if tuned is not None:
            lpp0, rpw0, var0, grp0 = tuned
            variant = variant or var0
            if groups_per_wg is None and var0 == variant:
                groups_per_wg = grp0
        else:
            lpp0, rpw0 = default_tiling(pb, spec.mode, B, opt.early_stop)
            if variant == 0 and not streamed and mfma_by_default(pb, spec.mode, opt.early_stop):
                variant = 3
    else:
        lpp0, rpw0 = default_tiling(pb, spec.mode, B, opt.early_stop)
    if variant == 3 and spec.mode not in (_lib.MODE_PHARM, _lib.MODE_ESP_COMBO) and lanes_per_pose is None:
        lpp0, rpw0 = 64, (rpw0 if lpp0 == 64 else 4)
    lpp = lanes_per_pose or lpp0
    rpw = restarts_per_wg or (B if opt.early_stop else min(rpw0, B))
    grp = int(groups_per_wg or 0)
    if variant == 3 and lanes_per_pose is None and opt.early_stop and B > 16 and not grp:
        # the MFMA loop needs a wavefront per pose: with every restart of a pair in one
        # workgroup, 50 restarts run as 8 wavefronts of 6-7 restarts each (two such workgroups
        # fill a CU's 16 wavefront slots; 425 ms against 626 ms for 5 wavefronts of 10 on the
        # 4096-pair surface + ESP batch)
        grp = 8
    ppw = int(pairs_per_wg or 1)
    check_tiling(B, lpp, rpw, opt.early_stop, grp, ppw)
    order_t = None
    if order == "cost" and P > 1:
        c = pb.cost(spec.mode)
        if c.min() != c.max():
            order_t = torch.from_numpy(np.argsort(-c, kind="stable").astype(np.int32)).to(dev)
    pose_score = torch.empty((P, B), dtype=torch.float64, device=dev)
    best = torch.empty(P, dtype=torch.float64, device=dev)
    bidx = torch.empty(P, dtype=torch.int32, device=dev)
    tf = torch.empty((P, 4, 4), dtype=torch.float32, device=dev)
    nst = torch.empty(P, dtype=torch.int32, device=dev)
    ap = _lib.FssAlignParams()
    ap.n_restarts, ap.max_steps = B, int(opt.max_steps)
    ap.lr, ap.beta1, ap.beta2, ap.eps = opt.lr, opt.beta1, opt.beta2, opt.eps
    ap.early_stop, ap.es_tol, ap.es_patience = int(bool(opt.early_stop)), opt.es_tol, int(opt.es_patience)
    ap.lanes_per_pose, ap.restarts_per_wg = int(lpp), int(rpw)
    ap.pair_order = _lib.ptr(order_t)
    ap.variant = int(variant)
    ap.groups_per_wg = grp
    ap.pairs_per_wg = ppw
    ap.tile_rows = int(tile_rows)
    bs, sp = pb.struct(), spec.struct()
    plan = _lib.FssAlignPlanInfo()
    _lib.call("fss_align_plan", _lib.byref(bs), _lib.byref(sp), _lib.byref(ap), _lib.byref(plan))
    out = _lib.FssAlignOut()
    out.q, out.t = _lib.ptr(q), _lib.ptr(t)
    out.pose_score, out.best_score, out.best_idx = _lib.ptr(pose_score), _lib.ptr(best), _lib.ptr(bidx)
    out.best_tf, out.n_steps = _lib.ptr(tf), _lib.ptr(nst)
    if events is None:
        _lib.call("fss_align", _lib.byref(bs), _lib.byref(sp), _lib.byref(ap), _lib.ptr(self_ov),
                  _lib.byref(out), _stream(stream))
    else:
        ts = stream if stream is not None else torch.cuda.current_stream()
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
        ap.skip_finish = 1
        ev[0].record(ts)
        _lib.call("fss_align", _lib.byref(bs), _lib.byref(sp), _lib.byref(ap), _lib.ptr(self_ov),
                  _lib.byref(out), _stream(ts))
        ev[1].record(ts)
        _lib.call("fss_align_finish", _lib.byref(bs), _lib.byref(sp), B, _lib.ptr(self_ov),
                  _lib.byref(out), _stream(ts))
        ev[2].record(ts)
        events.append(tuple(ev))
    return AlignResult(best, bidx, tf, q, t, pose_score, nst,
                       extra={"lanes_per_pose": lpp, "restarts_per_wg": rpw, "variant": int(plan.variant),
                              "groups_per_wg": grp or rpw, "pairs_per_wg": ppw, "order": order_t,
                              "tile_ref": int(plan.tile_ref), "tile_fit": int(plan.tile_fit)})


class SteppedAligner:
    """All pairs advance one optimiser step per (loss+grad, Adam, early-stop) launch triple.

    The step sequence touches only buffers allocated here, so ``capture()`` records it once in
    a HIP graph (torch.cuda.CUDAGraph) and ``run`` replays it ``max_steps`` times.
    """

    def __init__(self, pb: PairBatch, spec: ScoreSpec, q0: torch.Tensor, t0: torch.Tensor,
                 opt: OptimSpec = OptimSpec(), self_ov: Optional[torch.Tensor] = None):
        self.pb, self.spec, self.opt = pb, spec, opt
        self.q, self.t, self.B = _prep_seeds(pb, q0, t0)
        P, B, dev = pb.n_pairs, self.B, pb.device
        self.P = P
        self.self_ov = self_ov if self_ov is not None else self_overlaps(pb, spec)
        self.pose_pair = torch.arange(P, device=dev, dtype=torch.int32).repeat_interleave(B)
        self.loss = torch.zeros(P * B, dtype=torch.float32, device=dev)
        self.grad = torch.zeros((P * B, 7), dtype=torch.float32, device=dev)
        self.m = torch.zeros((P * B, 7), dtype=torch.float32, device=dev)
        self.v = torch.zeros((P * B, 7), dtype=torch.float32, device=dev)
        self.last = torch.full((P,), es_start(spec.mode), dtype=torch.float64, device=dev)
        self.counter = torch.zeros(P, dtype=torch.int32, device=dev)
        self.active = torch.ones(P, dtype=torch.int32, device=dev)
        self.n_steps = torch.zeros(P, dtype=torch.int32, device=dev)
        self._bs, self._sp = pb.struct(), spec.struct()
        self.graph = None

    def step(self, k: int, stream=None) -> None:
        """Optimiser step k (1-based) for every still-active pair."""
        s = _stream(stream)
        P, B = self.P, self.B
        _lib.call("fss_pose_loss_grad", _lib.byref(self._bs), _lib.byref(self._sp), P * B,
                  _lib.ptr(self.pose_pair), B, _lib.ptr(self.self_ov), _lib.ptr(self.q),
                  _lib.ptr(self.t), _lib.ptr(self.loss), _lib.ptr(self.grad), s)
        act = self.active if self.opt.early_stop else None
        _lib.call("fss_adam_step", P * B, _lib.ptr(self.q), _lib.ptr(self.t), _lib.ptr(self.grad),
                  _lib.ptr(self.m), _lib.ptr(self.v), int(k), self.opt.lr, self.opt.beta1,
                  self.opt.beta2, self.opt.eps, _lib.ptr(act), B, s)
        if self.opt.early_stop:
            _lib.call("fss_early_stop_update", P, B, _lib.ptr(self.loss), _lib.ptr(self.last),
                      _lib.ptr(self.counter), _lib.ptr(self.active), _lib.ptr(self.n_steps),
                      int(k), self.opt.es_tol, int(self.opt.es_patience), s)

    def capture(self) -> None:
        """Record all max_steps steps as one HIP graph (Adam's step count is baked per node)."""
        stream = torch.cuda.Stream()
        stream.wait_stream(torch.cuda.current_stream())
        g = torch.cuda.CUDAGraph()
        with torch.cuda.stream(stream):
            with torch.cuda.graph(g, stream=stream):
                for k in range(1, self.opt.max_steps + 1):
                    self.step(k, stream)
        torch.cuda.current_stream().wait_stream(stream)
        self.graph = g

    def run(self, use_graph: bool = False) -> AlignResult:
        if use_graph:
            if self.graph is None:
                self.capture()
            self.graph.replay()
        else:
            for k in range(1, self.opt.max_steps + 1):
                self.step(k)
        if not self.opt.early_stop:
            self.n_steps.fill_(self.opt.max_steps)
        return finish(self.pb, self.spec, self.q.view(self.P, self.B, 4),
                      self.t.view(self.P, self.B, 3), self.self_ov, self.n_steps)


def finish(pb: PairBatch, spec: ScoreSpec, q: torch.Tensor, t: torch.Tensor,
           self_ov: torch.Tensor, n_steps: torch.Tensor, stream=None) -> AlignResult:
    """Final scores and the best restart of every pair, as the persistent launch takes them
    (fss_align_finish): the FP32 similarity of every final pose, FP64 for the restarts within
    1e-3 of the pair's best (every restart in the ShaEP mode), the FP64 argmax over the
    restarts and the SE(3) matrix of the best.  q (P, B, 4), t (P, B, 3) float32."""
    P, B = q.shape[0], q.shape[1]
    dev = pb.device
    q, t = q.contiguous(), t.contiguous()
    ps = torch.empty((P, B), dtype=torch.float64, device=dev)
    best = torch.empty(P, dtype=torch.float64, device=dev)
    bidx = torch.empty(P, dtype=torch.int32, device=dev)
    tf = torch.empty((P, 4, 4), dtype=torch.float32, device=dev)
    out = _lib.FssAlignOut()
    out.q, out.t = _lib.ptr(q), _lib.ptr(t)
    out.pose_score, out.best_score, out.best_idx = _lib.ptr(ps), _lib.ptr(best), _lib.ptr(bidx)
    out.best_tf, out.n_steps = _lib.ptr(tf), _lib.ptr(n_steps)
    bs, sp = pb.struct(), spec.struct()
    _lib.call("fss_align_finish", _lib.byref(bs), _lib.byref(sp), B, _lib.ptr(self_ov), _lib.byref(out),
              _stream(stream))
    return AlignResult(best, bidx, tf, q, t, ps, n_steps)


def align_stepped(pb, spec, q0, t0, opt: OptimSpec = OptimSpec(), use_graph: bool = True,
                  self_ov=None) -> AlignResult:
    return SteppedAligner(pb, spec, q0, t0, opt, self_ov).run(use_graph=use_graph)


def align_plain(pb: PairBatch, spec: ScoreSpec, q0: torch.Tensor, t0: torch.Tensor,
                opt: OptimSpec = OptimSpec(), self_ov=None) -> AlignResult:
    """One pair per launch sequence, one launch pair per step, host early-stop test per step."""
    q, t, B = _prep_seeds(pb, q0, t0)
    P, dev = pb.n_pairs, pb.device
    if self_ov is None:
        self_ov = self_overlaps(pb, spec)
    bs, sp = pb.struct(), spec.struct()
    s = _stream()
    loss = torch.zeros(B, dtype=torch.float32, device=dev)
    grad = torch.zeros((B, 7), dtype=torch.float32, device=dev)
    m = torch.zeros((B, 7), dtype=torch.float32, device=dev)
    v = torch.zeros((B, 7), dtype=torch.float32, device=dev)
    n_steps = np.zeros(P, dtype=np.int32)
    for p in range(P):
        pose_pair = torch.full((B,), p, dtype=torch.int32, device=dev)
        qp, tp = q[p], t[p]
        m.zero_()
        v.zero_()
        last, counter, done = es_start(spec.mode), 0, opt.max_steps
        for k in range(1, opt.max_steps + 1):
            _lib.call("fss_pose_loss_grad", _lib.byref(bs), _lib.byref(sp), B, _lib.ptr(pose_pair),
                      B, _lib.ptr(self_ov), _lib.ptr(qp), _lib.ptr(tp), _lib.ptr(loss),
                      _lib.ptr(grad), s)
            _lib.call("fss_adam_step", B, _lib.ptr(qp), _lib.ptr(tp), _lib.ptr(grad), _lib.ptr(m),
                      _lib.ptr(v), k, opt.lr, opt.beta1, opt.beta2, opt.eps, 0, B, s)
            if opt.early_stop:
                L = float(loss.double().mean().item())
                counter = 0 if abs(L - last) > opt.es_tol else counter + 1
                last = L
                if counter > opt.es_patience:
                    done = k
                    break
        n_steps[p] = done
    return finish(pb, spec, q, t, self_ov, torch.from_numpy(n_steps).to(dev))


def se3_matrix(q: torch.Tensor, t: torch.Tensor) -> torch.Tensor:
    """(K,4) quaternion (normalised here) + (K,3) translation -> (K,4,4) [[R, t],[0, 1]]."""
    qn = q / q.norm(dim=1, keepdim=True)
    w, x, y, z = qn.unbind(1)
    R = torch.stack([1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w),
                     2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w),
                     2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)],
                    dim=1).view(-1, 3, 3)
    T = torch.zeros((q.shape[0], 4, 4), dtype=q.dtype, device=q.device)
    T[:, :3, :3] = R
    T[:, :3, 3] = t
    T[:, 3, 3] = 1
    return T


def overlap_grad(pb: PairBatch, spec: ScoreSpec, pose_pair: torch.Tensor, q: torch.Tensor,
                 t: torch.Tensor, need_grad: bool = True, stream=None):
    """Per-pose channel overlaps (and gradients): V (n,2), gR (n,2,3,3), gT (n,2,3), gQ (n,2,4)."""
    n = int(q.shape[0])
    dev = pb.device
    V = torch.zeros((n, 2), dtype=torch.float32, device=dev)
    gR = torch.zeros((n, 2, 3, 3), dtype=torch.float32, device=dev) if need_grad else None
    gT = torch.zeros((n, 2, 3), dtype=torch.float32, device=dev) if need_grad else None
    gQ = torch.zeros((n, 2, 4), dtype=torch.float32, device=dev) if need_grad else None
    pp = pose_pair.to(device=dev, dtype=torch.int32).contiguous()
    qq = q.to(device=dev, dtype=torch.float32).contiguous()
    tt = t.to(device=dev, dtype=torch.float32).contiguous()
    bs, sp = pb.struct(), spec.struct()
    _lib.call("fss_overlap_grad", _lib.byref(bs), _lib.byref(sp), n, _lib.ptr(pp), _lib.ptr(qq),
              _lib.ptr(tt), int(need_grad), _lib.ptr(V), _lib.ptr(gR), _lib.ptr(gT), _lib.ptr(gQ),
              _stream(stream))
    return V, gR, gT, gQ


def pose_loss_grad(pb: PairBatch, spec: ScoreSpec, pose_pair, B: int, q, t, self_ov, stream=None):
    """Per-pose loss (n,) and SE(3) gradient (n, 7) of one optimiser step."""
    n = int(q.shape[0])
    dev = pb.device
    loss = torch.zeros(n, dtype=torch.float32, device=dev)
    grad = torch.zeros((n, 7), dtype=torch.float32, device=dev)
    pp = pose_pair.to(device=dev, dtype=torch.int32).contiguous()
    qq = q.to(device=dev, dtype=torch.float32).contiguous()
    tt = t.to(device=dev, dtype=torch.float32).contiguous()
    bs, sp = pb.struct(), spec.struct()
    _lib.call("fss_pose_loss_grad", _lib.byref(bs), _lib.byref(sp), n, _lib.ptr(pp), int(B),
              _lib.ptr(self_ov), _lib.ptr(qq), _lib.ptr(tt), _lib.ptr(loss), _lib.ptr(grad),
              _stream(stream))
    return loss, grad


def pose_score(pb: PairBatch, spec: ScoreSpec, pose_pair, q, t, self_ov, stream=None):
    """FP64 similarity of arbitrary poses (n,)."""
    n = int(q.shape[0])
    out = torch.empty(n, dtype=torch.float64, device=pb.device)
    pp = pose_pair.to(device=pb.device, dtype=torch.int32).contiguous()
    qq = q.to(device=pb.device, dtype=torch.float32).contiguous()
    tt = t.to(device=pb.device, dtype=torch.float32).contiguous()
    bs, sp = pb.struct(), spec.struct()
    _lib.call("fss_pose_score", _lib.byref(bs), _lib.byref(sp), n, _lib.ptr(pp), 1,
              _lib.ptr(self_ov), _lib.ptr(qq), _lib.ptr(tt), _lib.ptr(out), _stream(stream))
    return out
