"""fss_align_noalloc (no device allocation; self-overlaps and re-scoring buffer in the library's
static per-GPU scratch) against fss_align with caller-allocated buffers (needs an MI355X).

Three calls in a row through the C ABI, each with a smaller batch (and another mode) than the
one before, so each runs on scratch that still holds the earlier calls' contents: each must
equal fss_align bit for bit (the same kernels on the same inputs).  Oversized batches and
skip_finish are rejected before any launch.
"""
import numpy as np
import pytest
import torch

from fast_shepherd_score_amd import _lib
from fast_shepherd_score_amd.align import OptimSpec, align
from tests._cases import make_case, random_seeds, to_dev

pytestmark = pytest.mark.gpu


def _params(B, opt, lpp, rpw, variant, skip_finish=0):
    ap = _lib.FssAlignParams()
    ap.n_restarts, ap.max_steps = B, opt.max_steps
    ap.lr, ap.beta1, ap.beta2, ap.eps = opt.lr, opt.beta1, opt.beta2, opt.eps
    ap.early_stop, ap.es_tol, ap.es_patience = int(opt.early_stop), opt.es_tol, opt.es_patience
    ap.lanes_per_pose, ap.restarts_per_wg, ap.variant = lpp, rpw, variant
    ap.pair_order, ap.skip_finish = None, skip_finish
    return ap


def noalloc(pb, spec, q0, t0, opt, lpp, rpw, variant, skip_finish=0):
    """One fss_align_noalloc call: only the seeds and the outputs are passed (caller-owned)."""
    P, B = q0.shape[:2]
    dev = q0.device
    q, t = q0.clone(), t0.clone()
    best = torch.empty(P, dtype=torch.float64, device=dev)
    bidx = torch.empty(P, dtype=torch.int32, device=dev)
    tf = torch.empty((P, 4, 4), dtype=torch.float32, device=dev)
    nst = torch.empty(P, dtype=torch.int32, device=dev)
    out = _lib.FssAlignOut()
    out.q, out.t, out.pose_score = _lib.ptr(q), _lib.ptr(t), None
    out.best_score, out.best_idx, out.best_tf, out.n_steps = (_lib.ptr(best), _lib.ptr(bidx), _lib.ptr(tf),
                                                              _lib.ptr(nst))
    bs, sp = pb.struct(), spec.struct()
    ap = _params(B, opt, lpp, rpw, variant, skip_finish)
    _lib.call("fss_align_noalloc", _lib.byref(bs), _lib.byref(sp), _lib.byref(ap), _lib.byref(out),
              int(torch.cuda.current_stream().cuda_stream))
    return dict(score=best, best_idx=bidx, transform=tf, q=q, t=t, n_steps=nst)


CALLS = [  # (mode, pairs, restarts, steps, early stop, lanes per pose, restarts per workgroup, variant)
    (_lib.MODE_SHAPE_ESP, 12, 10, 60, True, 16, 10, 2),
    (_lib.MODE_SHAPE, 5, 6, 40, False, 64, 2, 3),
    (_lib.MODE_SHAPE_PHARM, 4, 5, 30, True, 1, 5, 2),
]


def test_noalloc_matches_fss_align_calls_in_a_row(dev):
    for k, (mode, n, B, steps, es, lpp, rpw, var) in enumerate(CALLS):
        print(f"case noalloc call {k}: mode {mode}, {n} pairs x {B} restarts", flush=True)
        pb, spec, _ = make_case(mode, n, dev, seed=60 + k)
        q, t = random_seeds(n, B, seed=61 + k, t_scale=0.3)
        q0, t0 = to_dev(q, dev), to_dev(t, dev)
        opt = OptimSpec(max_steps=steps, early_stop=es)
        got = noalloc(pb, spec, q0, t0, opt, lpp, rpw, var)
        want = align(pb, spec, q0, t0, opt, lanes_per_pose=lpp, restarts_per_wg=rpw, variant=var, order="none")
        torch.cuda.synchronize()
        for name in ("score", "best_idx", "transform", "q", "t", "n_steps"):
            assert torch.equal(got[name], getattr(want, name)), (k, name)
        assert np.isfinite(got["score"].cpu().numpy()).all()


def test_noalloc_rejects_oversized_batches_and_skip_finish(dev):
    print("case noalloc rejections", flush=True)
    pb, spec, _ = make_case(_lib.MODE_SHAPE, 2, dev, seed=70)
    q, t = random_seeds(2, 4, seed=71)
    q0, t0 = to_dev(q, dev), to_dev(t, dev)
    with pytest.raises(_lib.FssError, match="skip_finish"):
        noalloc(pb, spec, q0, t0, OptimSpec(max_steps=5), 16, 4, 2, skip_finish=1)
    # 2 pairs x 600000 restarts need 9.6 MB of pose scores: refused before anything is launched
    B = 600_000
    with pytest.raises(_lib.FssError, match="scratch"):
        ap_q = torch.zeros((2, B, 4), dtype=torch.float32, device=dev)
        ap_t = torch.zeros((2, B, 3), dtype=torch.float32, device=dev)
        noalloc(pb, spec, ap_q, ap_t, OptimSpec(max_steps=1, early_stop=False), 64, 16, 2)
