"""HIP kernels vs the FP64 oracle on identical seeded inputs (needs an MI355X).

Tolerances: per-pose scores within 1e-4 absolute; best-restart scores
within 1e-3 under identical restart seeds.  Gradients of the FP32 kernels: 1e-4 relative to the
largest component.
"""
import numpy as np
import pytest
import torch

from fast_shepherd_score_amd import _lib
from fast_shepherd_score_amd.align import (OptimSpec, align, align_plain, align_stepped, overlap_grad,
                                           pose_loss_grad, pose_score, self_overlaps)
from oracle import fss_oracle as orc
from tests._cases import make_case, random_seeds, to_dev

pytestmark = pytest.mark.gpu
MODES = [_lib.MODE_SHAPE, _lib.MODE_SHAPE_ESP, _lib.MODE_PHARM, _lib.MODE_SHAPE_PHARM]
DT = torch.float64


def _se3(q, t):
    return torch.cat([torch.as_tensor(q, dtype=DT), torch.as_tensor(t, dtype=DT)], dim=-1)


@pytest.mark.parametrize("mode", MODES)
def test_self_overlaps_fp64(dev, mode):
    pb, spec, pairs = make_case(mode, 5, dev, seed=1)
    so = self_overlaps(pb, spec).cpu().numpy()
    for i, p in enumerate(pairs):
        o = p.self_overlaps()
        if p.has_shape:
            np.testing.assert_allclose(so[i, :2], [float(o["VAA_s"]), float(o["VBB_s"])], rtol=1e-12)
        if p.has_pharm:
            np.testing.assert_allclose(so[i, 2:], [float(o["VAA_p"]), float(o["VBB_p"])], rtol=1e-6)


@pytest.mark.parametrize("mode", MODES)
def test_overlap_and_gradients_per_pose(dev, mode):
    print(f"case overlap_grad mode={mode}", flush=True)
    pb, spec, pairs = make_case(mode, 4, dev, seed=2)
    B = 6
    q, t = random_seeds(4, B, seed=3, t_scale=0.5)
    pp = torch.arange(4, dtype=torch.int32).repeat_interleave(B)
    V, gR, gT, gQ = overlap_grad(pb, spec, pp, to_dev(q.reshape(-1, 4), dev), to_dev(t.reshape(-1, 3), dev))
    V, gR, gT = V.cpu().numpy(), gR.cpu().numpy(), gT.cpu().numpy()
    for i, p in enumerate(pairs):
        qn = torch.tensor(q[i], dtype=DT)
        qn = qn / qn.norm(dim=1, keepdim=True)
        R = orc.rot_from_unit_quat(qn)
        ch = p.channel_overlaps(R, torch.tensor(t[i], dtype=DT))
        for c, key in enumerate(("shape", "pharm")):
            if key not in ch:
                continue
            O, GR, GT = (x.numpy() for x in ch[key])
            sl = slice(i * B, (i + 1) * B)
            np.testing.assert_allclose(V[sl, c], O, rtol=2e-5, atol=1e-6)
            scale = max(np.abs(GR).max(), np.abs(GT).max(), 1e-6)
            np.testing.assert_allclose(gR[sl, c], GR, atol=1e-4 * scale)
            np.testing.assert_allclose(gT[sl, c], GT, atol=1e-4 * scale)


@pytest.mark.parametrize("mode", MODES)
def test_pose_loss_grad_and_score(dev, mode):
    print(f"case pose_loss_grad mode={mode}", flush=True)
    pb, spec, pairs = make_case(mode, 3, dev, seed=4)
    B = 5
    q, t = random_seeds(3, B, seed=5, t_scale=0.5)
    q = q * 1.3  # raw (non-unit) quaternions, as the optimiser carries them
    so = self_overlaps(pb, spec)
    pp = torch.arange(3, dtype=torch.int32).repeat_interleave(B)
    qd, td = to_dev(q.reshape(-1, 4), dev), to_dev(t.reshape(-1, 3), dev)
    loss, grad = pose_loss_grad(pb, spec, pp, B, qd, td, so)
    S = pose_score(pb, spec, pp, qd, td, so).cpu().numpy()
    loss, grad = loss.cpu().numpy(), grad.cpu().numpy()
    for i, p in enumerate(pairs):
        se3 = _se3(q[i], t[i])
        _, g, per = p.loss_grad_se3(se3, p.self_overlaps())
        sl = slice(i * B, (i + 1) * B)
        np.testing.assert_allclose(loss[sl], per.numpy(), atol=1e-4)
        scale = max(np.abs(g.numpy()).max(), 1e-8)
        np.testing.assert_allclose(grad[sl], g.numpy(), atol=2e-4 * scale)
        np.testing.assert_allclose(S[sl], p.score_poses(se3).numpy(), atol=1e-9)


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("early_stop", [True, False])
def test_persistent_align_matches_oracle(dev, mode, early_stop):
    print(f"case align mode={mode} es={early_stop}", flush=True)
    n, B, steps = 6, 10, 120
    pb, spec, pairs = make_case(mode, n, dev, seed=6 + mode)
    q, t = random_seeds(n, B, seed=7)
    opt = OptimSpec(max_steps=steps, lr=0.1, early_stop=early_stop)
    res = align(pb, spec, to_dev(q, dev), to_dev(t, dev), opt)
    torch.cuda.synchronize()
    score = res.score.cpu().numpy()
    pose = res.pose_score.cpu().numpy()
    nst = res.n_steps.cpu().numpy()
    for i, p in enumerate(pairs):
        o = orc.optimize(p, q[i], t[i], lr=0.1, max_steps=steps, early_stop=early_stop)
        assert abs(score[i] - o["score"]) < 1e-3, (i, score[i], o["score"])
        # the HIP pose scores are FP64 re-scorings of the HIP poses: consistent with its argmax
        assert score[i] == pose[i].max()
        if not early_stop:
            assert nst[i] == steps


@pytest.mark.parametrize("mode", [_lib.MODE_SHAPE, _lib.MODE_SHAPE_ESP, _lib.MODE_PHARM, _lib.MODE_SHAPE_PHARM])
def test_fp32_prepass_keeps_the_fp64_argmax(dev, mode):
    """Every mode but ShaEP re-scores in FP64 only the restarts within 1e-3 of the pair's FP32
    best: the best restart and its score equal a full FP64 re-scoring of every restart
    (fss_pose_score), the candidates' scores are bitwise the FP64 ones and the rest are within
    1e-5 of FP64."""
    print(f"case fp32 prepass mode={mode}", flush=True)
    n, B, steps = 8, 16, 60
    pb, spec, pairs = make_case(mode, n, dev, seed=21 + mode)
    q, t = random_seeds(n, B, seed=22)
    res = align(pb, spec, to_dev(q, dev), to_dev(t, dev), OptimSpec(max_steps=steps, early_stop=False))
    so = self_overlaps(pb, spec)
    pp = torch.arange(n, dtype=torch.int32, device=dev).repeat_interleave(B)
    full = pose_score(pb, spec, pp, res.q.reshape(-1, 4).contiguous(), res.t.reshape(-1, 3).contiguous(),
                      so).cpu().numpy().reshape(n, B)
    got = res.pose_score.cpu().numpy().reshape(n, B)
    np.testing.assert_allclose(got, full, atol=1e-5)
    assert (got.argmax(1) == full.argmax(1)).all()
    np.testing.assert_array_equal(res.score.cpu().numpy(), full.max(1))
    cand = full >= full.max(1, keepdims=True) - 1e-3 + 2e-5      # safely inside the margin
    np.testing.assert_array_equal(got[cand], full[cand])
    assert (~cand).any()                                          # the pre-pass skipped some


def test_workgroup_size_does_not_change_results(dev):
    """Surface clouds (> 64 points): restarts_per_wg = B and 4 B (one wavefront vs four per
    workgroup; the extra restart groups idle) give bitwise the same alignment, because the
    staging centroid is summed the same way for any workgroup size."""
    n, B = 5, 8
    pb, spec, pairs = make_case(_lib.MODE_SHAPE_ESP, n, dev, seed=41, points=(70, 130))
    assert pb.ref_counts.min() > 64
    q, t = random_seeds(n, B, seed=42)
    opt = OptimSpec(max_steps=40, lr=0.1, early_stop=True)
    a, b = (align(pb, spec, to_dev(q, dev), to_dev(t, dev), opt, lanes_per_pose=8, restarts_per_wg=r)
            for r in (B, 4 * B))
    for name in ("score", "q", "t", "pose_score", "n_steps"):
        assert torch.equal(getattr(a, name), getattr(b, name)), name


def test_paths_agree(dev):
    """Persistent, graph-captured stepped and plain per-pair paths: same optimiser."""
    n, B = 4, 8
    pb, spec, pairs = make_case(_lib.MODE_SHAPE, n, dev, seed=9)
    q, t = random_seeds(n, B, seed=10)
    opt = OptimSpec(max_steps=80, early_stop=True)
    a = align(pb, spec, to_dev(q, dev), to_dev(t, dev), opt)
    b = align_stepped(pb, spec, to_dev(q, dev), to_dev(t, dev), opt, use_graph=True)
    c = align_plain(pb, spec, to_dev(q, dev), to_dev(t, dev), opt)
    torch.cuda.synchronize()
    for r in (b, c):
        np.testing.assert_allclose(r.score.cpu().numpy(), a.score.cpu().numpy(), atol=1e-3)


def test_alone_equals_in_mix_bitwise(dev):
    """A pair's result does not depend on the other pairs of the launch (fixed tiling)."""
    n, B = 8, 10
    pb, spec, _ = make_case(_lib.MODE_SHAPE, n, dev, seed=11)
    q, t = random_seeds(n, B, seed=12)
    opt = OptimSpec(max_steps=60, early_stop=True)
    tile = dict(lanes_per_pose=8, restarts_per_wg=B)
    full = align(pb, spec, to_dev(q, dev), to_dev(t, dev), opt, **tile)
    from fast_shepherd_score_amd.packing import PairBatch
    for i in (0, 5):
        one = PairBatch(pb.ref, pb.fit, fit_idx=[i], ref_idx=[i])
        r = align(one, spec, to_dev(q[i:i + 1], dev), to_dev(t[i:i + 1], dev), opt, **tile)
        assert float(r.score[0]) == float(full.score[i])
        assert torch.equal(r.q[0], full.q[i])


def test_deterministic_repeat(dev):
    pb, spec, _ = make_case(_lib.MODE_SHAPE_ESP, 4, dev, seed=13)
    q, t = random_seeds(4, 8, seed=14)
    opt = OptimSpec(max_steps=40)
    a = align(pb, spec, to_dev(q, dev), to_dev(t, dev), opt)
    b = align(pb, spec, to_dev(q, dev), to_dev(t, dev), opt)
    assert torch.equal(a.pose_score, b.pose_score) and torch.equal(a.q, b.q)


@pytest.mark.parametrize("lpp,rpw", [(1, 10), (4, 10), (16, 10), (64, 10), (8, 3), (32, 1), (6, 10), (3, 10), (5, 10)])
def test_tilings_agree(dev, lpp, rpw):
    """Every legal (lanes_per_pose, restarts_per_wg) computes the same optimisation, including
    lane groups of non-power-of-two width inside one wavefront (6, 3 or 5 lanes x 10 restarts)."""
    n, B = 3, 10
    pb, spec, _ = make_case(_lib.MODE_SHAPE, n, dev, seed=15)
    q, t = random_seeds(n, B, seed=16)
    es = rpw >= B
    opt = OptimSpec(max_steps=50, early_stop=es)
    ref = align(pb, spec, to_dev(q, dev), to_dev(t, dev), opt, lanes_per_pose=16, restarts_per_wg=B)
    r = align(pb, spec, to_dev(q, dev), to_dev(t, dev), opt, lanes_per_pose=lpp, restarts_per_wg=rpw)
    np.testing.assert_allclose(r.pose_score.cpu().numpy(), ref.pose_score.cpu().numpy(), atol=1e-4)


@pytest.mark.parametrize("mode", [_lib.MODE_SHAPE, _lib.MODE_SHAPE_ESP, _lib.MODE_SHAPE_PHARM])
@pytest.mark.parametrize("lpp,ppw", [(1, 6), (2, 3), (3, 2), (1, 4), (6, 1)])
@pytest.mark.parametrize("es", [True, False])
def test_pairs_per_workgroup(dev, mode, lpp, ppw, es):
    """Several pairs per workgroup (each with all 10 restarts) compute the same optimisation as
    one pair per workgroup, with per-pair early stop, uneven clouds in one workgroup and a last
    workgroup holding fewer pairs (7 pairs); the best restarts match the FP64 oracle."""
    print(f"case pairs_per_wg mode={mode} lpp={lpp} ppw={ppw} es={es}", flush=True)
    n, B = 7, 10
    pb, spec, pairs = make_case(mode, n, dev, seed=41, atoms=(10, 40), pharms=(2, 12))
    q, t = random_seeds(n, B, seed=42)
    opt = OptimSpec(max_steps=60, early_stop=es)
    ref = align(pb, spec, to_dev(q, dev), to_dev(t, dev), opt, lanes_per_pose=16, restarts_per_wg=B)
    r = align(pb, spec, to_dev(q, dev), to_dev(t, dev), opt, lanes_per_pose=lpp, restarts_per_wg=B,
              pairs_per_wg=ppw)
    assert r.extra["pairs_per_wg"] == ppw
    np.testing.assert_allclose(r.pose_score.cpu().numpy(), ref.pose_score.cpu().numpy(), atol=1e-4)
    if not es:
        assert (r.n_steps.cpu() == 60).all()
    if mode == _lib.MODE_SHAPE:
        for i in (0, 3, 6):
            o = orc.optimize(pairs[i], q[i].astype(np.float64), t[i].astype(np.float64), max_steps=60,
                             early_stop=es)
            assert abs(float(r.score[i]) - o["score"]) < 1e-3


def test_edge_cases(dev):
    from fast_shepherd_score_amd.align import ScoreSpec
    from fast_shepherd_score_amd.packing import CloudSet, PairBatch
    # single-atom fit, single restart, uneven sizes in one launch
    ref = [np.zeros((1, 3), np.float32), np.random.default_rng(0).normal(size=(37, 3)).astype(np.float32)]
    fit = [np.array([[0.5, 0.0, 0.0]], np.float32), np.random.default_rng(1).normal(size=(5, 3)).astype(np.float32)]
    pb = PairBatch(CloudSet.from_arrays(ref, device=dev), CloudSet.from_arrays(fit, device=dev))
    q = np.tile(np.array([1, 0, 0, 0], np.float32), (2, 1, 1))
    t = np.zeros((2, 1, 3), np.float32)
    r = align(pb, ScoreSpec(), to_dev(q, dev), to_dev(t, dev), OptimSpec(max_steps=30))
    r3 = align(pb, ScoreSpec(), to_dev(q, dev), to_dev(t, dev), OptimSpec(max_steps=30), variant=3)
    torch.cuda.synchronize()
    for i in range(2):
        o = orc.optimize(orc.Pair("shape", ref=ref[i], fit=fit[i]), q[i], t[i], max_steps=30)
        assert abs(float(r.score[i]) - o["score"]) < 1e-3
        assert abs(float(r3.score[i]) - o["score"]) < 1e-3
    # the MFMA inner loop needs a whole wavefront per pose
    with pytest.raises(_lib.FssError, match="lanes_per_pose must be 64"):
        align(pb, ScoreSpec(), to_dev(q, dev), to_dev(t, dev), OptimSpec(max_steps=1), variant=3,
              lanes_per_pose=16, restarts_per_wg=1)
    # clouds beyond LDS: the shape modes stream them through LDS tiles; several pairs per
    # workgroup cannot, and are refused with a clear error
    big = [np.zeros((20000, 3), np.float32)]
    pbb = PairBatch(CloudSet.from_arrays(big, device=dev), CloudSet.from_arrays(big, device=dev))
    so = torch.ones((1, 4), dtype=torch.float64, device=dev)
    rb = align(pbb, ScoreSpec(), to_dev(q[:1], dev), to_dev(t[:1], dev), OptimSpec(max_steps=1), self_ov=so)
    assert rb.extra["variant"] == 4 and rb.extra["tile_ref"] == 1024
    with pytest.raises(_lib.FssError, match="LDS"):
        align(pbb, ScoreSpec(), to_dev(q[:1], dev), to_dev(t[:1], dev), OptimSpec(max_steps=1), self_ov=so,
              lanes_per_pose=1, restarts_per_wg=1, pairs_per_wg=2)


def test_graph_replays_match_eager(dev):
    """The captured step sequence, replayed several times with eager work in between, gives the
    eager result on every replay (no memset nodes: all state resets are kernels)."""
    from fast_shepherd_score_amd.align import SteppedAligner
    n, B = 5, 8
    pb, spec, _ = make_case(_lib.MODE_SHAPE_ESP, n, dev, seed=21)
    q, t = random_seeds(n, B, seed=22)
    qd, td = to_dev(q, dev), to_dev(t, dev)
    opt = OptimSpec(max_steps=30, early_stop=True)
    eager = SteppedAligner(pb, spec, qd, td, opt).run(use_graph=False)
    st = SteppedAligner(pb, spec, qd, td, opt)
    st.capture()
    for rep in range(3):
        st.q.copy_(qd)
        st.t.copy_(td)
        st.m.zero_()
        st.v.zero_()
        st.last.fill_(1.0)
        st.counter.zero_()
        st.active.fill_(1)
        st.n_steps.zero_()
        r = st.run(use_graph=True)
        torch.cuda.synchronize()
        assert torch.equal(r.pose_score, eager.pose_score), rep
        assert torch.equal(r.n_steps, eager.n_steps), rep
        # eager work between replays
        align(pb, spec, qd, td, OptimSpec(max_steps=5))
        torch.cuda.synchronize()


@pytest.mark.parametrize("mode", [_lib.MODE_SHAPE, _lib.MODE_SHAPE_ESP, _lib.MODE_SHAPE_PHARM])
@pytest.mark.parametrize("early_stop", [False, True])
def test_inner_loop_variants_agree(dev, mode, early_stop):
    """The direct, expanded-form and split-bf16 MFMA shape inner loops compute the same
    optimisation, each within the best-restart bar of the FP64 oracle."""
    print(f"case variants mode={mode} es={early_stop}", flush=True)
    pb, spec, pairs = make_case(mode, 4, dev, seed=23)
    q, t = random_seeds(4, 8, seed=24)
    opt = OptimSpec(max_steps=60, early_stop=early_stop)
    res = {v: align(pb, spec, to_dev(q, dev), to_dev(t, dev), opt, variant=v) for v in (1, 2, 3)}
    assert res[3].extra["lanes_per_pose"] == 64
    ref = res[2].pose_score.cpu().numpy()
    if not early_stop:
        for v in (1, 3):
            np.testing.assert_allclose(res[v].pose_score.cpu().numpy(), ref, atol=2e-4)
    for i, p in enumerate(pairs):
        o = orc.optimize(p, q[i], t[i], max_steps=60, early_stop=early_stop)
        for v in (1, 2, 3):
            assert abs(float(res[v].score[i]) - o["score"]) < 1e-3, (v, i)


@pytest.mark.parametrize("mode", [_lib.MODE_SHAPE, _lib.MODE_SHAPE_ESP])
def test_mfma_loop_far_rows_match_oracle(dev, mode):
    """Elongated clouds (rows up to ~30 A from the ref centroid, where the MFMA loop's column
    shift is active) through the persistent optimiser on the split-bf16 MFMA inner loop."""
    from fast_shepherd_score_amd.align import ScoreSpec
    from fast_shepherd_score_amd.constants import LAM_SCALING
    from fast_shepherd_score_amd.packing import CloudSet, PairBatch
    print(f"case mfma_far_rows mode={mode}", flush=True)
    rng = np.random.default_rng(42)
    ref, fit, rw, fw = [], [], [], []
    for n, m in ((70, 45), (37, 90), (130, 17)):
        axis = np.linspace(-28.0, 28.0, n)[:, None] * np.array([[1.0, 0.2, 0.1]])
        ref.append((axis + rng.normal(scale=1.0, size=(n, 3))).astype(np.float32))
        axm = np.linspace(-26.0, 26.0, m)[:, None] * np.array([[1.0, 0.25, -0.1]])
        fit.append((axm + rng.normal(scale=1.0, size=(m, 3))).astype(np.float32))
        rw.append(rng.normal(scale=0.1, size=n).astype(np.float32))
        fw.append(rng.normal(scale=0.1, size=m).astype(np.float32))
    esp = mode == _lib.MODE_SHAPE_ESP
    lam = 0.3 * LAM_SCALING
    spec = ScoreSpec(mode=mode, alpha=0.81, lam=lam)
    pb = PairBatch(CloudSet.from_arrays(ref, rw if esp else None, device=dev),
                   CloudSet.from_arrays(fit, fw if esp else None, device=dev))
    B, steps = 6, 50
    q, t = random_seeds(3, B, seed=43, t_scale=3.0)
    opt = OptimSpec(max_steps=steps, early_stop=False)
    r3 = align(pb, spec, to_dev(q, dev), to_dev(t, dev), opt, variant=3, restarts_per_wg=2)
    r2 = align(pb, spec, to_dev(q, dev), to_dev(t, dev), opt, variant=2)
    np.testing.assert_allclose(r3.pose_score.cpu().numpy(), r2.pose_score.cpu().numpy(), atol=2e-4)
    for i in range(3):
        if esp:
            p = orc.Pair("esp", ref=ref[i], fit=fit[i], ref_esp=rw[i], fit_esp=fw[i], lam=lam)
        else:
            p = orc.Pair("shape", ref=ref[i], fit=fit[i])
        o = orc.optimize(p, q[i], t[i], max_steps=steps, early_stop=False)
        assert abs(float(r3.score[i]) - o["score"]) < 1e-3, (i, float(r3.score[i]), o["score"])


def test_device_seeds_match_reference_rule(dev):
    """fss_seed_poses with the reference rule == the host restatement of _initialize_se3_params:
    identity and Fibonacci rows exactly, the four principal-axis rows as a set."""
    from fast_shepherd_score_amd.packing import CloudSet, PairBatch
    from fast_shepherd_score_amd.seeds import initial_seeds, pad, quats_from_fibo, rotmat
    from fast_shepherd_score_amd import synth
    ref, fit = synth.vol_pairs(6, 15, 40, seed=25)
    fit = [f + np.array([0.5, -1.0, 2.0], np.float32) for f in fit]
    pb = PairBatch(CloudSet.from_arrays(ref, device=dev), CloudSet.from_arrays(fit, device=dev))
    B = 12
    table = np.zeros((B, 4), np.float32)
    table[5:] = quats_from_fibo(B - 5)
    table_d = to_dev(table, dev)
    q = torch.empty((6, B, 4), dtype=torch.float32, device=dev)
    t = torch.empty((6, B, 3), dtype=torch.float32, device=dev)
    bs = pb.struct()
    _lib.call("fss_seed_poses", _lib.byref(bs), 2, B, _lib.ptr(table_d), _lib.ptr(q), _lib.ptr(t),
              int(torch.cuda.current_stream().cuda_stream))
    q, t = q.cpu().numpy(), t.cpu().numpy()
    (R, rn), (F, fn) = pad(ref), pad(fit)
    qh, th = initial_seeds(R, rn, F, fn, B)
    np.testing.assert_allclose(q[:, 0], qh[:, 0], atol=1e-7)
    np.testing.assert_allclose(q[:, 5:], qh[:, 5:], atol=1e-6)
    np.testing.assert_allclose(t[:, [0] + list(range(5, B))], th[:, [0] + list(range(5, B))], atol=1e-4)
    for i in range(6):
        Rd = rotmat(q[i, 1:5].astype(np.float64))
        Rh = rotmat(qh[i, 1:5].astype(np.float64))
        for k in range(4):      # every host PCA rotation appears among the device ones
            assert min(np.abs(Rd[m] - Rh[k]).max() for m in range(4)) < 1e-4, (i, k)


@pytest.mark.parametrize("mode", [_lib.MODE_SHAPE, _lib.MODE_SHAPE_ESP])
def test_far_rows_exponent_window(dev, mode):
    """Elongated clouds (rows up to ~30 A from the ref centroid) put fit rows in every regime of
    the expanded inner loop: inside the shifted exponent window, and beyond it (direct form)."""
    from fast_shepherd_score_amd.align import ScoreSpec
    from fast_shepherd_score_amd.constants import LAM_SCALING
    from fast_shepherd_score_amd.packing import CloudSet, PairBatch
    print(f"case far_rows mode={mode}", flush=True)
    rng = np.random.default_rng(40)
    ref, fit, rw, fw = [], [], [], []
    for n in (90, 130):
        axis = np.linspace(-28.0, 28.0, n)[:, None] * np.array([[1.0, 0.2, 0.1]])
        ref.append((axis + rng.normal(scale=1.0, size=(n, 3))).astype(np.float32))
        fit.append((axis[::-1] + rng.normal(scale=1.0, size=(n, 3))).astype(np.float32))
        rw.append(rng.normal(scale=0.1, size=n).astype(np.float32))
        fw.append(rng.normal(scale=0.1, size=n).astype(np.float32))
    esp = mode == _lib.MODE_SHAPE_ESP
    lam = 0.3 * LAM_SCALING
    spec = ScoreSpec(mode=mode, alpha=0.81, lam=lam)
    pb = PairBatch(CloudSet.from_arrays(ref, rw if esp else None, device=dev),
                   CloudSet.from_arrays(fit, fw if esp else None, device=dev))
    B = 4
    q, _ = random_seeds(2, B, seed=41)
    q[:, 0] = [1, 0, 0, 0]
    t = np.zeros((2, B, 3), np.float32)
    t[:, 1] = [6.0, 0, 0]
    t[:, 2] = [0, 9.0, -3.0]
    pp = torch.arange(2, dtype=torch.int32).repeat_interleave(B)
    V, gR, gT, _ = overlap_grad(pb, spec, pp, to_dev(q.reshape(-1, 4), dev), to_dev(t.reshape(-1, 3), dev))
    V, gR, gT = V.cpu().numpy(), gR.cpu().numpy(), gT.cpu().numpy()
    for i in range(2):
        if esp:
            p = orc.Pair("esp", ref=ref[i], fit=fit[i], ref_esp=rw[i], fit_esp=fw[i], lam=lam)
        else:
            p = orc.Pair("shape", ref=ref[i], fit=fit[i])
        qn = torch.tensor(q[i], dtype=DT)
        qn = qn / qn.norm(dim=1, keepdim=True)
        O, GR, GT = (x.numpy() for x in p.channel_overlaps(orc.rot_from_unit_quat(qn),
                                                           torch.tensor(t[i], dtype=DT))["shape"])
        sl = slice(i * B, (i + 1) * B)
        np.testing.assert_allclose(V[sl, 0], O, rtol=2e-5, atol=1e-6)
        scale = max(np.abs(GR).max(), np.abs(GT).max(), 1e-6)
        np.testing.assert_allclose(gR[sl, 0], GR, atol=1e-4 * scale)
        np.testing.assert_allclose(gT[sl, 0], GT, atol=1e-4 * scale)


@pytest.mark.parametrize("mode,lpp,groups,variant", [
    (_lib.MODE_SHAPE_ESP, 16, 5, 2), (_lib.MODE_SHAPE_ESP, 6, 5, 2), (_lib.MODE_SHAPE_ESP, 64, 5, 3),
    (_lib.MODE_SHAPE, 64, 2, 3), (_lib.MODE_SHAPE_PHARM, 8, 3, 2), (_lib.MODE_PHARM, 4, 4, 2)])
def test_groups_running_restarts_in_turn_are_bitwise_equal(dev, mode, lpp, groups, variant):
    """groups_per_wg < restarts_per_wg: each lane group runs several restarts in turn per step
    (parameters and moments through LDS).  Every restart's arithmetic is unchanged, so the
    results equal one restart per group bit for bit, early stop included."""
    print(f"case groups mode={mode} lpp={lpp} groups={groups} variant={variant}", flush=True)
    n, B = 3, 10
    pb, spec, _ = make_case(mode, n, dev, seed=81 + mode)
    q, t = random_seeds(n, B, seed=82)
    opt = OptimSpec(max_steps=70, early_stop=True)
    kw = dict(lanes_per_pose=lpp, restarts_per_wg=B, variant=variant)
    one = align(pb, spec, to_dev(q, dev), to_dev(t, dev), opt, **kw)
    turn = align(pb, spec, to_dev(q, dev), to_dev(t, dev), opt, groups_per_wg=groups, **kw)
    assert turn.extra["groups_per_wg"] == groups
    for name in ("score", "q", "t", "pose_score", "n_steps"):
        assert torch.equal(getattr(one, name), getattr(turn, name)), name
