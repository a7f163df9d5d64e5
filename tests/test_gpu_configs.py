"""The HIP optimiser at the benchmark configurations' real sizes against the FP64 oracle
(needs an MI355X).

The oracle's alignments of these pairs take minutes on a CPU, so tests/golden/make_config_vectors.py
ran them once and stored inputs and results (config_*.npz).  Each test aligns the same pairs from
the same seeds on the GPU and checks, per pair, that the best-restart score agrees with the
oracle's within 1e-3.  Live checks on the GPU: the FP32 per-pose loss and SE(3) gradient at the
seeds against the oracle's (loss within 1e-4, gradient within 2e-4 of its largest component), and
the FP64 re-scoring of the final HIP poses against the oracle's scoring restatement (1e-4).

  * C2 (the headline): surface shape + ESP pairs of 200-400 points, lam 0.3 * LAM_SCALING, 50
    restarts x 200 steps, on the split-bf16 MFMA loop at the benchmark's tiling and on the
    autotuned tiling; one pair with the drop-in default's per-pair early stop.
  * C1: 20-40 atom pairs, 50 x 200, persistent, graph-replayed stepped and plain paths.
  * C3: the 5- and 60-atom, 2- and 20-pharmacophore extremes of the C3 workload in one launch,
    directional shape + pharmacophore objective, 50 x 200, persistent (autotuned) and the
    graph-captured stepped path.
  * MoleculePairBatch.align_with_vol_color() with its defaults (directionless).
"""
import numpy as np
import pytest
import torch

from fast_shepherd_score_amd import _lib
from fast_shepherd_score_amd.align import (OptimSpec, ScoreSpec, SteppedAligner, align, align_plain,
                                           pose_loss_grad, self_overlaps)
from fast_shepherd_score_amd.packing import CloudSet, PairBatch, PharmSet
from oracle import fss_oracle as orc
from tests._cases import load_config as load, to_dev

pytestmark = pytest.mark.gpu
DT = torch.float64
BEST_TOL = 1e-3
POSE_TOL = 1e-4


def build(d, dev):
    """(PairBatch, ScoreSpec, oracle Pairs) of a fixture."""
    n = len(d["ref"])
    if d["kind"] == "esp":
        pb = PairBatch(CloudSet.from_arrays(d["ref"], d["ref_w"], device=dev),
                       CloudSet.from_arrays(d["fit"], d["fit_w"], device=dev))
        lam = float(d["lam"])
        spec = ScoreSpec(mode=_lib.MODE_SHAPE_ESP, alpha=0.81, lam=lam)
        pairs = [orc.Pair("esp", ref=d["ref"][i], fit=d["fit"][i], ref_esp=d["ref_w"][i],
                          fit_esp=d["fit_w"][i], lam=lam) for i in range(n)]
    elif d["kind"] == "shape":
        pb = PairBatch(CloudSet.from_arrays(d["ref"], device=dev), CloudSet.from_arrays(d["fit"], device=dev))
        spec = ScoreSpec(mode=_lib.MODE_SHAPE)
        pairs = [orc.Pair("shape", ref=d["ref"][i], fit=d["fit"][i]) for i in range(n)]
    else:
        dl = bool(d["directionless"])
        pb = PairBatch(CloudSet.from_arrays(d["ref"], device=dev), CloudSet.from_arrays(d["fit"], device=dev),
                       PharmSet.from_arrays(d["ref_types"], d["ref_anc"], d["ref_vec"], device=dev),
                       PharmSet.from_arrays(d["fit_types"], d["fit_anc"], d["fit_vec"], device=dev))
        spec = ScoreSpec(mode=_lib.MODE_SHAPE_PHARM, color_weight=0.5, directionless=dl)
        pairs = [orc.Pair("shape_pharm", ref=d["ref"][i], fit=d["fit"][i], ref_types=d["ref_types"][i],
                          fit_types=d["fit_types"][i], ref_anc=d["ref_anc"][i], fit_anc=d["fit_anc"][i],
                          ref_vec=d["ref_vec"][i], fit_vec=d["fit_vec"][i], directionless=dl)
                 for i in range(n)]
    return pb, spec, pairs


def opt_of(d):
    return OptimSpec(max_steps=int(d["steps"]), lr=float(d["lr"]), early_stop=bool(d["early_stop"]))


def check_best(label, res, d):
    got = res.score.cpu().numpy()
    err = np.abs(got - d["best"])
    print(f"{label}: best-restart |HIP - FP64| max {err.max():.2e} (per pair {np.array2string(err, precision=1)})",
          flush=True)
    assert (err <= BEST_TOL).all(), (label, got, d["best"])


def check_final_poses(res, pairs):
    """The HIP FP64 re-scoring of its own final poses equals the oracle's scoring restatement."""
    q, t, ps = res.q.cpu().numpy(), res.t.cpu().numpy(), res.pose_score.cpu().numpy()
    for i, p in enumerate(pairs):
        se3 = torch.cat([torch.as_tensor(q[i], dtype=DT), torch.as_tensor(t[i], dtype=DT)], dim=1)
        ref = p.score_poses(se3).numpy()
        np.testing.assert_allclose(ps[i], ref, atol=POSE_TOL)


def check_seed_losses(pb, spec, pairs, d, dev, pick=None):
    """FP32 per-pose loss and SE(3) gradient of the seeds against the oracle's FP64 ones."""
    n, B = d["q0"].shape[:2]
    so = self_overlaps(pb, spec)
    pp = torch.arange(n, dtype=torch.int32).repeat_interleave(B)
    loss, grad = pose_loss_grad(pb, spec, pp, B, to_dev(d["q0"].reshape(-1, 4), dev),
                                to_dev(d["t0"].reshape(-1, 3), dev), so)
    loss, grad = loss.cpu().numpy().reshape(n, B), grad.cpu().numpy().reshape(n, B, 7)
    for i in (range(n) if pick is None else pick):
        se3 = torch.cat([torch.as_tensor(d["q0"][i], dtype=DT), torch.as_tensor(d["t0"][i], dtype=DT)], dim=1)
        _, g, per = pairs[i].loss_grad_se3(se3, pairs[i].self_overlaps())
        np.testing.assert_allclose(loss[i], per.numpy(), atol=POSE_TOL)
        scale = max(float(g.abs().max()), 1e-8)
        np.testing.assert_allclose(grad[i], g.numpy(), atol=2e-4 * scale)


def test_c2_headline_size_matches_fp64(dev):
    print("case C2 fixture: 4 surface+ESP pairs, 50 x 200", flush=True)
    d = load("c2")
    assert d["ref_n"].min() >= 200 and max(d["ref_n"].max(), d["fit_n"].max()) <= 400
    pb, spec, pairs = build(d, dev)
    q0, t0 = to_dev(d["q0"], dev), to_dev(d["t0"], dev)
    opt = opt_of(d)
    check_seed_losses(pb, spec, pairs, d, dev)
    bench_tiling = align(pb, spec, q0, t0, opt, variant=3, lanes_per_pose=64, restarts_per_wg=4)
    assert bench_tiling.extra["variant"] == 3
    check_best("C2 MFMA loop, 64 lanes x 4 restarts per workgroup", bench_tiling, d)
    check_final_poses(bench_tiling, pairs)
    tuned = align(pb, spec, q0, t0, opt, tune=True)
    print(f"autotuned tiling {tuned.extra}", flush=True)
    check_best("C2 autotuned", tuned, d)
    check_best("C2 VALU expanded loop", align(pb, spec, q0, t0, opt, variant=2), d)


def test_c2_drop_in_default_early_stop_matches_fp64(dev):
    print("case C2 fixture with per-pair early stop", flush=True)
    d = load("c2_es")
    assert bool(d["early_stop"])
    pb, spec, pairs = build(d, dev)
    q0, t0 = to_dev(d["q0"], dev), to_dev(d["t0"], dev)
    res = align(pb, spec, q0, t0, opt_of(d))
    print(f"steps HIP {res.n_steps.cpu().numpy()} oracle {d['n_steps']}", flush=True)
    check_best("C2 early stop", res, d)
    check_final_poses(res, pairs)
    # the MFMA loop with every restart of the pair in one workgroup: 50 restarts as wavefronts
    # that each run several in turn
    mf = align(pb, spec, q0, t0, opt_of(d), variant=3)
    assert mf.extra["variant"] == 3 and mf.extra["groups_per_wg"] < 50
    check_best(f"C2 early stop, MFMA loop ({mf.extra['groups_per_wg']} wavefronts x "
               f"{50 // mf.extra['groups_per_wg']} restarts)", mf, d)
    tuned = align(pb, spec, q0, t0, opt_of(d), tune=True)
    print(f"autotuned with early stop: {tuned.extra}", flush=True)
    check_best("C2 early stop, autotuned", tuned, d)


def test_c1_size_all_paths_match_fp64(dev):
    print("case C1 fixture: 8 atom pairs, 50 x 200", flush=True)
    d = load("c1")
    pb, spec, pairs = build(d, dev)
    q0, t0 = to_dev(d["q0"], dev), to_dev(d["t0"], dev)
    opt = opt_of(d)
    check_seed_losses(pb, spec, pairs, d, dev)
    res = align(pb, spec, q0, t0, opt)
    check_best("C1 persistent", res, d)
    check_final_poses(res, pairs)
    check_best("C1 MFMA loop", align(pb, spec, q0, t0, opt, variant=3), d)
    check_best("C1 graph-replayed stepped", SteppedAligner(pb, spec, q0, t0, opt).run(use_graph=True), d)
    check_best("C1 plain per-pair", align_plain(pb, spec, q0, t0, opt), d)


def test_c3_uneven_extremes_in_one_launch_match_fp64(dev):
    print("case C3 fixture: 5-60 atoms / 2-20 pharmacophores in one launch, 50 x 200", flush=True)
    d = load("c3")
    assert d["ref_n"].min() <= 6 and d["fit_n"].max() >= 59
    assert min(d["ref_pn"].min(), d["fit_pn"].min()) == 2 and max(d["ref_pn"].max(), d["fit_pn"].max()) == 20
    pb, spec, pairs = build(d, dev)
    q0, t0 = to_dev(d["q0"], dev), to_dev(d["t0"], dev)
    opt = opt_of(d)
    check_seed_losses(pb, spec, pairs, d, dev)
    res = align(pb, spec, q0, t0, opt, tune=True)
    print(f"autotuned tiling {res.extra['lanes_per_pose']} lanes x {res.extra['restarts_per_wg']} "
          f"restarts, variant {res.extra['variant']}", flush=True)
    check_best("C3 persistent (autotuned)", res, d)
    check_final_poses(res, pairs)
    st = SteppedAligner(pb, spec, q0, t0, opt)
    st.capture()
    check_best("C3 graph-captured stepped", st.run(use_graph=True), d)


def test_vol_color_default_call_matches_fp64(dev):
    """MoleculePairBatch.align_with_vol_color() exactly as a caller writes it (device seeds from
    the reference's rule, 16 restarts, 40 steps, early stop, directionless), and again with the
    fixture's host-computed seeds passed in."""
    from fast_shepherd_score_amd.container import Molecule, MoleculePair, MoleculePairBatch
    print("case vol_color default call", flush=True)
    d = load("vol_color_default")
    assert bool(d["directionless"]) and int(d["restarts"]) == 16 and int(d["steps"]) == 40
    mol = lambda s, i: Molecule(d[s][i], pharm_types=d[f"{s}_types"][i], pharm_ancs=d[f"{s}_anc"][i],
                                pharm_vecs=d[f"{s}_vec"][i])
    pairs = [MoleculePair(mol("ref", i), mol("fit", i), device=dev) for i in range(len(d["ref"]))]
    batch = MoleculePairBatch(pairs)
    scores, aligned = batch.align_with_vol_color()
    assert aligned == [None] * len(pairs)
    err = np.abs(scores - d["best"])
    print(f"vol_color default: max |HIP - FP64| {err.max():.2e}", flush=True)
    assert (err <= BEST_TOL).all(), (scores, d["best"])
    assert [p.sim_aligned_vol_color for p in pairs] == list(scores)
    scores_h, _ = batch.align_with_vol_color(seeds=(d["q0"], d["t0"]))
    assert (np.abs(scores_h - d["best"]) <= BEST_TOL).all(), (scores_h, d["best"])
