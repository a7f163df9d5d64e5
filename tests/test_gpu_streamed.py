"""The streamed shape path (clouds of any size, staged through LDS tiles each step) against the
FP64 oracle (needs an MI355X).

  * Forced on small batches (variant 4, 48-row tiles, so both sides span several tiles with
    ragged last tiles; surface + ESP, atoms, and atoms + directional pharmacophores, whose
    pharmacophore rows stay staged whole): best restarts within 1e-3 of the oracle's from the same seeds, FP64
    re-scoring of the final poses within 1e-4 of the oracle's scoring, and close to the staged
    expanded form on the same batch.  Tilings with lanes that hold no restart (a partial last
    restart block), lane groups that run two or three restarts each, and lane groups of two and
    four wavefronts (pose sums across wavefronts through LDS); early stop on and off.
  * Clouds beyond one workgroup's LDS (config_large_*.npz, the oracle's alignments computed once
    by tests/golden/make_config_vectors.py): a 12000-point surface + ESP against a 300-point one,
    two 6000-point surfaces, a 9000-atom against a 2000-atom volume.  fss_align takes the
    streamed path by itself, with lane groups of up to 1024 lanes so that the few poses fill the
    GPU; the per-pose loss and SE(3) gradient at the seeds (the one-pose kernel, streamed as
    well) are checked against the oracle too.
"""
import numpy as np
import pytest

from fast_shepherd_score_amd import synth
from fast_shepherd_score_amd.align import OptimSpec, align
from fast_shepherd_score_amd.constants import LAM_SCALING
from fast_shepherd_score_amd.seeds import rotmat
from oracle import fss_oracle as orc
from tests._cases import load_config, to_dev
from tests.test_gpu_configs import build, check_best, check_final_poses, check_seed_losses, opt_of

pytestmark = pytest.mark.gpu


def small_case(kind, n=3, B=6, seed=31):
    if kind == "esp":
        r, re, f, fe = synth.surf_esp_pairs(n, atoms=(15, 40), points=(150, 260), seed=seed)
        d = dict(kind="esp", ref=r, fit=f, ref_w=re, fit_w=fe, lam=0.3 * LAM_SCALING)
    elif kind == "shape_pharm":
        ps = synth.pharm_shape_pairs(n, atoms=(40, 120), pharms=(4, 14), seed=seed)
        r, f = ps["ref_atoms"], ps["fit_atoms"]
        d = dict(kind="shape_pharm", ref=r, fit=f, ref_types=ps["ref_types"], fit_types=ps["fit_types"],
                 ref_anc=ps["ref_anchors"], fit_anc=ps["fit_anchors"], ref_vec=ps["ref_vectors"],
                 fit_vec=ps["fit_vectors"], directionless=False)
    else:
        r, f = synth.vol_pairs(n, 40, 120, seed=seed)
        d = dict(kind="shape", ref=r, fit=f)
    rng = np.random.default_rng(seed + 1)
    q0 = synth.random_quaternions(rng, n * B).reshape(n, B, 4)
    rc = np.stack([x.mean(0) for x in r]).astype(np.float64)
    fc = np.stack([x.mean(0) for x in f]).astype(np.float64)
    t0 = (rc[:, None] - np.einsum("nbij,nj->nbi", rotmat(q0.astype(np.float64)), fc)).astype(np.float32)
    d.update(q0=q0, t0=t0)
    return d


@pytest.mark.parametrize("kind", ["esp", "shape", "shape_pharm"])
@pytest.mark.parametrize("early_stop", [False, True])
def test_forced_streaming_matches_oracle_and_staged(dev, kind, early_stop):
    d = small_case(kind)
    print(f"case streamed-forced {kind} early_stop={early_stop}: 3 pairs x 6 restarts x 40 steps", flush=True)
    pb, spec, pairs = build(d, dev)
    q0, t0 = to_dev(d["q0"], dev), to_dev(d["t0"], dev)
    opt = OptimSpec(max_steps=40, lr=0.1, early_stop=early_stop)
    staged = align(pb, spec, q0, t0, opt, variant=2, lanes_per_pose=16, restarts_per_wg=6)
    assert staged.extra["variant"] == 2
    oracle = [orc.optimize(p, d["q0"][i], d["t0"][i], lr=0.1, max_steps=40, early_stop=early_stop)
              for i, p in enumerate(pairs)]
    want = np.array([o["score"] for o in oracle])
    tilings = [dict(lanes_per_pose=16, restarts_per_wg=6),
               dict(lanes_per_pose=64, restarts_per_wg=6, groups_per_wg=3),
               # lane groups of two wavefronts, each group running three restarts in turn
               dict(lanes_per_pose=128, restarts_per_wg=6, groups_per_wg=2)]
    if not early_stop:   # 6 restarts in blocks of 4: the second block's last two groups idle
        tilings += [dict(lanes_per_pose=32, restarts_per_wg=4), dict(lanes_per_pose=256, restarts_per_wg=4)]
    for kw in tilings:
        res = align(pb, spec, q0, t0, opt, variant=4, tile_rows=48, **kw)
        assert res.extra["variant"] == 4 and res.extra["tile_ref"] == res.extra["tile_fit"] == 48
        got = res.score.cpu().numpy()
        err = np.abs(got - want)
        print(f"  {kw}: |streamed - FP64| max {err.max():.2e}, |streamed - staged| max "
              f"{np.abs(got - staged.score.cpu().numpy()).max():.2e}", flush=True)
        assert (err <= 1e-3).all(), (kw, got, want)
        np.testing.assert_allclose(got, staged.score.cpu().numpy(), atol=1e-4)
        check_final_poses(res, pairs)


@pytest.mark.parametrize("name", ["large_esp", "large_vol"])
def test_clouds_beyond_lds_match_fp64(dev, name):
    d = load_config(name)
    print(f"case {name}: ref {list(d['ref_n'])} x fit {list(d['fit_n'])} rows, "
          f"{int(d['restarts'])} restarts x {int(d['steps'])} steps", flush=True)
    pb, spec, pairs = build(d, dev)
    check_seed_losses(pb, spec, pairs, d, dev)
    res = align(pb, spec, to_dev(d["q0"], dev), to_dev(d["t0"], dev), opt_of(d))
    print(f"  launch: {res.extra}", flush=True)
    assert res.extra["variant"] == 4 and res.extra["tile_ref"] > 0
    check_best(name, res, d)
    check_final_poses(res, pairs)
    print(f"  steps executed: HIP {res.n_steps.cpu().numpy()}, oracle {d['n_steps']}", flush=True)
