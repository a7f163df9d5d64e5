"""ShaEP score (vol_and_surf_esp, FSS_MODE_ESP_COMBO) on the GPU against the FP64 oracle, whose
gradient is torch autograd of the restated score (the reference optimiser's own method):
self-overlaps and normalisers, the ESP channel's value and analytic gradient per pose, the
per-pose loss and SE(3) gradient, FP64 re-scoring, the persistent optimiser and the container
method (needs an MI355X).  Tolerances: per-pose scores 1e-4, best-restart scores 1e-3."""
import numpy as np
import pytest
import torch

from fast_shepherd_score_amd import _lib, synth
from fast_shepherd_score_amd.align import (OptimSpec, ScoreSpec, align, overlap_grad, pose_loss_grad,
                                           pose_score, self_overlaps)
from fast_shepherd_score_amd.constants import LAM_SCALING
from fast_shepherd_score_amd.packing import CloudSet, PairBatch, ShaepSet
from oracle import fss_oracle as orc
from tests._cases import random_seeds, to_dev

pytestmark = pytest.mark.gpu
DT = torch.float64


def _case(n, seed, dev, alpha=0.81, w=0.5):
    rng = np.random.default_rng(seed)
    mols = [(synth.shaep_molecule(rng, int(rng.integers(8, 16)), int(rng.integers(60, 120))),
             synth.shaep_molecule(rng, int(rng.integers(8, 16)), int(rng.integers(60, 120)))) for _ in range(n)]
    key = "atom_pos" if alpha == 0.81 else "surf_pos"
    side = lambda k, i: [m[i][k] for m in mols]
    se = lambda i: ShaepSet.from_arrays(side("surf_pos", i), side("surf_esp", i), side("centers_w_H", i),
                                        side("partial_charges", i), side("radii", i), device=dev)
    pb = PairBatch(CloudSet.from_arrays(side(key, 0), device=dev), CloudSet.from_arrays(side(key, 1), device=dev),
                   ref_se=se(0), fit_se=se(1))
    spec = ScoreSpec(mode=_lib.MODE_ESP_COMBO, alpha=alpha, lam=LAM_SCALING * 0.001, esp_weight=w)
    pairs = [orc.ComboPair(ref_centers=a[key], fit_centers=b[key], ref_surf=a["surf_pos"], fit_surf=b["surf_pos"],
                           ref_surf_esp=a["surf_esp"], fit_surf_esp=b["surf_esp"], ref_atoms=a["centers_w_H"],
                           fit_atoms=b["centers_w_H"], ref_charges=a["partial_charges"],
                           fit_charges=b["partial_charges"], ref_radii=a["radii"], fit_radii=b["radii"],
                           alpha=alpha, esp_weight=w) for a, b in mols]
    return pb, spec, pairs, mols


def _poses(n, B, seed, scale=1.0):
    q, _ = random_seeds(n, B, seed=seed)
    t = (np.random.default_rng(seed + 1).normal(size=(n, B, 3)) * scale).astype(np.float32)
    return q, t


def test_self_overlaps_and_normalisers(dev):
    pb, spec, pairs, mols = _case(4, 1, dev)
    so = self_overlaps(pb, spec).cpu().numpy()
    for i, p in enumerate(pairs):
        np.testing.assert_allclose(so[i, 0], float(orc.vab(p.tref_centers[None], p.tref_centers[None], 0.81)[0]), rtol=1e-12)
        np.testing.assert_allclose(so[i, 1], float(orc.vab(p.tfit_centers[None], p.tfit_centers[None], 0.81)[0]), rtol=1e-12)
        assert so[i, 2] == len(mols[i][0]["surf_pos"]) and so[i, 3] == len(mols[i][1]["surf_pos"])


def test_esp_channel_value_and_gradient_per_pose(dev):
    print("case shaep overlap_grad", flush=True)
    n, B = 3, 6
    pb, spec, pairs, _ = _case(n, 2, dev)
    q, t = _poses(n, B, 3)
    pp = torch.arange(n, dtype=torch.int32).repeat_interleave(B)
    V, gR, gT, _ = overlap_grad(pb, spec, pp, to_dev(q.reshape(-1, 4), dev), to_dev(t.reshape(-1, 3), dev))
    V, gR, gT = V.cpu().numpy(), gR.cpu().numpy(), gT.cpu().numpy()
    for i, p in enumerate(pairs):
        qn = torch.tensor(q[i], dtype=DT)
        R = orc.rot_from_unit_quat(qn / qn.norm(dim=1, keepdim=True))
        v, GR, GT = p.esp_value_grad(R, torch.tensor(t[i], dtype=DT))
        sl = slice(i * B, (i + 1) * B)
        np.testing.assert_allclose(V[sl, 1], v.numpy(), rtol=1e-4, atol=1e-3)
        scale = max(np.abs(GR.numpy()).max(), np.abs(GT.numpy()).max(), 1e-6)
        np.testing.assert_allclose(gR[sl, 1].reshape(B, 3, 3), GR.numpy(), atol=2e-4 * scale)
        np.testing.assert_allclose(gT[sl, 1], GT.numpy(), atol=2e-4 * scale)
        assert v.min() > 0                       # the poses overlap: the channel is exercised


def test_pose_loss_grad_and_fp64_rescoring(dev):
    print("case shaep pose_loss_grad", flush=True)
    n, B = 3, 5
    pb, spec, pairs, _ = _case(n, 4, dev)
    q, t = _poses(n, B, 5)
    q = q * 1.4                                  # raw quaternions, as the optimiser carries them
    so = self_overlaps(pb, spec)
    pp = torch.arange(n, dtype=torch.int32).repeat_interleave(B)
    qd, td = to_dev(q.reshape(-1, 4), dev), to_dev(t.reshape(-1, 3), dev)
    loss, grad = pose_loss_grad(pb, spec, pp, B, qd, td, so)
    S = pose_score(pb, spec, pp, qd, td, so).cpu().numpy()
    loss, grad = loss.cpu().numpy(), grad.cpu().numpy()
    for i, p in enumerate(pairs):
        se3 = torch.cat([torch.tensor(q[i], dtype=DT), torch.tensor(t[i], dtype=DT)], 1)
        _, g, per = p.loss_grad_se3(se3, {})
        sl = slice(i * B, (i + 1) * B)
        np.testing.assert_allclose(loss[sl], per.numpy(), atol=1e-4)
        scale = max(np.abs(g.numpy()).max(), 1e-8)
        np.testing.assert_allclose(grad[sl], g.numpy(), atol=3e-4 * scale)
        np.testing.assert_allclose(S[sl], p.score_poses(se3).detach().numpy(), atol=1e-9)


@pytest.mark.parametrize("early_stop", [False, True])
@pytest.mark.parametrize("alpha", [0.81, 1.0])
def test_persistent_align_matches_oracle(dev, early_stop, alpha):
    print(f"case shaep align es={early_stop} alpha={alpha}", flush=True)
    n, B, steps = 4, 6, 60
    pb, spec, pairs, _ = _case(n, 6 + int(early_stop), dev, alpha=alpha)
    q, t = _poses(n, B, 8, scale=0.5)
    res = align(pb, spec, to_dev(q, dev), to_dev(t, dev), OptimSpec(max_steps=steps, early_stop=early_stop))
    score = res.score.cpu().numpy()
    for i, p in enumerate(pairs):
        o = orc.optimize(p, q[i], t[i], lr=0.1, max_steps=steps, early_stop=early_stop)
        assert abs(score[i] - o["score"]) < 1e-3, (i, score[i], o["score"])


def test_container_vol_and_surf_esp_matches_oracle(dev):
    from fast_shepherd_score_amd.container import Molecule, MoleculePair, MoleculePairBatch
    from fast_shepherd_score_amd.seeds import initial_seeds, pad
    print("case shaep container", flush=True)
    _, _, pairs_o, mols = _case(3, 9, dev)
    mk = lambda m: Molecule(atom_pos=m["atom_pos"], surf_pos=m["surf_pos"], surf_esp=m["surf_esp"],
                            centers_w_H=m["centers_w_H"], partial_charges=m["partial_charges"], radii=m["radii"])
    pairs = [MoleculePair(mk(a), mk(b), device=dev) for a, b in mols]
    R, S = 8, 50
    scores, aligned = MoleculePairBatch(pairs).align_with_vol_and_surf_esp(
        alpha=0.81, num_repeats=R, max_num_steps=S, backend="hip", return_aligned=True)
    assert scores.shape == (3,) and np.all(scores > 0)
    for i, (a, b) in enumerate(mols):
        T = pairs[i].transform_vol_and_surf_esp.numpy()
        np.testing.assert_allclose(aligned[i], b["surf_pos"] @ T[:3, :3].T + T[:3, 3], atol=1e-4)
        (Rp, rn), (Fp, fn) = pad([a["surf_pos"]]), pad([b["surf_pos"]])
        q0, t0 = initial_seeds(Rp, rn, Fp, fn, R)
        o = orc.optimize(pairs_o[i], q0[0], t0[0], lr=0.1, max_steps=S)
        assert abs(scores[i] - o["score"]) < 1e-3, (i, scores[i], o["score"])
        assert pairs[i].sim_aligned_vol_and_surf_esp == scores[i]
