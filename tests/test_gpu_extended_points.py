"""Pharmacophore extended points on the GPU vs the FP64 oracle (needs an MI355X).

With ``extended_points`` the acceptor, donor and halogen pairs lose their cosine weight (anchor
weight 1, or 0 with ``only_extended``) and add an isotropic overlap of the extended points
anchor + unit vector (pharmacophore_scoring.py:284-370, analytical_gradients/_torch.py:288-357).
Tolerances as for every other channel: per-pose scores within 1e-4 absolute (FP64 re-scoring:
1e-9), best-restart scores within 1e-3, FP32 gradients within 1e-4 of the largest component.
"""
import numpy as np
import pytest
import torch

from fast_shepherd_score_amd import _lib
from fast_shepherd_score_amd.align import (OptimSpec, align, overlap_grad, pose_loss_grad, pose_score,
                                           self_overlaps)
from oracle import fss_oracle as orc
from tests._cases import make_case, random_seeds, to_dev

pytestmark = pytest.mark.gpu
DT = torch.float64
# (mode, only_extended, similarity, vector scale)
CASES = [(_lib.MODE_PHARM, False, "tanimoto", 1.0), (_lib.MODE_PHARM, True, "tanimoto", 1.0),
         (_lib.MODE_PHARM, False, "tversky", 2.5), (_lib.MODE_SHAPE_PHARM, False, "tanimoto", 1.0),
         (_lib.MODE_SHAPE_PHARM, True, "tversky_fit", 0.5)]
IDS = ["pharm", "pharm-only", "pharm-tversky-scaled", "shape_pharm", "shape_pharm-only-tversky"]


def _case(dev, c, n, seed):
    mode, only, sim, scale = c
    return make_case(mode, n, dev, seed=seed, similarity=sim, extended_points=True, only_extended=only,
                     vec_scale=scale, pharms=(6, 14))


def _se3(q, t):
    return torch.cat([torch.as_tensor(q, dtype=DT), torch.as_tensor(t, dtype=DT)], dim=-1)


@pytest.mark.parametrize("c", CASES, ids=IDS)
def test_extended_self_overlaps(dev, c):
    print(f"case ext self_overlap {c}", flush=True)
    pb, spec, pairs = _case(dev, c, 5, seed=11)
    so = self_overlaps(pb, spec).cpu().numpy()
    for i, p in enumerate(pairs):
        o = p.self_overlaps()
        np.testing.assert_allclose(so[i, 2:], [float(o["VAA_p"]), float(o["VBB_p"])], rtol=1e-6)


@pytest.mark.parametrize("c", CASES, ids=IDS)
def test_extended_overlap_and_gradients_per_pose(dev, c):
    print(f"case ext overlap_grad {c}", flush=True)
    n, B = 4, 6
    pb, spec, pairs = _case(dev, c, n, seed=12)
    q, t = random_seeds(n, B, seed=13, t_scale=0.5)
    pp = torch.arange(n, dtype=torch.int32).repeat_interleave(B)
    V, gR, gT, _ = overlap_grad(pb, spec, pp, to_dev(q.reshape(-1, 4), dev), to_dev(t.reshape(-1, 3), dev))
    V, gR, gT = V.cpu().numpy(), gR.cpu().numpy(), gT.cpu().numpy()
    for i, p in enumerate(pairs):
        qn = torch.tensor(q[i], dtype=DT)
        R = orc.rot_from_unit_quat(qn / qn.norm(dim=1, keepdim=True))
        O, GR, GT = (x.numpy() for x in p.channel_overlaps(R, torch.tensor(t[i], dtype=DT))["pharm"])
        sl = slice(i * B, (i + 1) * B)
        np.testing.assert_allclose(V[sl, 1], O, rtol=2e-5, atol=1e-6)
        scale = max(np.abs(GR).max(), np.abs(GT).max(), 1e-6)
        np.testing.assert_allclose(gR[sl, 1], GR, atol=1e-4 * scale)
        np.testing.assert_allclose(gT[sl, 1], GT, atol=1e-4 * scale)


@pytest.mark.parametrize("c", CASES, ids=IDS)
def test_extended_loss_grad_and_score(dev, c):
    print(f"case ext pose_loss_grad {c}", flush=True)
    n, B = 3, 5
    pb, spec, pairs = _case(dev, c, n, seed=14)
    q, t = random_seeds(n, B, seed=15, t_scale=0.5)
    q = q * 1.3
    so = self_overlaps(pb, spec)
    pp = torch.arange(n, dtype=torch.int32).repeat_interleave(B)
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
        # FP64 re-scoring against the scoring restatement (pharmacophore_scoring.py:284-370)
        np.testing.assert_allclose(S[sl], p.score_poses(se3).numpy(), atol=1e-9)


@pytest.mark.parametrize("c", CASES, ids=IDS)
def test_extended_persistent_align_matches_oracle(dev, c):
    print(f"case ext align {c}", flush=True)
    n, B, steps = 5, 10, 100
    pb, spec, pairs = _case(dev, c, n, seed=16)
    q, t = random_seeds(n, B, seed=17)
    res = align(pb, spec, to_dev(q, dev), to_dev(t, dev), OptimSpec(max_steps=steps, lr=0.1))
    torch.cuda.synchronize()
    score = res.score.cpu().numpy()
    for i, p in enumerate(pairs):
        o = orc.optimize(p, q[i], t[i], lr=0.1, max_steps=steps)
        assert abs(score[i] - o["score"]) < 1e-3, (i, score[i], o["score"])


def test_extended_changes_the_score(dev):
    """The three pharmacophore forms give different scores on directional features, and
    directionless scoring ignores the extended-point switch (pharmacophore_scoring.py:517-523)."""
    print("case ext forms differ", flush=True)
    n, B = 3, 4
    q, t = random_seeds(n, B, seed=19, t_scale=0.3)
    pp = torch.arange(n, dtype=torch.int32).repeat_interleave(B)
    out = {}
    for key, kw in {"cos": {}, "ext": dict(extended_points=True),
                    "only": dict(extended_points=True, only_extended=True),
                    "dl": dict(directionless=True),
                    "dl_ext": dict(directionless=True, extended_points=True)}.items():
        pb, spec, _ = make_case(_lib.MODE_PHARM, n, dev, seed=18, pharms=(6, 14), **kw)
        so = self_overlaps(pb, spec)
        out[key] = pose_score(pb, spec, pp, to_dev(q.reshape(-1, 4), dev), to_dev(t.reshape(-1, 3), dev),
                              so).cpu().numpy()
    assert np.abs(out["cos"] - out["ext"]).max() > 1e-3
    assert np.abs(out["ext"] - out["only"]).max() > 1e-3
    np.testing.assert_array_equal(out["dl"], out["dl_ext"])
