"""The reference-named scoring functions (fast_shepherd_score_amd.score) on the GPU against the
FP64 oracle's scoring restatement (needs an MI355X).  Tolerances: FP32 similarities 1e-5
relative to the value, FP64 ones 1e-9; gradients of the FP64 kernels 1e-9 relative to the
largest component."""
import numpy as np
import pytest
import torch

from fast_shepherd_score_amd import score as fs
from fast_shepherd_score_amd import synth
from fast_shepherd_score_amd.constants import LAM_SCALING
from oracle import fss_oracle as orc

pytestmark = pytest.mark.gpu
DT = torch.float64


def _clouds(seed, B, N=30, M=40):
    rng = np.random.default_rng(seed)
    a = np.stack(synth.random_walks(rng, [N] * B)).astype(np.float32)
    b = np.stack(synth.random_walks(rng, [M] * B)).astype(np.float32)
    b += rng.normal(0, 0.7, size=(B, 1, 3)).astype(np.float32)
    qa = synth.esp_values(rng, B * N).reshape(B, N)
    qb = synth.esp_values(rng, B * M).reshape(B, M)
    return a, b, qa.astype(np.float32), qb.astype(np.float32)


def test_shape_and_esp_values_single_and_batched(dev):
    print("case score api shape/esp", flush=True)
    a, b, qa, qb = _clouds(1, 3)
    A, Bt = torch.tensor(a, dtype=DT), torch.tensor(b, dtype=DT)
    QA, QB = torch.tensor(qa, dtype=DT), torch.tensor(qb, dtype=DT)
    lam = 0.3 * LAM_SCALING
    ref_S = orc.tanimoto_shape(A, Bt, 0.81).numpy()
    ref_E = orc.tanimoto_esp(A, Bt, QA, QB, 0.81, lam).numpy()
    ref_V = orc.vab(A, Bt, 0.81).numpy()
    w = torch.exp(-((QA.unsqueeze(2) - QB.unsqueeze(1)) ** 2) / lam)
    ref_VE = orc.vab(A, Bt, 0.81, w).numpy()
    S = fs.get_overlap(torch.from_numpy(a), torch.from_numpy(b))               # CPU tensors in, CPU out
    assert S.device.type == "cpu" and S.dtype == torch.float32 and S.shape == (3,)
    np.testing.assert_allclose(S.numpy(), ref_S, rtol=1e-5)
    np.testing.assert_allclose(fs.get_overlap(a[1], b[1]).numpy(), ref_S[1], rtol=1e-5)   # numpy, single
    assert fs.get_overlap(a[1], b[1]).dim() == 0
    np.testing.assert_allclose(fs.VAB_2nd_order(a, b, 0.81).numpy(), ref_V, rtol=1e-5)
    np.testing.assert_allclose(fs.get_overlap_esp(a, b, qa, qb).numpy(), ref_E, rtol=1e-5)
    np.testing.assert_allclose(fs.VAB_2nd_order_esp(a, b, qa, qb, 0.81, lam).numpy(), ref_VE, rtol=1e-5)
    # float64 inputs run the FP64 kernels
    S64 = fs.get_overlap(A.to(dev), Bt.to(dev))
    assert S64.dtype == torch.float64 and S64.is_cuda
    np.testing.assert_allclose(S64.cpu().numpy(), ref_S, rtol=1e-12)


def test_get_overlap_gradients_match_autograd_of_the_oracle(dev):
    print("case score api gradients", flush=True)
    a, b, qa, qb = _clouds(2, 2, N=12, M=15)
    for esp in (False, True):
        A = torch.tensor(a, dtype=DT, device=dev, requires_grad=True)
        Bt = torch.tensor(b, dtype=DT, device=dev, requires_grad=True)
        if esp:
            S = fs.get_overlap_esp(A, Bt, torch.tensor(qa, dtype=DT), torch.tensor(qb, dtype=DT))
        else:
            S = fs.get_overlap(A, Bt)
        (S * torch.tensor([1.0, -2.0], dtype=DT, device=dev)).sum().backward()
        Ao = torch.tensor(a, dtype=DT, requires_grad=True)
        Bo = torch.tensor(b, dtype=DT, requires_grad=True)
        if esp:
            So = orc.tanimoto_esp(Ao, Bo, torch.tensor(qa, dtype=DT), torch.tensor(qb, dtype=DT), 0.81,
                                  0.3 * LAM_SCALING)
        else:
            So = orc.tanimoto_shape(Ao, Bo, 0.81)
        (So * torch.tensor([1.0, -2.0], dtype=DT)).sum().backward()
        for g, go in ((A.grad, Ao.grad), (Bt.grad, Bo.grad)):
            scale = go.abs().max().item()
            np.testing.assert_allclose(g.cpu().numpy(), go.numpy(), atol=1e-9 * scale)


def _pharm(seed, B, n=9, m=12):
    rng = np.random.default_rng(seed)
    atoms = synth.random_walks(rng, [20] * (2 * B))
    ph1 = [synth.pharmacophores(rng, atoms[i], n) for i in range(B)]
    ph2 = [synth.pharmacophores(rng, atoms[B + i], m) for i in range(B)]
    t1 = np.stack([p[0] for p in ph1]); a1 = np.stack([p[1] for p in ph1]).astype(np.float32)
    v1 = np.stack([p[2] for p in ph1]).astype(np.float32)
    t2 = np.stack([p[0] for p in ph2]); a2 = np.stack([p[1] for p in ph2]).astype(np.float32)
    v2 = np.stack([p[2] for p in ph2]).astype(np.float32)
    t2[:, : n // 2] = t1[:, : n // 2]       # shared types so every channel is populated
    c1 = np.stack(atoms[:B]).astype(np.float32)
    c2 = np.stack(atoms[B:]).astype(np.float32)
    return t1, t2, a1, a2, v1, v2, c1, c2


@pytest.mark.parametrize("kw", [dict(), dict(similarity="tversky_ref"), dict(extended_points=True),
                                dict(extended_points=True, only_extended=True, similarity="tversky"),
                                dict(directionless=True)],
                         ids=["tanimoto", "tversky_ref", "extended", "only_extended", "directionless"])
def test_get_overlap_pharm_matches_oracle(dev, kw):
    print(f"case score api pharm {kw}", flush=True)
    B = 3
    t1, t2, a1, a2, v1, v2, _, _ = _pharm(3, B)
    got = fs.get_overlap_pharm(t1, t2, a1, a2, v1, v2, **kw)
    assert got.shape == (B,) and got.dtype == torch.float32
    T = lambda x: torch.tensor(x, dtype=DT)
    for i in range(B):
        ref = orc.pharm_similarity(torch.tensor(t1[i]), torch.tensor(t2[i]), T(a1[i]), T(a2[i])[None], T(v1[i]),
                                   T(v2[i])[None], **kw)
        assert abs(float(got[i]) - float(ref[0])) < 1e-6, (i, float(got[i]), float(ref[0]))
    single = fs.get_overlap_pharm(t1[0], t2[0], a1[0], a2[0], v1[0], v2[0], **kw)
    assert single.dim() == 0 and abs(float(single) - float(got[0])) < 1e-7


def test_get_overlap_pharm_precomputed_self_overlaps(dev):
    print("case score api pharm precomputed", flush=True)
    t1, t2, a1, a2, v1, v2, _, _ = _pharm(4, 2)
    s0 = fs.get_overlap_pharm(t1, t2, a1, a2, v1, v2)
    T = lambda x: torch.tensor(x, dtype=DT)
    vaa = [[float(orc.overlap_grad_pharm(torch.eye(3, dtype=DT)[None], torch.zeros((1, 3), dtype=DT),
                                         torch.tensor(t[i])[None], torch.tensor(t[i])[None], T(a[i])[None],
                                         T(a[i])[None], T(v[i])[None], T(v[i])[None])[0][0])
            for i in range(2)] for t, a, v in ((t1, a1, v1), (t2, a2, v2))]
    s1 = fs.get_overlap_pharm(t1, t2, a1, a2, v1, v2, precomputed_self_overlaps=(vaa[0], vaa[1]))
    np.testing.assert_allclose(s1.numpy(), s0.numpy(), rtol=1e-6)
    # doubled self terms: VAB / (2 U - VAB), from the same VAB
    s2 = fs.get_overlap_pharm(t1, t2, a1, a2, v1, v2,
                              precomputed_self_overlaps=([2 * x for x in vaa[0]], [2 * x for x in vaa[1]]))
    U = np.array(vaa[0]) + np.array(vaa[1])
    vab = s0.numpy().astype(np.float64) * U / (1 + s0.numpy())
    np.testing.assert_allclose(s2.numpy(), vab / (2 * U - vab), rtol=1e-5)
    with pytest.raises(ValueError):
        fs.get_overlap_pharm(t1, t2, a1, a2, v1, v2, precomputed_self_overlaps=(vaa[0], vaa[1]), directionless=True)


@pytest.mark.parametrize("similarity", ["tanimoto", "tversky_fit"])
def test_get_pharm_combo_score_matches_oracle(dev, similarity):
    print(f"case score api combo {similarity}", flush=True)
    B = 2
    t1, t2, a1, a2, v1, v2, c1, c2 = _pharm(5, B)
    got = fs.get_pharm_combo_score(c1, c2, t1, t2, a1, a2, v1, v2, similarity=similarity, color_weight=0.3,
                                   extended_points=True)
    T = lambda x: torch.tensor(x, dtype=DT)
    for i in range(B):
        ab, aa, bb = (orc.vab(T(x)[None], T(y)[None], 0.81) for x, y in ((c1[i], c2[i]), (c1[i], c1[i]), (c2[i], c2[i])))
        shape = ab / (aa + bb - ab) if similarity == "tanimoto" else torch.clamp_max(ab / (0.05 * aa + 0.95 * bb), 1.0)
        ph = orc.pharm_similarity(torch.tensor(t1[i]), torch.tensor(t2[i]), T(a1[i]), T(a2[i])[None], T(v1[i]),
                                  T(v2[i])[None], similarity=similarity, extended_points=True)
        ref = 0.7 * float(shape[0]) + 0.3 * float(ph[0])
        assert abs(float(got[i]) - ref) < 1e-5, (i, float(got[i]), ref)


def test_esp_combo_score_matches_oracle(dev):
    print("case score api esp_combo", flush=True)
    rng = np.random.default_rng(6)
    mols = [(synth.shaep_molecule(rng, 12, 80), synth.shaep_molecule(rng, 12, 80)) for _ in range(2)]
    # a batch needs equal sizes: keep the first 20 atoms with hydrogens (12 heavy + 8 H) of each
    for pair in mols:
        for m in pair:
            for k in ("centers_w_H", "partial_charges", "radii"):
                m[k] = m[k][:20]
    st = lambda k, i: np.stack([m[i][k] for m in mols])
    got = fs.esp_combo_score(st("centers_w_H", 0), st("centers_w_H", 1), st("atom_pos", 0), st("atom_pos", 1),
                             st("surf_pos", 0), st("surf_pos", 1), st("partial_charges", 0),
                             st("partial_charges", 1), st("surf_esp", 0), st("surf_esp", 1), st("radii", 0),
                             st("radii", 1), alpha=0.81)
    ident = torch.tensor([[1.0, 0, 0, 0, 0, 0, 0]], dtype=DT)
    for i, (a, b) in enumerate(mols):
        p = orc.ComboPair(ref_centers=a["atom_pos"], fit_centers=b["atom_pos"], ref_surf=a["surf_pos"],
                          fit_surf=b["surf_pos"], ref_surf_esp=a["surf_esp"], fit_surf_esp=b["surf_esp"],
                          ref_atoms=a["centers_w_H"], fit_atoms=b["centers_w_H"], ref_charges=a["partial_charges"],
                          fit_charges=b["partial_charges"], ref_radii=a["radii"], fit_radii=b["radii"])
        ref = float(p.score_poses(ident)[0])
        assert abs(float(got[i]) - ref) < 1e-6, (i, float(got[i]), ref)
        one = fs.esp_combo_score(a["centers_w_H"], b["centers_w_H"], a["atom_pos"], b["atom_pos"], a["surf_pos"],
                                 b["surf_pos"], a["partial_charges"], b["partial_charges"], a["surf_esp"],
                                 b["surf_esp"], a["radii"], b["radii"], alpha=0.81)
        assert one.dim() == 0 and abs(float(one) - ref) < 1e-6
