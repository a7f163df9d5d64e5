"""The drop-in container: MoleculePairBatch.align_with_* through backend="hip" reproduces the
FP64 oracle optimiser run from the same default multi-start seeds, writes the reference's result
attributes, and returns aligned clouds consistent with the stored transforms (needs an MI355X)."""
import numpy as np
import pytest
import torch

from fast_shepherd_score_amd import synth
from fast_shepherd_score_amd.constants import LAM_SCALING
from fast_shepherd_score_amd.container import Molecule, MoleculePair, MoleculePairBatch
from fast_shepherd_score_amd.seeds import initial_seeds, pad
from oracle import fss_oracle as orc

pytestmark = pytest.mark.gpu


def _mols(n, seed):
    d = synth.pharm_shape_pairs(n, atoms=(12, 30), pharms=(4, 10), seed=seed)
    rng = np.random.default_rng(seed)
    out = []
    for side in ("ref", "fit"):
        ms = []
        for i in range(n):
            atoms = d[f"{side}_atoms"][i]
            surf = synth.surface_points(rng, atoms, 60)
            ms.append(Molecule(atom_pos=atoms, surf_pos=surf, surf_esp=synth.esp_values(rng, 60),
                               pharm_types=d[f"{side}_types"][i], pharm_ancs=d[f"{side}_anchors"][i],
                               pharm_vecs=d[f"{side}_vectors"][i],
                               atom_charges=synth.esp_values(rng, len(atoms)) * 3))
        out.append(ms)
    return out


@pytest.mark.parametrize("mode", ["vol", "surf_esp", "pharm", "vol_color", "vol_esp", "pharm_ext",
                                  "pharm_only_ext", "vol_color_ext"])
def test_align_with_matches_oracle(dev, mode):
    print(f"case container mode={mode}", flush=True)
    ext = dict(extended_points=mode.endswith("ext"), only_extended=mode == "pharm_only_ext")
    mode = mode.replace("_only_ext", "").replace("_ext", "")
    refs, fits = _mols(4, 30)
    pairs = [MoleculePair(r, f, device=dev) for r, f in zip(refs, fits)]
    batch = MoleculePairBatch(pairs)
    R, S = 12, 60
    if mode == "vol":
        scores, aligned = batch.align_with_vol(num_repeats=R, max_num_steps=S, backend="hip", return_aligned=True)
        attr, cloud = "transform_vol_noH", "atom_pos"
    elif mode == "vol_esp":
        scores, aligned = batch.align_with_vol_esp(lam=0.1, num_repeats=R, max_num_steps=S, backend="hip", return_aligned=True)
        attr, cloud = "transform_vol_esp_noH", "atom_pos"
    elif mode == "surf_esp":
        scores, aligned = batch.align_with_surf_esp(alpha=0.81, lam=0.3, num_repeats=R, max_num_steps=S,
                                                    backend="triton", return_aligned=True)
        attr, cloud = "transform_surf_esp", "surf_pos"
    elif mode == "pharm":
        scores, aligned, vecs = batch.align_with_pharm(num_repeats=R, max_num_steps=S, backend="hip",
                                                       return_aligned=True, **ext)
        attr, cloud = "transform_pharm", "pharm_ancs"
    else:
        scores, aligned = batch.align_with_vol_color(num_repeats=R, max_num_steps=S, backend="hip",
                                                     return_aligned=True, directionless=False, **ext)
        attr, cloud = "transform_vol_color", "atom_pos"
    assert scores.dtype == np.float64 and scores.shape == (4,)
    for i, p in enumerate(pairs):
        T = getattr(p, attr)
        assert isinstance(T, torch.Tensor) and T.shape == (4, 4)
        X = getattr(p.fit_molec, cloud).astype(np.float64)
        np.testing.assert_allclose(aligned[i], X @ T.numpy()[:3, :3].T + T.numpy()[:3, 3], atol=1e-5)
        r, f = refs[i], fits[i]
        seed_r = r.pharm_ancs if mode == "pharm" else (r.surf_pos if mode == "surf_esp" else r.atom_pos)
        seed_f = f.pharm_ancs if mode == "pharm" else (f.surf_pos if mode == "surf_esp" else f.atom_pos)
        (Rp, rn), (Fp, fn) = pad([seed_r]), pad([seed_f])
        q0, t0 = initial_seeds(Rp, rn, Fp, fn, R)
        kw = dict(ref_types=r.pharm_types, fit_types=f.pharm_types, ref_anc=r.pharm_ancs,
                  fit_anc=f.pharm_ancs, ref_vec=r.pharm_vecs, fit_vec=f.pharm_vecs)
        if mode == "vol":
            pair = orc.Pair("shape", ref=r.atom_pos, fit=f.atom_pos)
        elif mode == "vol_esp":
            pair = orc.Pair("esp", ref=r.atom_pos, fit=f.atom_pos, ref_esp=r.atom_charges,
                            fit_esp=f.atom_charges, lam=0.1)
        elif mode == "surf_esp":
            pair = orc.Pair("esp", ref=r.surf_pos, fit=f.surf_pos, ref_esp=r.surf_esp, fit_esp=f.surf_esp,
                            lam=0.3 * LAM_SCALING)
        elif mode == "pharm":
            pair = orc.Pair("pharm", **kw, **ext)
        else:
            pair = orc.Pair("shape_pharm", ref=r.atom_pos, fit=f.atom_pos, **kw, **ext)
        o = orc.optimize(pair, q0[0], t0[0], lr=0.1, max_steps=S)
        assert abs(scores[i] - o["score"]) < 1e-3, (mode, i, scores[i], o["score"])


@pytest.mark.parametrize("mode", ["vol", "pharm"])
def test_device_and_host_seeds_agree(dev, mode):
    """Default (device) seeds and the host FP64 restatement of the seed rule give the same best
    scores: the same restart set, the principal-axis rows possibly in another order."""
    refs, fits = _mols(4, 32)
    res = {}
    for seeds in (None, "host"):
        pairs = [MoleculePair(r, f, device=dev) for r, f in zip(refs, fits)]
        batch = MoleculePairBatch(pairs)
        fn = batch.align_with_vol if mode == "vol" else batch.align_with_pharm
        out = fn(num_repeats=20, max_num_steps=60, backend="hip", seeds=seeds)
        res[seeds] = out[0]
    np.testing.assert_allclose(res[None], res["host"], atol=1e-3)


def test_defaults_and_errors(dev):
    refs, fits = _mols(2, 31)
    batch = MoleculePairBatch([MoleculePair(r, f, device=dev) for r, f in zip(refs, fits)])
    scores, _ = batch.align_with_vol()          # MODE_SEEDS / MODE_STEPS defaults, backend None -> hip
    assert np.all((scores > 0) & (scores <= 1 + 1e-9))
    assert batch.pairs[0].sim_aligned_vol_noH == scores[0]
    with pytest.raises(ValueError):
        batch.align_with_vol(backend="numba")


@pytest.mark.parametrize("mode", ["vol", "surf_esp", "pharm"])
def test_trans_init_matches_oracle(dev, mode):
    """trans_init=True: the translation-centre seeds (5 + 10 per ref atom, or per ref anchor in
    pharm mode), one launch per restart count, against the FP64 oracle from the same seeds."""
    from fast_shepherd_score_amd.seeds import translation_seeds
    print(f"case container trans_init mode={mode}", flush=True)
    refs, fits = _mols(3, 34)
    pairs = [MoleculePair(r, f, device=dev) for r, f in zip(refs, fits)]
    batch = MoleculePairBatch(pairs)
    S = 25
    if mode == "vol":
        scores, _ = batch.align_with_vol(trans_init=True, max_num_steps=S, backend="hip")
    elif mode == "surf_esp":
        scores, _ = batch.align_with_surf_esp(alpha=0.81, lam=0.3, trans_init=True, max_num_steps=S, backend="hip")
    else:
        scores, _, _ = batch.align_with_pharm(trans_init=True, max_num_steps=S, backend="hip")
    assert len({len(r.atom_pos) for r in refs}) > 1 or mode == "pharm"     # several launch groups
    for i, (r, f) in enumerate(zip(refs, fits)):
        if mode == "vol":
            q0, t0 = translation_seeds(r.atom_pos, f.atom_pos, r.atom_pos)
            pair = orc.Pair("shape", ref=r.atom_pos, fit=f.atom_pos)
        elif mode == "surf_esp":
            q0, t0 = translation_seeds(r.surf_pos, f.surf_pos, r.atom_pos)
            pair = orc.Pair("esp", ref=r.surf_pos, fit=f.surf_pos, ref_esp=r.surf_esp, fit_esp=f.surf_esp,
                            lam=0.3 * LAM_SCALING)
        else:
            q0, t0 = translation_seeds(r.pharm_ancs, f.pharm_ancs, r.pharm_ancs)
            pair = orc.Pair("pharm", ref_types=r.pharm_types, fit_types=f.pharm_types, ref_anc=r.pharm_ancs,
                            fit_anc=f.pharm_ancs, ref_vec=r.pharm_vecs, fit_vec=f.pharm_vecs)
        o = orc.optimize(pair, q0, t0, lr=0.1, max_steps=S)
        assert abs(scores[i] - o["score"]) < 1e-3, (mode, i, scores[i], o["score"])


def _with_hydrogens(ms, seed):
    """Give each molecule atoms with hydrogens (heavy atoms + one H 1.0 A off each of a few of
    them) and partial charges for every atom."""
    rng = np.random.default_rng(seed)
    for m in ms:
        k = max(2, len(m.atom_pos) // 2)
        h = m.atom_pos[rng.integers(0, len(m.atom_pos), size=k)] + synth.random_quaternions(rng, k)[:, 1:] * 1.0
        m.centers_w_H = np.concatenate([m.atom_pos, h]).astype(np.float32)
        m.partial_charges = synth.esp_values(rng, len(m.centers_w_H)) * 3
    return ms


@pytest.mark.parametrize("mode", ["vol", "vol_esp"])
def test_batch_with_hydrogens_matches_oracle(dev, mode):
    """no_H=False aligns the atoms with hydrogens and writes transform_vol[_esp] /
    sim_aligned_vol[_esp] (_core.py:661-710)."""
    print(f"case container withH mode={mode}", flush=True)
    refs, fits = _mols(3, 35)
    _with_hydrogens(refs, 1), _with_hydrogens(fits, 2)
    pairs = [MoleculePair(r, f, device=dev) for r, f in zip(refs, fits)]
    batch = MoleculePairBatch(pairs)
    R, S = 10, 50
    fn = batch.align_with_vol if mode == "vol" else batch.align_with_vol_esp
    scores, aligned = fn(no_H=False, num_repeats=R, max_num_steps=S, backend="hip", return_aligned=True)
    attr = "transform_vol" if mode == "vol" else "transform_vol_esp"
    for i, p in enumerate(pairs):
        assert getattr(p, attr + "_noH") is None
        assert getattr(p, "sim_aligned" + attr[9:]) == scores[i]
        T = getattr(p, attr).numpy()
        np.testing.assert_allclose(aligned[i], fits[i].centers_w_H @ T[:3, :3].T + T[:3, 3], atol=1e-5)
        r, f = refs[i], fits[i]
        (Rp, rn), (Fp, fn_) = pad([r.centers_w_H]), pad([f.centers_w_H])
        q0, t0 = initial_seeds(Rp, rn, Fp, fn_, R)
        if mode == "vol":
            pair = orc.Pair("shape", ref=r.centers_w_H, fit=f.centers_w_H)
        else:
            pair = orc.Pair("esp", ref=r.centers_w_H, fit=f.centers_w_H, ref_esp=r.partial_charges,
                            fit_esp=f.partial_charges, lam=0.1)
        o = orc.optimize(pair, q0[0], t0[0], lr=0.1, max_steps=S)
        assert abs(scores[i] - o["score"]) < 1e-3, (mode, i, scores[i], o["score"])


@pytest.mark.parametrize("mode", ["vol", "surf_esp", "pharm", "vol_color"])
def test_single_pair_methods_match_the_batch(dev, mode):
    """MoleculePair.align_with_* (the reference's per-pair API) runs the same kernels on one pair:
    same score as the batch call, numpy transform and score, aligned cloud returned."""
    print(f"case single pair mode={mode}", flush=True)
    refs, fits = _mols(2, 36)
    kw = dict(num_repeats=10, max_num_steps=50)
    bpairs = [MoleculePair(r, f, device=dev) for r, f in zip(refs, fits)]
    batch = MoleculePairBatch(bpairs)
    for i in range(2):
        p = MoleculePair(refs[i], fits[i], device=dev)
        if mode == "vol":
            out = p.align_with_vol(**kw)
            bs = batch.align_with_vol(backend="hip", **kw)[0]
            attr, cloud = "vol_noH", fits[i].atom_pos
        elif mode == "surf_esp":
            out = p.align_with_surf_esp(alpha=0.81, lam=0.3, **kw)
            bs = batch.align_with_surf_esp(alpha=0.81, lam=0.3, backend="hip", **kw)[0]
            attr, cloud = "surf_esp", fits[i].surf_pos
        elif mode == "pharm":
            out, vecs = p.align_with_pharm(**kw)
            bs = batch.align_with_pharm(backend="hip", **kw)[0]
            attr, cloud = "pharm", fits[i].pharm_ancs
            Rm = p.transform_pharm[:3, :3].astype(np.float64)
            np.testing.assert_allclose(vecs, fits[i].pharm_vecs @ Rm.T, atol=1e-5)
        else:
            out = p.align_with_vol_color(directionless=False, **kw)
            bs = batch.align_with_vol_color(backend="hip", directionless=False, **kw)[0]
            attr, cloud = "vol_color", fits[i].atom_pos
        T = getattr(p, "transform_" + attr)
        sc = getattr(p, "sim_aligned_" + attr)
        assert isinstance(T, np.ndarray) and T.shape == (4, 4)
        assert isinstance(sc, np.ndarray)
        # the autotuner may tile one pair differently from two: same kernels, FP32 trajectories
        # equal to rounding, so the best restarts agree to the parity bar
        assert abs(float(sc) - bs[i]) < 1e-3, (mode, i, float(sc), bs[i])
        np.testing.assert_allclose(out, cloud @ T[:3, :3].T.astype(np.float64) + T[:3, 3], atol=1e-5)


def test_score_with_current_pose_matches_oracle(dev):
    """score_with_surf / _esp / _pharm score the pair as it stands (_core.py:1359-1560)."""
    print("case score_with", flush=True)
    refs, fits = _mols(2, 37)
    ident = torch.tensor([[1.0, 0, 0, 0, 0, 0, 0]], dtype=torch.float64)
    for r, f in zip(refs, fits):
        p = MoleculePair(r, f, device=dev)
        s = p.score_with_surf(0.81)
        o = orc.Pair("shape", ref=r.surf_pos, fit=f.surf_pos).score_poses(ident)
        np.testing.assert_allclose(s, o.numpy(), atol=1e-9)
        s = p.score_with_esp(0.81, lam=0.3, use="torch")
        o = orc.Pair("esp", ref=r.surf_pos, fit=f.surf_pos, ref_esp=r.surf_esp, fit_esp=f.surf_esp,
                     lam=0.3 * LAM_SCALING).score_poses(ident)
        np.testing.assert_allclose(s, o.numpy(), atol=1e-9)
        kw = dict(ref_types=r.pharm_types, fit_types=f.pharm_types, ref_anc=r.pharm_ancs,
                  fit_anc=f.pharm_ancs, ref_vec=r.pharm_vecs, fit_vec=f.pharm_vecs)
        for sim, ext, only in (("tanimoto", False, False), ("tversky", True, False), ("tanimoto", True, True)):
            s = p.score_with_pharm(similarity=sim, extended_points=ext, only_extended=only)
            o = orc.Pair("pharm", similarity=sim, extended_points=ext, only_extended=only, **kw).score_poses(ident)
            np.testing.assert_allclose(s, o.numpy(), atol=1e-9)
