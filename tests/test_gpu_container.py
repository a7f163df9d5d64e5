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
