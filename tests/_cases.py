"""Shared synthetic cases: the same seeded inputs as a device PairBatch and as oracle Pairs."""
import numpy as np
import torch

from fast_shepherd_score_amd import _lib, synth
from fast_shepherd_score_amd.align import ScoreSpec
from fast_shepherd_score_amd.constants import LAM_SCALING
from fast_shepherd_score_amd.packing import CloudSet, PairBatch, PharmSet
from oracle import fss_oracle as orc

MODE_NAME = {_lib.MODE_SHAPE: "shape", _lib.MODE_SHAPE_ESP: "esp", _lib.MODE_PHARM: "pharm",
             _lib.MODE_SHAPE_PHARM: "shape_pharm"}


def make_case(mode, n_pairs, dev, seed=0, atoms=(20, 40), points=(40, 80), pharms=(4, 12),
              similarity="tanimoto", directionless=False, alpha=0.81, extended_points=False,
              only_extended=False, vec_scale=1.0):
    """Returns (PairBatch, ScoreSpec, [oracle.Pair]).  ``vec_scale`` multiplies the
    pharmacophore vectors (both sides see the same non-unit vectors)."""
    lam = 0.3 * LAM_SCALING
    spec = ScoreSpec(mode=mode, alpha=alpha, lam=lam, similarity=similarity, directionless=directionless,
                     extended_points=extended_points, only_extended=only_extended)
    if mode == _lib.MODE_SHAPE:
        ref, fit = synth.vol_pairs(n_pairs, atoms[0], atoms[1], seed=seed)
        pb = PairBatch(CloudSet.from_arrays(ref, device=dev), CloudSet.from_arrays(fit, device=dev))
        pairs = [orc.Pair("shape", ref=r, fit=f, alpha=alpha) for r, f in zip(ref, fit)]
    elif mode == _lib.MODE_SHAPE_ESP:
        r, re, f, fe = synth.surf_esp_pairs(n_pairs, atoms=(8, 16), points=points, seed=seed)
        pb = PairBatch(CloudSet.from_arrays(r, re, device=dev), CloudSet.from_arrays(f, fe, device=dev))
        pairs = [orc.Pair("esp", ref=a, fit=b, ref_esp=c, fit_esp=d, alpha=alpha, lam=lam)
                 for a, c, b, d in zip(r, re, f, fe)]
    else:
        d = synth.pharm_shape_pairs(n_pairs, atoms=atoms, pharms=pharms, seed=seed)
        for side in ("ref", "fit"):
            d[f"{side}_vectors"] = [np.asarray(v, np.float32) * np.float32(vec_scale)
                                    for v in d[f"{side}_vectors"]]
        ref_ph = PharmSet.from_arrays(d["ref_types"], d["ref_anchors"], d["ref_vectors"], device=dev)
        fit_ph = PharmSet.from_arrays(d["fit_types"], d["fit_anchors"], d["fit_vectors"], device=dev)
        ref = CloudSet.from_arrays(d["ref_atoms"], device=dev)
        fit = CloudSet.from_arrays(d["fit_atoms"], device=dev)
        pb = PairBatch(ref, fit, ref_ph, fit_ph)
        name = MODE_NAME[mode]
        pairs = [orc.Pair(name, ref=d["ref_atoms"][i], fit=d["fit_atoms"][i], ref_types=d["ref_types"][i],
                          fit_types=d["fit_types"][i], ref_anc=d["ref_anchors"][i], fit_anc=d["fit_anchors"][i],
                          ref_vec=d["ref_vectors"][i], fit_vec=d["fit_vectors"][i], alpha=alpha,
                          similarity=similarity, directionless=directionless,
                          extended_points=extended_points, only_extended=only_extended)
                 for i in range(n_pairs)]
    return pb, spec, pairs


def random_seeds(n_pairs, B, seed=0, t_scale=0.0):
    rng = np.random.default_rng(seed)
    q = synth.random_quaternions(rng, n_pairs * B).reshape(n_pairs, B, 4)
    t = (rng.normal(size=(n_pairs, B, 3)) * t_scale).astype(np.float32)
    return q, t


def to_dev(a, dev):
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)
