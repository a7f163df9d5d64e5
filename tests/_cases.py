"""Shared synthetic cases: the same seeded inputs as a device PairBatch and as oracle Pairs."""
import os

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


# ---------------------------------------------------------------------------------------------
# Early-stop pins.  In each case every restart keeps one exact loss at every step and an exactly
# zero gradient (Adam then leaves the pose unchanged), so the rule alone fixes the step count:
# stop after the first step at which |L_k - L_(k-1)| <= 1e-5 has held more than 10 times running,
# with L_(-1) = 1.0 (analytic optimisers, _torch_analytical.py:110, 132-139) or +inf (ShaEP,
# alignment/_torch.py:922, 967-974).
#   * "optimum": a one-atom cloud aligned with its copy at the identity: L = 0 every step; the
#     first step resets the counter, steps 2..12 count 1..11, so 12 steps run.
#   * "apart": clouds 150 A apart: L = 1 every step; from L_(-1) = 1 step 1 already counts, so 11
#     steps run.  ">=" in place of ">" would give 11 and 10.
#   * "shaep_apart": the ShaEP score of molecules 150 A apart with surface ESP 10 and zero charges
#     (esp terms exp(-100 / lam) vanish, and zero charges make the ESP gradient exactly 0): L = 1,
#     and from L_(-1) = +inf the first step resets, so 12 steps run (11 from 1.0).
ES_EXPECTED = {"optimum": 12, "apart": 11, "shaep_apart": 12}
ES_SHIFT = np.array([150.0, 0.0, 0.0])


def es_shape_clouds(name):
    """(ref, fit) float64 clouds of the shape cases."""
    if name == "optimum":
        return np.zeros((1, 3)), np.zeros((1, 3))
    ref = synth.random_walks(np.random.default_rng(3), [6])[0].astype(np.float64)
    fit = synth.random_walks(np.random.default_rng(4), [5])[0].astype(np.float64) + ES_SHIFT
    return ref, fit


def es_shaep_molecules():
    """(ref, fit) ShaEP molecule dicts of the "shaep_apart" case."""
    rng = np.random.default_rng(6)
    a, b = synth.shaep_molecule(rng, 6, 30), synth.shaep_molecule(rng, 5, 24)
    for m in (a, b):
        m["surf_esp"] = np.full(len(m["surf_pos"]), 10.0, np.float32)
        m["partial_charges"] = np.zeros(len(m["centers_w_H"]), np.float32)
    for k in ("atom_pos", "surf_pos", "centers_w_H"):
        b[k] = (b[k].astype(np.float64) + ES_SHIFT).astype(np.float32)
    return a, b


def es_seeds(name, B):
    """Identity seeds for "optimum" (the pose must stay the exact optimum), random rotations
    about the origin with zero translation otherwise."""
    if name == "optimum":
        q = np.tile(np.array([1.0, 0.0, 0.0, 0.0], np.float32), (1, B, 1))
    else:
        q = synth.random_quaternions(np.random.default_rng(7), B).reshape(1, B, 4).astype(np.float32)
    return q, np.zeros((1, B, 3), np.float32)


# ---------------------------------------------------------------------------------------------
# Config-sized oracle fixtures (tests/golden/make_config_vectors.py)
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load_config(name):
    """A config_<name>.npz fixture as a dict, ragged arrays split per pair."""
    z = np.load(os.path.join(GOLD, f"config_{name}.npz"))
    split = lambda a, n: np.split(z[a], np.cumsum(z[n])[:-1])
    d = {k: z[k] for k in z.files}
    d["kind"] = str(z["kind"])
    d["ref"], d["fit"] = split("ref_xyz", "ref_n"), split("fit_xyz", "fit_n")
    if d["kind"] == "esp":
        d["ref_w"], d["fit_w"] = split("ref_w", "ref_n"), split("fit_w", "fit_n")
    if d["kind"] == "shape_pharm":
        for side in ("ref", "fit"):
            for k in ("types", "anc", "vec"):
                d[f"{side}_{k}"] = split(f"{side}_{k}", f"{side}_pn")
    return d
