"""Write the FP64 oracle's alignments of workload-sized pairs as golden fixtures.

The HIP optimiser is checked at the benchmark configurations' real sizes (200-400 surface points
with ESP, 50 restarts x 200 steps; uneven 5-60 atom / 2-20 pharmacophore pairs; 20-40 atom
pairs), where one oracle run takes minutes on a CPU.  This script runs the oracle
(oracle/fss_oracle.py, torch float64) once, from the same generators, seeds and restart draws
as bench.py's legs, on the pairs of each workload that span its size range, and stores the
inputs with the oracle's results, so a GPU test can compare against them without rerunning it.

Each fixture ``config_<name>.npz`` (numpy, no pickles) holds, per pair i of n:
  ref_xyz / fit_xyz (sum of counts, 3) float32 with ref_n / fit_n counts, ESP values ref_w /
  fit_w (esp), pharmacophore types / anchors / vectors and counts (shape_pharm), the seeds q0
  (n, B, 4) / t0 (n, B, 3) float32, and the oracle's best score, best restart, per-restart final
  scores (n, B), final poses se3 (n, B, 7) and executed steps, plus the run's constants.

    python tests/golden/make_config_vectors.py [--jobs 4] [--only c2]
"""
from __future__ import annotations

import argparse
import os
import sys
import time
from concurrent.futures import ProcessPoolExecutor
from functools import lru_cache

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
HERE = os.path.dirname(os.path.abspath(__file__))

from fast_shepherd_score_amd import synth                     # noqa: E402
from fast_shepherd_score_amd.constants import LAM_SCALING     # noqa: E402
from fast_shepherd_score_amd.seeds import initial_seeds, pad, rotmat   # noqa: E402


def com_seeds(ref, fit, q):
    """t0 = ref_com - R(q) fit_com for every restart (the benchmarks' random SO(3) seeds)."""
    rc = np.stack([np.asarray(x, np.float64).mean(0) for x in ref])
    fc = np.stack([np.asarray(x, np.float64).mean(0) for x in fit])
    return (rc[:, None, :] - np.einsum("nbij,nj->nbi", rotmat(q.astype(np.float64)), fc)).astype(np.float32)


def extremes(*sizes, n_random=0, seed=0):
    """Indices of the pairs holding the minimum and maximum of each size array, then n_random
    more drawn uniformly (duplicates removed, order kept)."""
    pick = []
    for s in sizes:
        pick += [int(np.argmin(s)), int(np.argmax(s))]
    rng = np.random.default_rng(seed)
    pick += [int(i) for i in rng.integers(0, len(sizes[0]), size=n_random)]
    return list(dict.fromkeys(pick))


# ------------------------------------------------------------------------------- workloads
@lru_cache(maxsize=None)
def c2():
    """Surface shape + ESP (bench.py default workload, seed 1): 4096 pairs generated, the
    largest / smallest N*M pairs and the pairs holding the extreme ref and fit sizes."""
    n_all, B = 4096, 50
    r, re, f, fe = synth.surf_esp_pairs(n_all, atoms=(15, 50), points=(200, 400), seed=1)
    rng = np.random.default_rng(1 + 1000)
    q = synth.random_quaternions(rng, n_all * B).reshape(n_all, B, 4)
    nr, nf = np.array([len(x) for x in r]), np.array([len(x) for x in f])
    idx = extremes(nr * nf)[:2] + extremes(nr, nf)
    idx = list(dict.fromkeys(idx))[:4]
    sel = lambda xs: [xs[i] for i in idx]
    return dict(kind="esp", idx=idx, ref=sel(r), fit=sel(f), ref_w=sel(re), fit_w=sel(fe),
                q0=q[idx], t0=com_seeds(sel(r), sel(f), q[idx]), lam=0.3 * LAM_SCALING,
                steps=200, early_stop=False)


def c2_es():
    """The largest C2 pair under the drop-in default: per-pair early stop on."""
    d = c2()
    k = int(np.argmax(np.array([len(r) * len(f) for r, f in zip(d["ref"], d["fit"])])))
    return dict(d, idx=d["idx"][k:k + 1], ref=d["ref"][k:k + 1], fit=d["fit"][k:k + 1],
                ref_w=d["ref_w"][k:k + 1], fit_w=d["fit_w"][k:k + 1], q0=d["q0"][k:k + 1],
                t0=d["t0"][k:k + 1], early_stop=True)


@lru_cache(maxsize=None)
def c1():
    """Volumetric shape (the C1 leg's 64 pairs, seed 3): eight pairs spanning the sizes."""
    n_all, B = 64, 50
    r, f = synth.vol_pairs(n_all, 20, 40, seed=3)
    rng = np.random.default_rng(3 + 1000)
    q = synth.random_quaternions(rng, n_all * B).reshape(n_all, B, 4)
    nr, nf = np.array([len(x) for x in r]), np.array([len(x) for x in f])
    idx = extremes(nr, nf, nr * nf, n_random=4, seed=3)[:8]
    sel = lambda xs: [xs[i] for i in idx]
    return dict(kind="shape", idx=idx, ref=sel(r), fit=sel(f), q0=q[idx],
                t0=com_seeds(sel(r), sel(f), q[idx]), steps=200, early_stop=False)


@lru_cache(maxsize=None)
def c3():
    """Shape + directional pharmacophores (the C3 leg's 16384 pairs, seed 4; restarts from rng
    44): the pairs holding the extreme atom and pharmacophore counts of either side."""
    n_all, B = 16384, 50
    d = synth.pharm_shape_pairs(n_all, seed=4)
    rng = np.random.default_rng(44)
    q = synth.random_quaternions(rng, n_all * B).reshape(n_all, B, 4)
    na_r = np.array([len(x) for x in d["ref_atoms"]])
    na_f = np.array([len(x) for x in d["fit_atoms"]])
    np_r = np.array([len(x) for x in d["ref_types"]])
    np_f = np.array([len(x) for x in d["fit_types"]])
    idx = extremes(na_r, na_f, np_r, np_f)
    sel = lambda k: [d[k][i] for i in idx]
    return dict(kind="shape_pharm", idx=idx, ref=sel("ref_atoms"), fit=sel("fit_atoms"),
                ref_types=sel("ref_types"), fit_types=sel("fit_types"),
                ref_anc=sel("ref_anchors"), fit_anc=sel("fit_anchors"),
                ref_vec=sel("ref_vectors"), fit_vec=sel("fit_vectors"), q0=q[idx],
                t0=com_seeds(sel("ref_atoms"), sel("fit_atoms"), q[idx]), steps=200,
                early_stop=False, directionless=False)


@lru_cache(maxsize=None)
def vol_color_default():
    """MoleculePairBatch.align_with_vol_color() with its defaults: directionless, colour weight
    0.5, 16 restarts from the reference's seed rule (identity, principal axes, Fibonacci), 40
    steps, per-pair early stop.  Six pairs of the C3 generator (seed 9)."""
    d = synth.pharm_shape_pairs(6, seed=9)
    (R, rn), (F, fn) = pad(d["ref_atoms"]), pad(d["fit_atoms"])
    q0, t0 = initial_seeds(R, rn, F, fn, 16)
    k = lambda key: d[key]
    return dict(kind="shape_pharm", idx=list(range(6)), ref=k("ref_atoms"), fit=k("fit_atoms"),
                ref_types=k("ref_types"), fit_types=k("fit_types"), ref_anc=k("ref_anchors"),
                fit_anc=k("fit_anchors"), ref_vec=k("ref_vectors"), fit_vec=k("fit_vectors"),
                q0=q0, t0=t0, steps=40, early_stop=True, directionless=True)


def _large_seeds(ref, fit, B, seed):
    rng = np.random.default_rng(seed)
    q = synth.random_quaternions(rng, len(ref) * B).reshape(len(ref), B, 4)
    return q, com_seeds(ref, fit, q)


@lru_cache(maxsize=None)
def large_esp():
    """Surface shape + ESP clouds beyond one workgroup's LDS (the streamed path; seed 21): a
    12000-point surface of a 600-atom random walk against a 300-point ligand surface, and two
    6000-point surfaces of 300-atom walks.  2 restarts x 30 steps."""
    rng = np.random.default_rng(21)
    atoms = synth.random_walks(rng, [600, 40, 300, 300])
    pts = [synth.surface_points(rng, a, n) for a, n in zip(atoms, [12000, 300, 6000, 6000])]
    esp = [synth.esp_values(rng, len(x)) for x in pts]
    ref, fit = [pts[0], pts[2]], [pts[1], pts[3]]
    q0, t0 = _large_seeds(ref, fit, 2, 22)
    return dict(kind="esp", idx=[0, 1], ref=ref, fit=fit, ref_w=[esp[0], esp[2]], fit_w=[esp[1], esp[3]],
                q0=q0, t0=t0, lam=0.3 * LAM_SCALING, steps=30, early_stop=False)


@lru_cache(maxsize=None)
def large_vol():
    """Volumetric shape beyond one workgroup's LDS (seed 23): a 9000-atom random walk against a
    2000-atom one (11000 rows of 16 B).  2 restarts x 30 steps, per-pair early stop."""
    rng = np.random.default_rng(23)
    ref, fit = synth.random_walks(rng, [9000, 2000])
    q0, t0 = _large_seeds([ref], [fit], 2, 24)
    return dict(kind="shape", idx=[0], ref=[ref], fit=[fit], q0=q0, t0=t0, steps=30, early_stop=True)


CONFIGS = {"c1": c1, "c2": c2, "c2_es": c2_es, "c3": c3, "vol_color_default": vol_color_default,
           "large_esp": large_esp, "large_vol": large_vol}


# ------------------------------------------------------------------------------- oracle runs
def oracle_pair(d, i):
    from oracle import fss_oracle as orc
    if d["kind"] == "esp":
        p = orc.Pair("esp", ref=d["ref"][i], fit=d["fit"][i], ref_esp=d["ref_w"][i], fit_esp=d["fit_w"][i],
                     lam=float(d["lam"]))
    elif d["kind"] == "shape":
        p = orc.Pair("shape", ref=d["ref"][i], fit=d["fit"][i])
    else:
        p = orc.Pair("shape_pharm", ref=d["ref"][i], fit=d["fit"][i], ref_types=d["ref_types"][i],
                     fit_types=d["fit_types"][i], ref_anc=d["ref_anc"][i], fit_anc=d["fit_anc"][i],
                     ref_vec=d["ref_vec"][i], fit_vec=d["fit_vec"][i], directionless=bool(d["directionless"]))
    return p


def pair_data(d, i):
    """The arrays of pair i (what one worker needs)."""
    one = {k: (v[i] if isinstance(v, list) and k != "idx" else v) for k, v in d.items()}
    one["q0"], one["t0"] = d["q0"][i], d["t0"][i]
    return one


def run_one(args):
    name, i, threads, one = args
    import torch
    torch.set_num_threads(threads)
    from oracle import fss_oracle as orc
    wrap = {k: ([v] if k in LISTS else v) for k, v in one.items()}
    p = oracle_pair(wrap, 0)
    t0 = time.time()
    o = orc.optimize(p, one["q0"], one["t0"], lr=0.1, max_steps=one["steps"], early_stop=one["early_stop"])
    return name, i, o, time.time() - t0


LISTS = ("ref", "fit", "ref_w", "fit_w", "ref_types", "fit_types", "ref_anc", "fit_anc", "ref_vec", "fit_vec")


def flat(xs, dtype, width=None):
    xs = [np.asarray(x, dtype) for x in xs]
    cat = np.concatenate(xs) if xs else np.zeros((0,), dtype)
    return cat.reshape(-1, width) if width else cat


def save(name, d, outs):
    n = len(d["idx"])
    arr = dict(idx=np.array(d["idx"], np.int64), kind=np.array(d["kind"]), steps=d["steps"],
               early_stop=d["early_stop"], lr=0.1, restarts=d["q0"].shape[1],
               ref_xyz=flat(d["ref"], np.float32, 3), fit_xyz=flat(d["fit"], np.float32, 3),
               ref_n=np.array([len(x) for x in d["ref"]], np.int64),
               fit_n=np.array([len(x) for x in d["fit"]], np.int64),
               q0=np.asarray(d["q0"], np.float32), t0=np.asarray(d["t0"], np.float32),
               best=np.array([outs[i]["score"] for i in range(n)]),
               best_idx=np.array([outs[i]["best_idx"] for i in range(n)], np.int64),
               scores=np.stack([outs[i]["scores"] for i in range(n)]),
               se3=np.stack([outs[i]["se3"] for i in range(n)]),
               n_steps=np.array([outs[i]["n_steps"] for i in range(n)], np.int64))
    if d["kind"] == "esp":
        arr.update(ref_w=flat(d["ref_w"], np.float32), fit_w=flat(d["fit_w"], np.float32), lam=d["lam"])
    if d["kind"] == "shape_pharm":
        arr.update(ref_types=flat(d["ref_types"], np.int64), fit_types=flat(d["fit_types"], np.int64),
                   ref_anc=flat(d["ref_anc"], np.float32, 3), fit_anc=flat(d["fit_anc"], np.float32, 3),
                   ref_vec=flat(d["ref_vec"], np.float32, 3), fit_vec=flat(d["fit_vec"], np.float32, 3),
                   ref_pn=np.array([len(x) for x in d["ref_types"]], np.int64),
                   fit_pn=np.array([len(x) for x in d["fit_types"]], np.int64),
                   directionless=d["directionless"])
    path = os.path.join(HERE, f"config_{name}.npz")
    np.savez_compressed(path, **arr)
    print("wrote", path, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=4)
    ap.add_argument("--only", nargs="*", default=sorted(CONFIGS))
    ap.add_argument("--threads", type=int, default=0, help="torch threads per job (0: cpus / jobs)")
    a = ap.parse_args()
    threads = a.threads or max(1, (os.cpu_count() or 8) // a.jobs)
    data = {name: CONFIGS[name]() for name in a.only}
    work = [(name, i, threads, pair_data(data[name], i)) for name in a.only
            for i in range(len(data[name]["idx"]))]
    outs = {name: {} for name in a.only}
    with ProcessPoolExecutor(max_workers=a.jobs) as ex:
        for name, i, o, dt in ex.map(run_one, work):
            outs[name][i] = o
            print(f"{name} pair {i}: best {o['score']:.6f} after {o['n_steps']} steps ({dt:.1f} s)", flush=True)
    for name in a.only:
        save(name, data[name], outs[name])


if __name__ == "__main__":
    main()
