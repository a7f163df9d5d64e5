"""Screening service on the GPU: self-screen ranks the query's own conformer first, per-pair
scores match the FP64 oracle under the same seeds, and a query's hits are bit-identical alone
and in a mix of concurrent queries (continuous batching)."""
import threading
import time

import numpy as np
import pytest
import torch

from fast_shepherd_score_amd import synth
from fast_shepherd_score_amd.screen import ConformerLibrary, ScreenService, restart_table, screen
from fast_shepherd_score_amd.seeds import rotmat
from oracle import fss_oracle as orc

pytestmark = pytest.mark.gpu


def _library(n, seed, dev):
    rng = np.random.default_rng(seed)
    mols = synth.random_walks(rng, rng.integers(10, 31, size=n))
    return mols, ConformerLibrary.from_arrays(points=mols, device=dev)


def test_self_screen_ranks_first(dev):
    mols, lib = _library(300, 0, dev)
    rng = np.random.default_rng(1)
    k = 137
    q = synth.random_quaternions(rng, 1).astype(np.float64)
    query = {"points": (mols[k] @ rotmat(q)[0].T + np.array([0.3, -0.2, 0.1])).astype(np.float32)}
    hits = screen(query, lib, "vol", top_k=5, num_repeats=24, max_num_steps=150)
    assert hits[0].id == k
    assert hits[0].score == pytest.approx(1.0, abs=1e-2)
    assert all(hits[i].score >= hits[i + 1].score for i in range(len(hits) - 1))


def test_screen_scores_match_oracle(dev):
    mols, lib = _library(40, 2, dev)
    query = {"points": mols[5]}
    B, steps = 8, 60
    hits = screen(query, lib, "vol", top_k=40, num_repeats=B, max_num_steps=steps)
    table = restart_table(B).astype(np.float64)
    got = {h.id: h.score for h in hits}
    for i in (0, 7, 21, 39):
        ref, fit = mols[5].astype(np.float64), mols[i].astype(np.float64)
        t0 = ref.mean(0)[None] - np.einsum("bij,j->bi", rotmat(table), fit.mean(0))
        o = orc.optimize(orc.Pair("shape", ref=ref, fit=fit), table, t0, max_steps=steps)
        assert abs(got[i] - o["score"]) < 1e-3, (i, got[i], o["score"])


def test_alone_equals_in_mix(dev):
    mols, lib = _library(2000, 3, dev)
    queries = [{"points": mols[i]} for i in (1, 50, 999)]
    kw = dict(num_repeats=6, max_num_steps=40, top_k=25, launch_pairs=1500)
    with ScreenService(lib, "vol", **kw) as svc:
        alone = svc.submit(queries[0]).result(timeout=300)
    with ScreenService(lib, "vol", **kw) as svc:
        futs = []
        for qq in queries:             # staggered arrivals: later queries join running launches
            futs.append(svc.submit(qq))
            time.sleep(0.02)
        mixed = [f.result(timeout=300) for f in futs]
        assert svc.launches > 3
    a = alone.hits
    b = mixed[0].hits
    assert [h.score for h in a] == [h.score for h in b]
    assert [h.id for h in a] == [h.id for h in b]
    for r in mixed:
        assert r.latency_ms > 0 and r.pairs == 2000
