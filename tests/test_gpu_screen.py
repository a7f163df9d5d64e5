"""Screening service on the GPU: self-screen ranks the query's own conformer first, per-pair
scores match the FP64 oracle under the same seeds, and a query's hits are bit-identical alone
and in a mix of concurrent queries (continuous batching)."""
import threading
import time

import numpy as np
import pytest
import torch

from fast_shepherd_score_amd import synth
from fast_shepherd_score_amd.screen import ConformerLibrary, ScreenService, restart_table, screen, screen_many
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
    hits = screen(query, lib, "vol", top_k=40, num_repeats=B, max_num_steps=steps, do_center=False)
    table = restart_table(B).astype(np.float64)
    got = {h.id: h.score for h in hits}
    for i in (0, 7, 21, 39):
        ref, fit = mols[5].astype(np.float64), mols[i].astype(np.float64)
        t0 = ref.mean(0)[None] - np.einsum("bij,j->bi", rotmat(table), fit.mean(0))
        o = orc.optimize(orc.Pair("shape", ref=ref, fit=fit), table, t0, max_steps=steps)
        assert abs(got[i] - o["score"]) < 1e-3, (i, got[i], o["score"])


def test_surface_screen_scores_match_oracle(dev):
    """surf_esp screening runs the split-bf16 MFMA loop (two wavefronts per pair) by default; its
    per-entry scores agree with the FP64 oracle's alignment under the same seeds."""
    from fast_shepherd_score_amd.constants import LAM_SCALING
    print("case surf_esp screen vs oracle: 12 entries of 60-110 surface points, 6 x 50", flush=True)
    rng = np.random.default_rng(9)
    atoms = synth.random_walks(rng, rng.integers(12, 25, size=12))
    surf = [synth.surface_points(rng, a, int(n)) for a, n in zip(atoms, rng.integers(60, 111, size=12))]
    esp = [synth.esp_values(rng, len(x)) for x in surf]
    lib = ConformerLibrary.from_arrays(points=surf, esp=esp, device=dev)
    B, steps, lam = 6, 50, 0.3
    with ScreenService(lib, "surf_esp", num_repeats=B, max_num_steps=steps, top_k=12, lam=lam, alpha=0.81) as svc:
        assert svc.variant == 3 and svc.lanes_per_pose == 64 and svc.groups_per_wg == 2
        hits = svc.submit({"points": surf[3], "esp": esp[3]}).result(timeout=120).hits
    table = restart_table(B).astype(np.float64)
    got = {h.id: h.score for h in hits}
    assert len(got) == 12
    for i in (0, 3, 7, 11):
        ref, fit = surf[3].astype(np.float64), surf[i].astype(np.float64)
        t0 = ref.mean(0)[None] - np.einsum("bij,j->bi", rotmat(table), fit.mean(0))
        o = orc.optimize(orc.Pair("esp", ref=ref, fit=fit, ref_esp=esp[3], fit_esp=esp[i], lam=lam * LAM_SCALING),
                         table, t0, max_steps=steps)
        assert abs(got[i] - o["score"]) < 1e-3, (i, got[i], o["score"])


def test_empty_library_returns_no_hits(dev):
    """A query against a library with no conformers finishes (no launch) with zero hits, alone
    and beside a query on the same service."""
    print("case screen on an empty library", flush=True)
    empty = ConformerLibrary.from_arrays(points=[], device=dev)
    assert empty.n == 0
    query = {"points": synth.random_walks(np.random.default_rng(5), [12])[0]}
    assert screen(query, empty, "vol", top_k=5, num_repeats=4, max_num_steps=10) == []
    with ScreenService(empty, "vol", num_repeats=4, max_num_steps=10, top_k=5) as svc:
        res = [svc.submit(query).result(timeout=60) for _ in range(2)]
        assert svc.launches == 0
    for r in res:
        assert r.hits == [] and r.pairs == 0


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


def test_memory_budget_defers_admission_and_keeps_hits(dev):
    """A query budget of two maximum-size queries: queries of mixed sizes wait for pool room
    with slots free, are admitted by their own rows, and get the same hits as on a service that
    admits them all at once."""
    print("case screen with a two-query row budget", flush=True)
    mols, lib = _library(600, 11, dev)
    rng = np.random.default_rng(12)
    queries = [{"points": synth.random_walks(rng, [int(n)])[0]} for n in rng.integers(8, 33, size=8)]
    kw = dict(num_repeats=6, max_num_steps=30, top_k=10, launch_pairs=700, max_active=8,
              max_query_points=32)
    with ScreenService(lib, "vol", **kw) as svc:
        free = [f.result(timeout=300) for f in [svc.submit(q) for q in queries]]
        adm = svc.admission()
    assert adm["pool_rows"] == [8 * 32, 0, 0, 0] and adm["memory_waits"] == 0
    with ScreenService(lib, "vol", query_budget_bytes=2 * 32 * 16, **kw) as svc:
        tight = [f.result(timeout=300) for f in [svc.submit(q) for q in queries]]
        adm = svc.admission()
    assert adm["pool_rows"] == [64, 0, 0, 0]
    assert adm["memory_waits"] > 0 and 2 <= adm["max_admitted"] < 8, adm
    for a, b in zip(free, tight):
        assert [h.id for h in a.hits] == [h.id for h in b.hits]
        assert [h.score for h in a.hits] == [h.score for h in b.hits]
        assert b.pairs == 600


@pytest.mark.parametrize("lpp,ppw", [(4, 1), (2, 3)])
def test_native_service_topk_is_exact(dev, lpp, ppw):
    """The native service's hits (uploaded from host arrays through the C ABI, several launches,
    device top-k merges) are exactly the top-k of aligning every pair directly with the same seeds
    and tiling, ties (duplicated conformers) broken by the lower library index.  With several
    pairs per workgroup the service orders each launch by entry size; a pair's result does not
    depend on the pairs it shares a workgroup with, so the hits still match bit for bit."""
    import ctypes
    from fast_shepherd_score_amd import _lib
    from fast_shepherd_score_amd.align import OptimSpec, ScoreSpec, align, self_overlaps
    from fast_shepherd_score_amd.packing import CloudSet, PairBatch
    print(f"case native screen topk lpp={lpp} ppw={ppw}", flush=True)
    rng = np.random.default_rng(11)
    base = synth.random_walks(rng, rng.integers(10, 31, size=700))
    mols = base + base[:300]                   # entries 700..999 duplicate 0..299: exact ties
    n, B, steps, K = len(mols), 6, 40, 50
    counts = np.array([len(m) for m in mols], np.int32)
    rows = np.zeros((int(counts.sum()), 4), np.float32)
    rows[:, :3] = np.concatenate(mols).astype(np.float32)
    lib = _lib.load()
    spec = _lib.FssLibSpec()
    spec.device, spec.channels, spec.max_conformers, spec.max_points = 0, _lib.LIB_SHAPE, n, len(rows)
    h = ctypes.c_void_p()
    _lib.call("fss_lib_create", _lib.byref(spec), ctypes.byref(h))
    first = ctypes.c_int64()
    half = int(counts[:500].sum())             # two uploads from host memory
    _lib.call("fss_lib_upload", h, 500, rows[:half].ctypes.data, counts[:500].ctypes.data,
              None, None, None, None, ctypes.byref(first))
    _lib.call("fss_lib_upload", h, n - 500, rows[half:].ctypes.data, counts[500:].ctypes.data,
              None, None, None, None, ctypes.byref(first))
    assert first.value == 500
    table = restart_table(B)
    sp = ScoreSpec().struct()
    prm = _lib.FssScreenParams()
    prm.n_restarts, prm.max_steps, prm.lr, prm.beta1, prm.beta2, prm.eps = B, steps, 0.1, 0.9, 0.999, 1e-8
    prm.early_stop, prm.es_tol, prm.es_patience = 1, 1e-5, 10
    prm.top_k, prm.launch_pairs, prm.max_active, prm.max_query_points = K, 333, 4, 64
    prm.lanes_per_pose, prm.restarts_per_wg, prm.quat_table = lpp, B, table.ctypes.data
    prm.pairs_per_wg = ppw
    svc = ctypes.c_void_p()
    _lib.call("fss_screen_create", h, _lib.byref(sp), _lib.byref(prm), ctypes.byref(svc))
    qxyz = mols[17]
    qrows = np.zeros((len(qxyz), 4), np.float32)
    qrows[:, :3] = qxyz
    q = _lib.FssQuery()
    q.n_points, q.pts = len(qrows), qrows.ctypes.data
    ticket = ctypes.c_int64()
    _lib.call("fss_screen_submit", svc, ctypes.byref(q), ctypes.byref(ticket))
    score = np.zeros(K, np.float64)
    index = np.zeros(K, np.int64)
    tf = np.zeros((K, 16), np.float32)
    hits = _lib.FssHits()
    hits.capacity, hits.score, hits.index, hits.transform = K, score.ctypes.data, index.ctypes.data, tf.ctypes.data
    _lib.call("fss_screen_poll", svc, ticket.value, -1, ctypes.byref(hits))
    launches, pairs, pair_steps = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
    _lib.call("fss_screen_stats", svc, ctypes.byref(launches), ctypes.byref(pairs), ctypes.byref(pair_steps))
    _lib.call("fss_screen_destroy", svc)
    _lib.call("fss_lib_destroy", h)
    assert hits.n == K and launches.value == 4          # ceil(1000 / 333)
    # direct: every pair in one align() call with the same seeds and tiling
    libset = CloudSet.from_arrays(mols, device=dev)
    pb = PairBatch(CloudSet.from_arrays([qxyz], device=dev), libset,
                   ref_idx=np.zeros(n, np.int64), fit_idx=np.arange(n))
    q0 = torch.empty((n, B, 4), dtype=torch.float32, device=dev)
    t0 = torch.empty((n, B, 3), dtype=torch.float32, device=dev)
    bs = pb.struct()
    _lib.call("fss_seed_poses", _lib.byref(bs), 0, B, _lib.ptr(torch.from_numpy(table).to(dev)),
              _lib.ptr(q0), _lib.ptr(t0), int(torch.cuda.current_stream().cuda_stream))
    r = align(pb, ScoreSpec(), q0, t0, OptimSpec(max_steps=steps, lr=0.1, early_stop=True),
              lanes_per_pose=lpp, restarts_per_wg=B, pairs_per_wg=ppw, order="none",
              self_ov=self_overlaps(pb, ScoreSpec()))
    s = r.score.cpu().numpy()
    assert pairs.value == n and pair_steps.value == int(r.n_steps.sum())     # the executed steps
    order = np.lexsort((np.arange(n), -s))[:K]
    np.testing.assert_array_equal(index, order)
    np.testing.assert_array_equal(score, s[order])
    np.testing.assert_array_equal(tf.reshape(K, 4, 4), r.transform.cpu().numpy()[order])
    assert any(s[i] == s[i + 700] for i in range(300))   # ties were present


@pytest.mark.parametrize("mode", ["vol", "surf_esp", "pharm"])
def test_store_library_screens_like_in_memory(dev, tmp_path, mode):
    """A ProfileStore streamed into HBM by store.load_library screens exactly like an in-memory
    library built from the same (float16-rounded, pre-centred) arrays; screen() also takes the
    store path directly."""
    import os
    from fast_shepherd_score_amd.store import MoleculeProfile, ProfileStore, load_library
    print(f"case store screen mode={mode}", flush=True)
    rng = np.random.default_rng(21)
    S = 50                      # ALPHA(num_surf_points) is tabulated for 50-400 points
    atoms = synth.random_walks(rng, rng.integers(10, 26, size=260))
    path = os.path.join(tmp_path, "lib.fss")
    with ProfileStore.create(path, num_surf_points=S, modes=("vol", "surf_esp", "pharm"),
                             shard_size=100) as st:
        for i, a in enumerate(atoms):
            t, anc, vec = synth.pharmacophores(rng, a, int(rng.integers(3, 9)))
            st.add(MoleculeProfile(atom_pos=a, surf_pos=synth.surface_points(rng, a, S),
                                   surf_esp=synth.esp_values(rng, S), pharm_types=t.astype(np.int32),
                                   pharm_ancs=anc, pharm_vecs=vec, id=f"mol{i}"))
    store = ProfileStore.open(path)
    profs = [p for shard in store.iter_profiles() for p in shard]
    if mode == "vol":
        mem = ConformerLibrary.from_arrays(points=[p.atom_pos for p in profs], ids=None, device=dev)
        query = {"points": profs[33].atom_pos}
    elif mode == "surf_esp":
        mem = ConformerLibrary.from_arrays(points=[p.surf_pos for p in profs],
                                           esp=[p.surf_esp for p in profs], device=dev)
        query = {"points": profs[33].surf_pos, "esp": profs[33].surf_esp}
    else:
        mem = ConformerLibrary.from_arrays(pharm_types=[p.pharm_types for p in profs],
                                           pharm_anchors=[p.pharm_ancs for p in profs],
                                           pharm_vectors=[p.pharm_vecs for p in profs], device=dev)
        query = {"pharm_types": profs[33].pharm_types, "pharm_anchors": profs[33].pharm_ancs,
                 "pharm_vectors": profs[33].pharm_vecs}
    kw = dict(top_k=15, num_repeats=6, max_num_steps=40, launch_pairs=100)
    lib = load_library(store, mode)
    assert lib.n == 260 and lib.nbytes() > 0
    a = screen(query, lib, mode, **kw)
    b = screen(query, mem, mode, **kw)
    c = screen(query, path, mode, **kw)
    lib.close()
    assert [h.score for h in a] == [h.score for h in b] == [h.score for h in c]
    assert [h.id for h in a] == [f"mol{h.id}" for h in b] == [h.id for h in c]
    assert a[0].id == "mol33"


def test_one_library_serves_atom_and_surface_modes(dev):
    """A library with heavy atoms, pharmacophores and surface rows (points + ESP, no ShaEP atoms)
    screens vol, vol_color, pharm, surf and surf_esp with the same hits as single-purpose
    libraries holding each mode's rows in the shape channel."""
    print("case one library, five modes", flush=True)
    rng = np.random.default_rng(31)
    S = 50
    atoms = synth.random_walks(rng, rng.integers(10, 26, size=180))
    surf = [synth.surface_points(rng, a, S) for a in atoms]
    esp = [synth.esp_values(rng, S) for _ in atoms]
    ph = [synth.pharmacophores(rng, a, int(rng.integers(1, 17))) for a in atoms]
    pk = dict(pharm_types=[p[0].astype(np.int32) for p in ph], pharm_anchors=[p[1] for p in ph],
              pharm_vectors=[p[2] for p in ph])
    full = ConformerLibrary.from_arrays(points=atoms, surf_points=surf, surf_esp=esp, device=dev, **pk)
    assert full.has_surface_rows and full.num_surf_points == S and full.shaep is None
    single = {"vol": ConformerLibrary.from_arrays(points=atoms, device=dev),
              "vol_color": ConformerLibrary.from_arrays(points=atoms, device=dev, **pk),
              "pharm": ConformerLibrary.from_arrays(device=dev, **pk),
              "surf": ConformerLibrary.from_arrays(points=surf, device=dev),
              "surf_esp": ConformerLibrary.from_arrays(points=surf, esp=esp, device=dev)}
    k = 17
    qph = {"pharm_types": ph[k][0].astype(np.int32), "pharm_anchors": ph[k][1], "pharm_vectors": ph[k][2]}
    queries = {"vol": {"points": atoms[k]}, "vol_color": {"points": atoms[k], **qph}, "pharm": qph,
               "surf": {"points": surf[k], "atom_pos": atoms[k]},
               "surf_esp": {"points": surf[k], "esp": esp[k], "atom_pos": atoms[k]}}
    kw = dict(top_k=12, num_repeats=6, max_num_steps=30, launch_pairs=128)
    for mode, lib in single.items():
        a = screen(queries[mode], full, mode, **kw)
        b = screen(queries[mode], lib, mode, **kw)
        assert [h.id for h in a] == [h.id for h in b], mode
        assert [h.score for h in a] == [h.score for h in b], mode
        assert a[0].id == k, mode


def test_screen_terms_count_every_executed_row_pair(dev):
    """With early stop off every pair runs max_steps: the service's term counts are exactly
    steps x restarts x query rows x entry rows summed over the library, per score term; the
    screening library builder's entries are rigid copies of its base molecules."""
    from fast_shepherd_score_amd.screen import ConformerLibrary, ScreenService
    print("case screen term counts", flush=True)
    shape, pharm, surface = synth.screening_library(60, dev, base=25, seed=3)
    lib = ConformerLibrary(shape=shape, pharm=pharm, surface=surface)
    lib.num_surf_points = 200
    base = synth.screening_molecules(25, seed=3)
    assert list(shape.counts) == [len(base[i % 25]["atom_pos"]) for i in range(60)]
    s, c = shape.rng[37].tolist()
    a, b = shape.pts[s:s + c, :3].cpu().numpy(), base[37 % 25]["atom_pos"]
    d = lambda x: np.linalg.norm(x[:, None] - x[None], axis=-1)
    np.testing.assert_allclose(d(a), d(b), atol=1e-4)
    q = synth.screening_molecules(1, seed=99)[0]
    B, steps = 4, 20
    kw = dict(num_repeats=B, max_num_steps=steps, top_k=10, early_stop=False, launch_pairs=64)
    with ScreenService(lib, "vol_color", **kw) as svc:
        svc.submit({"points": q["atom_pos"], "pharm_types": q["pharm_types"], "pharm_anchors": q["pharm_ancs"],
                    "pharm_vectors": q["pharm_vecs"]}).result(timeout=120)
        st = svc.stats()
    assert st["pairs"] == 60 and st["pair_steps"] == 60 * steps
    assert st["shape_terms"] == steps * B * len(q["atom_pos"]) * int(shape.counts.sum())
    assert st["pharm_terms"] == steps * B * len(q["pharm_types"]) * int(pharm.counts.sum())
    with ScreenService(lib, "surf_esp", **kw) as svc:
        hits = svc.submit({"points": q["surf_pos"], "esp": q["surf_esp"], "atom_pos": q["atom_pos"]}).result(timeout=120)
        st = svc.stats()
    assert st["shape_terms"] == steps * B * 200 * 200 * 60 and st["pharm_terms"] == 0
    assert len(hits.hits) == 10


@pytest.mark.parametrize("only", [False, True])
def test_pharm_screen_extended_points_match_oracle(dev, only):
    """pharm screening with extended points (screen.py:866-867): per-pair best scores equal the
    FP64 oracle run from the screen's seeds (restart table, anchor centroids)."""
    print(f"case screen pharm extended only={only}", flush=True)
    rng = np.random.default_rng(8)
    atoms = synth.random_walks(rng, rng.integers(10, 25, size=30))
    f32 = lambda x: np.asarray(x, np.float32).astype(np.float64)   # the values the device holds
    ph = [(t, f32(a), f32(v)) for t, a, v in (synth.pharmacophores(rng, m, int(rng.integers(4, 10)))
                                               for m in atoms)]
    lib = ConformerLibrary.from_arrays(pharm_types=[p[0].astype(np.int32) for p in ph],
                                       pharm_anchors=[p[1] for p in ph], pharm_vectors=[p[2] for p in ph],
                                       device=dev)
    qi = 4
    query = {"pharm_types": ph[qi][0].astype(np.int32), "pharm_anchors": ph[qi][1], "pharm_vectors": ph[qi][2]}
    B, steps = 8, 60
    hits = screen(query, lib, "pharm", top_k=30, num_repeats=B, max_num_steps=steps, do_center=False,
                  directionless=False, extended_points=True, only_extended=only)
    table = restart_table(B).astype(np.float64)
    got = {h.id: h.score for h in hits}
    for i in (0, 4, 13, 29):
        ra, fa = ph[qi][1], ph[i][1]
        t0 = ra.mean(0)[None] - np.einsum("bij,j->bi", rotmat(table), fa.mean(0))
        pair = orc.Pair("pharm", ref_types=ph[qi][0], fit_types=ph[i][0], ref_anc=ra, fit_anc=fa,
                        ref_vec=ph[qi][2], fit_vec=ph[i][2], extended_points=True, only_extended=only)
        o = orc.optimize(pair, table, t0, max_steps=steps)
        assert abs(got[i] - o["score"]) < 1e-3, (i, got[i], o["score"])


def test_scores_out_object_queries_and_default_alpha(dev):
    """scores_out receives every entry's best score in library order (screen.py:1174-1176) and
    agrees with the hits; a container Molecule query equals its dict; surf_esp takes
    ALPHA(num_surf_points) when alpha is not given."""
    from fast_shepherd_score_amd.constants import ALPHA
    from fast_shepherd_score_amd.container import Molecule
    print("case screen scores_out / objects / alpha", flush=True)
    rng = np.random.default_rng(12)
    S = 50
    atoms = synth.random_walks(rng, rng.integers(10, 25, size=150))
    surf = [synth.surface_points(rng, a, S) for a in atoms]
    esp = [synth.esp_values(rng, S) for _ in atoms]
    lib = ConformerLibrary.from_arrays(points=surf, esp=esp, device=dev)
    qm = Molecule(atoms[9], surf_pos=surf[9], surf_esp=esp[9])
    kw = dict(top_k=20, num_repeats=6, max_num_steps=30, launch_pairs=64)
    out = np.full(150, -1.0)
    hits = screen(qm, lib, "surf_esp", scores_out=out, **kw)
    assert np.isfinite(out).all() and (out >= 0).all() and (out <= 1 + 1e-9).all()
    for h in hits:
        assert out[h.id] == h.score
    np.testing.assert_array_equal(np.sort(out)[::-1][:20], [h.score for h in hits])
    as_dict = screen({"points": surf[9], "esp": esp[9], "atom_pos": atoms[9]}, lib, "surf_esp", **kw)
    explicit = screen(qm, lib, "surf_esp", alpha=ALPHA(S), **kw)
    assert [h.score for h in as_dict] == [h.score for h in hits] == [h.score for h in explicit]
    assert [h.id for h in as_dict] == [h.id for h in hits]
    with pytest.raises(ValueError, match="explicit alpha"):
        screen(qm, lib, "vol_and_surf_esp", **kw)
    # a panel with per-query score arrays
    outs = [np.zeros(150), np.zeros(150)]
    res = screen_many([qm, Molecule(atoms[3], surf_pos=surf[3], surf_esp=esp[3])], lib, "surf_esp",
                      scores_out=outs, **kw)
    np.testing.assert_array_equal(outs[0], out)
    assert res[1][0].id == 3 and outs[1][3] == res[1][0].score
