"""vol_and_surf_esp (ShaEP) screening on the GPU: the native service's hits are exactly the top-k
of aligning every pair directly with the same seeds and tiling, its scores match the FP64 oracle
under the same seeds, and a with-H ProfileStore screens like the in-memory library (needs an
MI355X)."""
import os

import numpy as np
import pytest
import torch

from fast_shepherd_score_amd import _lib, synth
from fast_shepherd_score_amd.align import OptimSpec, ScoreSpec, align, self_overlaps
from fast_shepherd_score_amd.constants import LAM_SCALING
from fast_shepherd_score_amd.packing import CloudSet, PairBatch, ShaepSet
from fast_shepherd_score_amd.screen import ConformerLibrary, ScreenService, restart_table, screen
from fast_shepherd_score_amd.seeds import rotmat
from oracle import fss_oracle as orc

pytestmark = pytest.mark.gpu
MODE = "vol_and_surf_esp"


def _mols(n, seed):
    rng = np.random.default_rng(seed)
    return [synth.shaep_molecule(rng, int(rng.integers(8, 16)), int(rng.integers(40, 80))) for _ in range(n)]


def _library(mols, dev):
    col = lambda k: [m[k] for m in mols]
    return ConformerLibrary.from_arrays(points=col("atom_pos"), surf_points=col("surf_pos"),
                                        surf_esp=col("surf_esp"), atoms_w_H=col("centers_w_H"),
                                        charges=col("partial_charges"), radii=col("radii"), device=dev)


def _query(m):
    return {"points": m["atom_pos"], "surf_points": m["surf_pos"], "surf_esp": m["surf_esp"],
            "atoms_w_H": m["centers_w_H"], "charges": m["partial_charges"], "radii": m["radii"]}


@pytest.mark.parametrize("alpha", [0.81, 1.0])
def test_shaep_service_topk_is_exact(dev, alpha):
    """Three launches and device top-k merges against one align() over every pair (heavy-atom
    shape at alpha 0.81, surface shape otherwise; seeds from the surface centroids)."""
    print(f"case shaep screen topk alpha={alpha}", flush=True)
    mols = _mols(90, 31)
    n, B, steps, K = len(mols), 4, 20, 30
    lib = _library(mols, dev)
    qm = mols[7]
    with ScreenService(lib, MODE, num_repeats=B, max_num_steps=steps, top_k=K, launch_pairs=40,
                       alpha=alpha) as svc:
        res = svc.submit(_query(qm)).result(timeout=300)
        assert svc.launches == 3
    hits = res.hits
    assert len(hits) == K
    se_q = ShaepSet.from_arrays([qm["surf_pos"]], [qm["surf_esp"]], [qm["centers_w_H"]],
                                [qm["partial_charges"]], [qm["radii"]], device=dev)
    if alpha == 0.81:
        ref, fit = CloudSet.from_arrays([qm["atom_pos"]], device=dev), lib.shape
    else:
        ref, fit = se_q.srf, lib.shaep.srf
    pb = PairBatch(ref, fit, ref_idx=np.zeros(n, np.int64), fit_idx=np.arange(n),
                   ref_se=se_q, fit_se=lib.shaep)
    spec = ScoreSpec(mode=_lib.MODE_ESP_COMBO, alpha=alpha, lam=LAM_SCALING * 0.001)
    table = torch.from_numpy(restart_table(B)).to(dev)
    q0 = torch.empty((n, B, 4), dtype=torch.float32, device=dev)
    t0 = torch.empty((n, B, 3), dtype=torch.float32, device=dev)
    bs = pb.struct()
    _lib.call("fss_seed_poses", _lib.byref(bs), _lib.SEED_SURF, B, _lib.ptr(table), _lib.ptr(q0),
              _lib.ptr(t0), int(torch.cuda.current_stream().cuda_stream))
    r = align(pb, spec, q0, t0, OptimSpec(max_steps=steps, lr=0.1, early_stop=True),
              lanes_per_pose=32, restarts_per_wg=B, order="none", self_ov=self_overlaps(pb, spec))
    s = r.score.cpu().numpy()
    order = np.lexsort((np.arange(n), -s))[:K]
    assert [h.id for h in hits] == order.tolist()
    assert [h.score for h in hits] == s[order].tolist()
    np.testing.assert_array_equal(np.stack([h.transform for h in hits]), r.transform.cpu().numpy()[order])
    assert hits[0].id == 7


def test_shaep_screen_scores_match_oracle(dev):
    """Per-hit scores against the FP64 oracle optimiser run from the service's seeds (the
    restart table, t0 = ref surface centroid - R fit surface centroid)."""
    print("case shaep screen oracle", flush=True)
    mols = _mols(6, 32)
    B, steps = 4, 30
    lib = _library(mols, dev)
    qm = mols[2]
    # do_center=False: the oracle runs in the query's own frame, from the same seeds
    hits = screen(_query(qm), lib, MODE, top_k=6, num_repeats=B, max_num_steps=steps, early_stop=False,
                  alpha=0.81, do_center=False)
    got = {h.id: h.score for h in hits}
    table = restart_table(B).astype(np.float64)
    for i in (0, 2, 5):
        m = mols[i]
        pair = orc.ComboPair(ref_centers=qm["atom_pos"], fit_centers=m["atom_pos"], ref_surf=qm["surf_pos"],
                             fit_surf=m["surf_pos"], ref_surf_esp=qm["surf_esp"], fit_surf_esp=m["surf_esp"],
                             ref_atoms=qm["centers_w_H"], fit_atoms=m["centers_w_H"],
                             ref_charges=qm["partial_charges"], fit_charges=m["partial_charges"],
                             ref_radii=qm["radii"], fit_radii=m["radii"], alpha=0.81)
        rc = np.asarray(qm["surf_pos"], np.float64).mean(0)
        fc = np.asarray(m["surf_pos"], np.float64).mean(0)
        t0 = rc[None] - np.einsum("bij,j->bi", rotmat(table), fc)
        o = orc.optimize(pair, table, t0, lr=0.1, max_steps=steps, early_stop=False)
        assert abs(got[i] - o["score"]) < 1e-3, (i, got[i], o["score"])


def test_shaep_store_screens_like_in_memory(dev, tmp_path):
    """A with-H ProfileStore (float32, not pre-centred) loaded by store.load_library screens
    exactly like the in-memory library of the same arrays; screen() also takes the path."""
    from fast_shepherd_score_amd.store import MoleculeProfile, ProfileStore, load_library
    print("case shaep store screen", flush=True)
    rng = np.random.default_rng(33)
    S = 48
    mols = [synth.shaep_molecule(rng, int(rng.integers(8, 16)), S) for _ in range(70)]
    path = os.path.join(tmp_path, "lib.fss")
    with ProfileStore.create(path, num_surf_points=S, modes=(MODE,), dtype="float32",
                             pre_centered=False, shard_size=30) as st:
        for i, m in enumerate(mols):
            nh = len(m["atom_pos"])
            st.add(MoleculeProfile(atom_pos=m["atom_pos"], surf_pos=m["surf_pos"], surf_esp=m["surf_esp"],
                                   partial_charges=m["partial_charges"], radii=m["radii"],
                                   centers_w_H=m["centers_w_H"], nonH_atoms_idx=np.arange(nh), id=i))
    kw = dict(top_k=12, num_repeats=4, max_num_steps=20, launch_pairs=25, alpha=0.81, do_center=False)
    lib = load_library(path, MODE)
    assert lib.n == 70 and lib.nbytes() > 0
    a = screen(_query(mols[11]), lib, MODE, **kw)
    lib.close()
    b = screen(_query(mols[11]), _library(mols, dev), MODE, **kw)
    c = screen(_query(mols[11]), path, MODE, **kw)
    assert [h.score for h in a] == [h.score for h in b] == [h.score for h in c]
    assert [h.id for h in a] == [h.id for h in b] == [h.id for h in c]
    assert a[0].id == 11
    # default centring (query, and this store on load) moves the frames; the scores agree to the
    # best-restart bar (FP32 trajectories in another frame; the ESP mask is discontinuous)
    kw.pop("do_center")
    d = screen(_query(mols[11]), path, MODE, **kw)
    np.testing.assert_allclose([h.score for h in d], [h.score for h in a], atol=1e-3)
