"""Host-side logic (no GPU): seeds, packing, tiling model, synthetic generators, container API."""
import numpy as np
import pytest
import torch

from fast_shepherd_score_amd import _lib, synth
from fast_shepherd_score_amd.align import (ScoreSpec, candidate_tilings, check_tiling, default_tiling,
                                           tiling_efficiency)
from fast_shepherd_score_amd.constants import ALPHA, LAM_SCALING, P_TYPES, pharm_tables
from fast_shepherd_score_amd.container import MoleculePairBatch, resolve_backend
from fast_shepherd_score_amd.modes import HIP_MODES, MODE_ATTRS, MODE_SEEDS, MODE_STEPS, canonical
from fast_shepherd_score_amd.packing import CloudSet, PairBatch, PharmSet
from fast_shepherd_score_amd.seeds import _principal_axes, initial_seeds, pad, quats_from_fibo, rotmat


# ------------------------------------------------------------------------------- seeds
def test_seed_zero_is_identity_with_com_alignment():
    ref, fit = synth.vol_pairs(3, 10, 20, seed=1)
    fit = [f + np.array([1.0, -2.0, 0.5], np.float32) for f in fit]
    (R, rn), (F, fn) = pad(ref), pad(fit)
    q, t = initial_seeds(R, rn, F, fn, 50)
    assert q.shape == (3, 50, 4) and t.shape == (3, 50, 3)
    np.testing.assert_allclose(q[:, 0], np.tile([1, 0, 0, 0], (3, 1)))
    for i in range(3):
        np.testing.assert_allclose(t[i, 0], ref[i].mean(0) - fit[i].mean(0), atol=1e-5)
        # every seed puts the fit centroid on the ref centroid
        moved = np.einsum("sij,j->si", rotmat(q[i].astype(np.float64)), fit[i].mean(0)) + t[i]
        np.testing.assert_allclose(moved, np.tile(ref[i].mean(0), (50, 1)), atol=1e-4)


def test_pca_seed_aligns_principal_axes():
    """Seed 1 rotates the fit's principal axes onto the ref's (up to the axis sign flips)."""
    rng = np.random.default_rng(2)
    ref = rng.normal(size=(40, 3)) * np.array([4.0, 2.0, 0.7])
    q = synth.random_quaternions(rng, 1).astype(np.float64)
    fit = (ref - ref.mean(0)) @ rotmat(q)[0].T
    (R, rn), (F, fn) = pad([ref]), pad([fit])
    qs, _ = initial_seeds(R, rn, F, fn, 5)
    ax_ref = _principal_axes(R, rn)[0]
    for s in range(1, 5):
        Rs = rotmat(qs[0, s].astype(np.float64))
        moved = fit @ Rs.T
        ax_fit = _principal_axes(moved[None], fn)[0]
        for k in range(3):
            assert abs(abs(ax_fit[k] @ ax_ref[k]) - 1.0) < 1e-4


def test_seed_count_rules():
    ref, fit = synth.vol_pairs(1, 8, 8, seed=3)
    (R, rn), (F, fn) = pad(ref), pad(fit)
    assert initial_seeds(R, rn, F, fn, 1)[0].shape[1] == 1
    assert initial_seeds(R, rn, F, fn, 3)[0].shape[1] == 5       # 1 < S < 5 is raised to 5
    q, _ = initial_seeds(R, rn, F, fn, 12)
    np.testing.assert_allclose(q[0, 5:], quats_from_fibo(7), atol=1e-7)


# ------------------------------------------------------------------------------- packing
def test_cloudset_and_pairbatch_ranges():
    xyz = [np.ones((3, 3)), np.zeros((0, 3)), 2 * np.ones((5, 3))]
    cs = CloudSet.from_arrays(xyz, [np.arange(3), None, np.arange(5)], device="cpu")
    assert cs.rng.tolist() == [[0, 3], [3, 0], [3, 5]]
    assert cs.pts.shape == (8, 4) and cs.pts[4, 3] == 1.0
    pb = PairBatch(cs, cs, ref_idx=[2, 0], fit_idx=[0, 2])
    assert pb.n_pairs == 2 and pb.ref_rng.tolist() == [[3, 5], [0, 3]]
    assert pb.cost(_lib.MODE_SHAPE).tolist() == [15.0, 15.0]
    s = pb.struct()
    assert s.max_ref == 5 and s.max_fit == 5 and s.n_pairs == 2
    ph = PharmSet.from_arrays([np.array([0, 2]), np.array([7])], [np.zeros((2, 3)), np.ones((1, 3))],
                              [np.eye(3)[:2], np.zeros((1, 3))], device="cpu")
    assert ph.typ.tolist() == [0, 2, 7] and ph.rng.tolist() == [[0, 2], [2, 1]]


def test_padded_packing_equals_list_packing():
    rng = np.random.default_rng(0)
    counts = np.array([3, 1, 4])
    P = rng.normal(size=(3, 4, 3)).astype(np.float32)
    W = rng.normal(size=(3, 4)).astype(np.float32)
    a = CloudSet.from_padded(P, counts, W, device="cpu")
    b = CloudSet.from_arrays([P[i, :counts[i]] for i in range(3)], [W[i, :counts[i]] for i in range(3)], device="cpu")
    assert torch.equal(a.pts, b.pts) and torch.equal(a.rng, b.rng)


# ------------------------------------------------------------------------------- tiling
def test_tiling_choices_are_legal():
    rng = np.random.default_rng(0)
    for n, lo, hi, B, es in [(64, 20, 40, 50, True), (4096, 200, 400, 50, False), (3, 1, 5, 1, True),
                             (1000, 10, 50, 10, True), (10, 2, 20, 32, False)]:
        fit = [np.zeros((int(k), 3)) for k in rng.integers(lo, hi + 1, n)]
        ref = [np.zeros((int(k), 3)) for k in rng.integers(lo, hi + 1, n)]
        pb = PairBatch(CloudSet.from_arrays(ref, device="cpu"), CloudSet.from_arrays(fit, device="cpu"))
        lpp, rpw = default_tiling(pb, _lib.MODE_SHAPE, B, es)
        check_tiling(B, lpp, rpw, es)
        assert 0 < tiling_efficiency(pb.fit_counts, B, lpp, rpw) <= 1
    with pytest.raises(ValueError):
        check_tiling(50, 3, 50, True)
    with pytest.raises(ValueError):
        check_tiling(50, 16, 10, True)
    assert all(l * r <= 1024 for l, r in candidate_tilings(50, False))


# ------------------------------------------------------------------------------- synth
def test_random_walk_geometry():
    rng = np.random.default_rng(0)
    mols = synth.random_walks(rng, [30, 5, 1])
    for m in mols:
        np.testing.assert_allclose(m.mean(0), 0, atol=1e-5)
        if len(m) > 1:
            steps = np.linalg.norm(np.diff(m, axis=0), axis=1)
            np.testing.assert_allclose(steps, 1.5, atol=1e-4)
            d = np.linalg.norm(m[:, None] - m[None], axis=-1) + 10 * np.eye(len(m))
            assert d.min() >= 1.2 - 1e-4


def test_surface_and_esp_and_pharm():
    rng = np.random.default_rng(1)
    atoms = synth.random_walks(rng, [12])[0].astype(np.float64)
    pts = synth.surface_points(rng, atoms, 150).astype(np.float64)
    d = np.linalg.norm(pts[:, None] - atoms[None], axis=-1)
    r = 1.2 * 1.7
    assert (d.min(1) >= r - 1e-4).all() and np.isclose(d.min(1), r, atol=1e-4).all()
    e = synth.esp_values(rng, 1000)
    assert e.min() >= -0.3 and e.max() <= 0.3 and abs(e.std() - 0.1) < 0.02
    t, a, v = synth.pharmacophores(rng, atoms.astype(np.float32), 50)
    assert t.min() >= 0 and t.max() <= 7
    directional = np.isin(t, synth.DIRECTIONAL_TYPE_IDS)
    np.testing.assert_allclose(np.linalg.norm(v[directional], axis=1), 1, atol=1e-5)
    assert (v[~directional] == 0).all()


# ------------------------------------------------------------------------------- constants / API
def test_constants_and_tables():
    assert P_TYPES[-1] == "Dummy" and len(P_TYPES) == 9
    assert LAM_SCALING == pytest.approx(207.3535, rel=1e-6)
    a, k, c = pharm_tables()
    assert c == [1, 1, 2, 0, 1, 0, 0, 0, 3]
    assert pharm_tables(directionless=True)[2] == [0, 0, 0, 0, 0, 0, 0, 0, 3]
    # extended points: acceptor, donor, halogen -> 4 (anchor + extended) or 5 (extended only);
    # directionless scoring turns them off (pharmacophore_scoring.py:517-523)
    assert pharm_tables(extended_points=True)[2] == [4, 4, 2, 0, 4, 0, 0, 0, 3]
    assert pharm_tables(extended_points=True, only_extended=True)[2] == [5, 5, 2, 0, 5, 0, 0, 0, 3]
    assert pharm_tables(True, True, True)[2] == pharm_tables(directionless=True)[2]
    assert list(ScoreSpec(mode=_lib.MODE_PHARM, extended_points=True).struct().ph_cat) == [4, 4, 2, 0, 4, 0, 0, 0, 3]
    assert a[2] == 0.7 and a[3] == 0.7 and a[0] == 1.0
    assert ALPHA(200) == pytest.approx(1.118, abs=1e-9)
    s = ScoreSpec(similarity="tversky_fit").struct()
    assert s.similarity == _lib.SIM_TVERSKY and s.sigma == pytest.approx(0.05)
    with pytest.raises(ValueError):
        ScoreSpec(similarity="dice").struct()


def test_mode_registry_and_backend_names():
    assert canonical("esp") == "surf_esp" and canonical("esp_combo") == "vol_and_surf_esp"
    assert set(MODE_ATTRS) == set(MODE_SEEDS) == set(MODE_STEPS)
    for b in (None, "hip", "triton", "cuda", "gpu"):
        assert resolve_backend(b) == "hip"
    for b in ("numba", "cpu", "jax", "nope"):
        with pytest.raises(ValueError):
            resolve_backend(b)
    # every registered mode is served by the HIP backend, through the legacy aliases too
    assert set(HIP_MODES) == set(MODE_ATTRS)
    assert MoleculePairBatch.align_with_esp_combo is MoleculePairBatch.align_with_vol_and_surf_esp
    scores, aligned = MoleculePairBatch([]).align_with_vol_and_surf_esp(0.81, trans_init=True)
    assert len(scores) == 0 and aligned == []


def test_translation_seeds_rule():
    """trans_init seeds (alignment/_torch.py:329-398): 5 + 10 per centre; the base five are the
    ordinary rule's (identity + principal axes, centroid onto centroid); then the Fibonacci
    rotations repeat centre by centre and put the fit centroid on each centre."""
    from fast_shepherd_score_amd.seeds import (initial_seeds, quats_from_fibo, reference_table, rotmat,
                                               translation_seeds)
    rng = np.random.default_rng(3)
    ref, fit = rng.normal(size=(14, 3)) * 2, rng.normal(size=(11, 3)) + 1.5
    cen = rng.normal(size=(4, 3)) * 3
    q, t = translation_seeds(ref, fit, cen)
    assert q.shape == (45, 4) and t.shape == (45, 3) and q.dtype == np.float32
    q5, t5 = initial_seeds(ref[None], np.array([14]), fit[None], np.array([11]), 5)
    np.testing.assert_array_equal(q[:5], q5[0])
    np.testing.assert_allclose(t[:5], t5[0], atol=1e-6)
    np.testing.assert_array_equal(q[5:], np.tile(quats_from_fibo(10).astype(np.float32), (4, 1)))
    moved = np.einsum("sij,j->si", rotmat(q.astype(np.float64)), fit.mean(0)) + t
    np.testing.assert_allclose(moved[0], ref.mean(0), atol=1e-5)
    np.testing.assert_allclose(moved[5:], np.repeat(cen, 10, axis=0), atol=1e-5)
    q45, _ = translation_seeds(ref, fit, cen[:1], per_trans=45)   # the 50-restart table's rotations
    np.testing.assert_array_equal(q45[5:], reference_table(50)[5:])


# Parameter names of the reference's single-pair methods (container/_core.py:612-1560), in order.
PAIR_API = {
    "align_with_vol": ["no_H", "num_repeats", "trans_init", "lr", "max_num_steps", "use_jax",
                       "use_analytical", "verbose"],
    "align_with_vol_esp": ["lam", "no_H", "num_repeats", "trans_init", "lr", "max_num_steps",
                           "use_jax", "use_analytical", "verbose"],
    "align_with_surf": ["alpha", "num_repeats", "trans_init", "lr", "max_num_steps", "use_jax",
                        "use_analytical", "verbose"],
    "align_with_surf_esp": ["alpha", "lam", "num_repeats", "trans_init", "lr", "max_num_steps",
                            "use_jax", "use_analytical", "verbose"],
    "align_with_vol_and_surf_esp": ["alpha", "lam", "probe_radius", "esp_weight", "num_repeats",
                                    "trans_init", "lr", "max_num_steps", "use_jax", "verbose"],
    "align_with_vol_color": ["color_weight", "alpha", "similarity", "directionless",
                             "extended_points", "only_extended", "num_repeats", "trans_init", "lr",
                             "max_num_steps", "verbose"],
    "align_with_pharm": ["similarity", "extended_points", "only_extended", "num_repeats",
                         "trans_init", "lr", "max_num_steps", "use_jax", "verbose",
                         "use_vectorized", "use_analytical"],
    "score_with_surf": ["alpha", "use"],
    "score_with_esp": ["alpha", "lam", "use"],
    "score_with_pharm": ["similarity", "extended_points", "only_extended", "use"],
}


def test_single_pair_api_matches_reference_signatures():
    import inspect
    from fast_shepherd_score_amd.container import MoleculePair
    for name, params in PAIR_API.items():
        got = list(inspect.signature(getattr(MoleculePair, name)).parameters)[1:]
        assert got == params, name
    assert MoleculePair.align_with_esp is MoleculePair.align_with_surf_esp
    assert MoleculePair.align_with_esp_combo is MoleculePair.align_with_vol_and_surf_esp


def test_result_attributes_with_and_without_hydrogens():
    from fast_shepherd_score_amd.container import result_attrs
    assert result_attrs("vol") == ("transform_vol_noH", "sim_aligned_vol_noH")
    assert result_attrs("vol", no_H=False) == ("transform_vol", "sim_aligned_vol")
    assert result_attrs("vol_esp", no_H=False) == ("transform_vol_esp", "sim_aligned_vol_esp")
    assert result_attrs("surf", no_H=False) == ("transform_surf", "sim_aligned_surf")


def test_score_with_rejects_unknown_implementation():
    from fast_shepherd_score_amd.container import Molecule, MoleculePair
    m = Molecule(np.zeros((2, 3)), surf_pos=np.zeros((4, 3)))
    with pytest.raises(ValueError, match="use"):
        MoleculePair(m, m, device="cpu").score_with_surf(0.81, use="numba")


# Parameter names of the reference's scoring functions (score/gaussian_overlap.py,
# electrostatic_scoring.py, pharmacophore_scoring.py), in order.
SCORE_API = {
    "VAB_2nd_order": ["centers_1", "centers_2", "alpha"],
    "shape_tanimoto": ["centers_1", "centers_2", "alpha"],
    "get_overlap": ["centers_1", "centers_2", "alpha"],
    "VAB_2nd_order_esp": ["centers_1", "centers_2", "charges_1", "charges_2", "alpha", "lam"],
    "shape_tanimoto_esp": ["centers_1", "centers_2", "charges_1", "charges_2", "alpha", "lam"],
    "get_overlap_esp": ["centers_1", "centers_2", "charges_1", "charges_2", "alpha", "lam"],
    "esp_combo_score": ["centers_w_H_1", "centers_w_H_2", "centers_1", "centers_2", "points_1",
                        "points_2", "partial_charges_1", "partial_charges_2", "point_charges_1",
                        "point_charges_2", "radii_1", "radii_2", "alpha", "lam", "probe_radius",
                        "esp_weight"],
    "tanimoto_func": ["VAB", "VAA", "VBB"],
    "tversky_func": ["VAB", "VAA", "VBB", "sigma"],
    "get_overlap_pharm": ["ptype_1", "ptype_2", "anchors_1", "anchors_2", "vectors_1", "vectors_2",
                          "similarity", "extended_points", "only_extended",
                          "precomputed_self_overlaps", "directionless"],
    "get_pharm_combo_score": ["centers_1", "centers_2", "ptype_1", "ptype_2", "anchors_1",
                              "anchors_2", "vectors_1", "vectors_2", "alpha", "similarity",
                              "extended_points", "only_extended", "color_weight", "directionless"],
}


def test_score_module_matches_reference_signatures():
    import inspect
    from fast_shepherd_score_amd import score
    for name, params in SCORE_API.items():
        assert list(inspect.signature(getattr(score, name)).parameters) == params, name
    assert sorted(score.__all__) == sorted(SCORE_API)


def test_score_module_host_checks():
    import torch
    from fast_shepherd_score_amd import score
    v = torch.tensor([2.0, 1.0])
    assert torch.allclose(score.tanimoto_func(v, torch.tensor([3.0, 3.0]), torch.tensor([3.0, 3.0])),
                          torch.tensor([0.5, 0.2]))
    assert torch.allclose(score.tversky_func(v, torch.tensor([1.0, 2.0]), torch.tensor([1.0, 2.0]), 0.95),
                          torch.tensor([1.0, 0.5]))           # clamped at 1
    with pytest.raises(ValueError, match="similarity"):
        score.get_overlap_pharm(np.zeros(2), np.zeros(2), np.zeros((2, 3)), np.zeros((2, 3)),
                                np.zeros((2, 3)), np.zeros((2, 3)), similarity="dice")
    a = torch.zeros((2, 3), requires_grad=True)
    with pytest.raises(NotImplementedError, match="gradient"):
        score.get_overlap_pharm(np.zeros(2), np.zeros(2), a, a, a, a)


def test_score_as_loss_rejects_mismatched_shapes_before_any_launch():
    """gaussian_similarity / overlaps check the cloud and ESP shapes before packing: CPU tensors
    reach the shape check first, so each mismatch raises ValueError with no device involved."""
    from fast_shepherd_score_amd import autograd as ag
    ref, fit = torch.zeros((3, 5, 3)), torch.zeros((3, 7, 3))
    bad = [
        (dict(ref=ref, fit=torch.zeros((2, 7, 3))), "one fit cloud per ref cloud"),
        (dict(ref=ref, fit=torch.zeros((4, 7, 3))), "one fit cloud per ref cloud"),
        (dict(ref=torch.zeros((5, 3)), fit=fit), r"\(P, N, 3\)"),
        (dict(ref=ref, fit=torch.zeros((3, 7, 4))), r"\(P, N, 3\)"),
        (dict(ref=ref, fit=fit, ref_esp=torch.zeros((3, 4)), fit_esp=torch.zeros((3, 7))), "ref_esp"),
        (dict(ref=ref, fit=fit, ref_esp=torch.zeros((3, 5)), fit_esp=torch.zeros((2, 7))), "fit_esp"),
        (dict(ref=ref, fit=fit, ref_esp=torch.zeros((3, 5))), "both clouds"),
    ]
    for kw, msg in bad:
        for fn in (ag.gaussian_similarity, ag.overlaps):
            with pytest.raises(ValueError, match=msg):
                fn(**kw)
    # well-shaped CPU inputs pass the shape check and stop at the device check
    with pytest.raises(RuntimeError, match="GPU only"):
        ag.gaussian_similarity(ref, fit, torch.zeros((3, 5)), torch.zeros((3, 7)))


# Parameter names of the reference's MoleculePairBatch methods (container/_batch.py), in order;
# the HIP container may take further keywords after them.
BATCH_API = {
    "align_with_vol": ["no_H", "num_repeats", "trans_init", "lr", "max_num_steps", "num_workers",
                       "use_shmap", "num_buckets", "verbose", "backend", "alpha", "return_aligned"],
    "align_with_vol_esp": ["lam", "no_H", "num_repeats", "trans_init", "lr", "max_num_steps",
                           "num_workers", "use_shmap", "num_buckets", "verbose", "backend", "alpha",
                           "return_aligned"],
    "align_with_surf": ["alpha", "num_repeats", "trans_init", "lr", "max_num_steps", "use_jax",
                        "use_analytical", "num_workers", "use_shmap", "verbose", "backend",
                        "return_aligned"],
    "align_with_surf_esp": ["alpha", "lam", "num_repeats", "trans_init", "lr", "max_num_steps",
                            "use_jax", "use_analytical", "num_workers", "use_shmap", "verbose",
                            "backend", "return_aligned"],
    "align_with_vol_and_surf_esp": ["alpha", "lam", "probe_radius", "esp_weight", "num_repeats",
                                    "trans_init", "lr", "max_num_steps", "verbose", "backend",
                                    "return_aligned"],
    "align_with_vol_color": ["color_weight", "alpha", "num_repeats", "trans_init", "lr",
                             "max_num_steps", "verbose", "backend", "return_aligned"],
    "align_with_pharm": ["similarity", "extended_points", "only_extended", "num_repeats",
                         "trans_init", "lr", "max_num_steps", "num_workers", "use_shmap",
                         "num_buckets", "verbose", "backend", "return_aligned"],
}


def test_batch_api_starts_with_the_reference_signatures():
    import inspect
    from fast_shepherd_score_amd.container import MoleculePairBatch
    for name, params in BATCH_API.items():
        got = [p for p in inspect.signature(getattr(MoleculePairBatch, name)).parameters][1:]
        assert got[:len(params)] == params, name


# Positional and keyword-only parameters of the reference's kernel table (accel/kernels/cpu.py,
# the same signatures as the Triton kernels).
KERNEL_API = {
    "overlap_score_grad_se3_batch": (['A', 'B', 'q', 't'], ['alpha', 'N_real', 'M_real', 'NEED_GRAD', 'BLOCK', 'num_warps', 'num_stages']),
    "fused_adam_qt_with_tangent_proj": (['q', 't', 'dQ', 'dT', 'm_q', 'v_q', 'm_t', 'v_t', 'lr', 'beta1', 'beta2', 'eps'], []),
    "fused_adam_qt": (['q', 't', 'dQ', 'dT', 'm_q', 'v_q', 'm_t', 'v_t', 'lr', 'beta1', 'beta2', 'eps'], []),
    "_batch_self_overlap": (['P_pad', 'N_real', 'alpha'], []),
    "overlap_score_grad_esp_se3_batch": (['A', 'B', 'charges_A', 'charges_B', 'q', 't'], ['alpha', 'lam', 'N_real', 'M_real', 'NEED_GRAD', 'BLOCK', 'num_warps', 'num_stages']),
    "esp_comparison_batch": (['points', 'atoms', 'charges', 'point_esp', 'radii'], ['N_real', 'M_real', 'probe_radius', 'lam', 'BLOCK', 'num_warps', 'num_stages']),
    "pharm_score_grad_se3_batch": (['R', 't', 'ref_types', 'fit_types', 'ref_anchors', 'fit_anchors', 'ref_vectors', 'fit_vectors', 'alphas', 'Ks', 'cats'], ['N_real', 'M_real', 'NEED_GRAD', 'BLOCK', 'num_warps']),
    "pharm_color_score_grad_se3_batch": (['A', 'B', 'q', 't', 'ref_types', 'fit_types', 'alphas', 'Ks', 'cats'], ['N_real', 'M_real', 'NEED_GRAD', 'BLOCK', 'num_warps', 'num_stages']),
    "pharm_grad_dq_se3_batch": (['q', 't', 'ref_types', 'fit_types', 'ref_anchors', 'fit_anchors', 'ref_vectors', 'fit_vectors', 'alphas', 'Ks', 'cats'], ['N_real', 'M_real', 'NEED_GRAD', 'BLOCK', 'num_warps', 'num_stages']),
    "_batch_self_overlap_esp": (['P_pad', 'charges', 'N_real', 'alpha', 'lam'], []),
}


def test_kernel_table_matches_reference_signatures():
    import inspect
    from fast_shepherd_score_amd import kernels
    for name, (pos, kwo) in KERNEL_API.items():
        ps = inspect.signature(getattr(kernels, name)).parameters.values()
        assert [p.name for p in ps if p.kind == p.POSITIONAL_OR_KEYWORD] == pos, name
        assert [p.name for p in ps if p.kind == p.KEYWORD_ONLY] == kwo, name
    assert sorted(kernels.__all__) == sorted(KERNEL_API)


def test_kernel_table_host_helpers():
    import torch
    from fast_shepherd_score_amd import kernels
    from oracle import fss_oracle as orc
    rng = np.random.default_rng(4)
    q = torch.tensor(rng.normal(size=(50, 4)))
    q = q / q.norm(dim=1, keepdim=True)
    R = orc.rot_from_unit_quat(q)
    q2 = kernels._quat_from_rot(R)                      # the same rotation, sign aside
    np.testing.assert_allclose(orc.rot_from_unit_quat(q2).numpy(), R.numpy(), atol=1e-12)
    rr = kernels._ranges(3, 7, torch.tensor([2, 9, 0]), "cpu")
    assert rr.tolist() == [[0, 2], [7, 7], [14, 0]]      # counts clamped to the padding
    with pytest.raises(RuntimeError, match="device tensors"):
        kernels.overlap_score_grad_se3_batch(torch.zeros(1, 2, 3), torch.zeros(1, 2, 3),
                                             torch.zeros(1, 4), torch.zeros(1, 3))
