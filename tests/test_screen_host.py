"""Host logic of the screening service (no GPU): launch planning and arena sizing."""
import numpy as np
import pytest

from fast_shepherd_score_amd.screen import arena_bytes, plan_launch, restart_table
from fast_shepherd_score_amd.seeds import quats_from_fibo


@pytest.mark.parametrize("rem,cap", [([10, 10, 10], 12), ([0, 5, 100], 64), ([3], 1000), ([7, 1], 2),
                                      ([1000] * 7, 65536), ([5, 0, 0, 9], 3)])
def test_plan_launch_fills_oldest_first(rem, cap):
    take = plan_launch(rem, cap)
    assert sum(take) == min(cap, sum(rem))
    assert all(0 <= t <= r for t, r in zip(take, rem))
    # a request gets capacity only once every older one is fully served in this launch
    for i, t in enumerate(take):
        if t > 0:
            assert all(take[j] == rem[j] for j in range(i))


def test_arena_bytes_scale_with_the_service_shape():
    """Device bytes reserved by a screening service: linear in launch pairs and query slots."""
    a = arena_bytes(1000, 10, 4, 100, 512, 0)
    assert arena_bytes(2000, 10, 4, 100, 512, 0) - a == arena_bytes(1000, 10, 0, 100, 512, 0) - arena_bytes(0, 10, 0, 100, 512, 0)
    assert arena_bytes(1000, 10, 8, 100, 512, 0) - a == 4 * arena_bytes(0, 0, 1, 100, 512, 0)
    # one slot holds the query rows and two top-k buffers of (score, index, 4x4 transform)
    assert arena_bytes(0, 0, 1, 100, 512, 0) == 512 * 16 + 2 * 100 * 80 + 64


def test_restart_table():
    t = restart_table(10)
    assert t.shape == (10, 4)
    np.testing.assert_allclose(t[0], [1, 0, 0, 0])
    np.testing.assert_allclose(t[1:], quats_from_fibo(9), atol=1e-7)
    np.testing.assert_allclose(np.linalg.norm(t, axis=1), 1.0, atol=1e-6)


def test_screen_argument_checks_follow_the_reference():
    """screen() validates like screen.py:936-959 before any device work."""
    from fast_shepherd_score_amd.screen import screen
    q = {"points": np.zeros((3, 3), np.float32)}
    with pytest.raises(ValueError, match="lam"):
        screen(q, object(), "vol_esp")
    with pytest.raises(ValueError, match="no_H"):
        screen(q, object(), "vol", no_H=False)
    with pytest.raises(ValueError, match="one GPU"):
        screen(q, object(), "vol", ndev=2)
    with pytest.raises(ValueError, match="backend"):
        screen(q, object(), "vol", backend="numba")


def test_resolve_alpha_rules():
    from types import SimpleNamespace
    from fast_shepherd_score_amd.constants import ALPHA
    from fast_shepherd_score_amd.screen import resolve_alpha
    lib = SimpleNamespace(num_surf_points=200, shape=None)
    assert resolve_alpha(lib, "surf_esp", None) == pytest.approx(ALPHA(200))
    assert resolve_alpha(lib, "surf", 0.9) == 0.9
    assert resolve_alpha(lib, "vol", None) == 0.81
    assert resolve_alpha(lib, "pharm", None) == 0.81
    with pytest.raises(ValueError, match="explicit alpha"):
        resolve_alpha(lib, "vol_and_surf_esp", None)
    with pytest.raises(ValueError, match="differ in size"):
        resolve_alpha(SimpleNamespace(num_surf_points=None, shape=None), "surf", None)


def test_query_arrays_and_centring():
    from fast_shepherd_score_amd.container import Molecule
    from fast_shepherd_score_amd.screen import centered_query, query_arrays
    rng = np.random.default_rng(0)
    atoms = rng.normal(size=(6, 3)) + 5.0
    m = Molecule(atoms, surf_pos=rng.normal(size=(10, 3)) + 5.0, surf_esp=rng.normal(size=10),
                 pharm_types=[0, 2], pharm_ancs=atoms[:2], pharm_vecs=[[0, 0, 1.0], [1.0, 0, 0]],
                 atom_charges=rng.normal(size=6))
    d = query_arrays(m, "surf_esp")
    assert d["points"] is m.surf_pos and d["esp"] is m.surf_esp
    assert query_arrays(m, "vol_esp")["esp"] is m.atom_charges
    with pytest.raises(ValueError, match="missing"):
        query_arrays(m, "vol_and_surf_esp")
    d = query_arrays(m, "vol_color")
    c = centered_query(d, "vol_color")
    np.testing.assert_allclose(c["atom_pos"].mean(0), 0, atol=1e-6)
    np.testing.assert_allclose(c["pharm_anchors"], c["atom_pos"][:2], atol=1e-6)
    np.testing.assert_array_equal(c["pharm_vectors"], d["pharm_vectors"])
    np.testing.assert_allclose(d["atom_pos"].mean(0), atoms.mean(0), rtol=1e-6)   # caller's arrays kept
