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
