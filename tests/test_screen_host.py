"""Host logic of the screening service (no GPU): launch planning and HBM admission."""
import numpy as np
import pytest

from fast_shepherd_score_amd.screen import Admission, plan_launch, restart_table
from fast_shepherd_score_amd.seeds import quats_from_fibo


@pytest.mark.parametrize("rem,cap", [([10, 10, 10], 12), ([0, 5, 100], 64), ([3], 1000), ([7, 1], 2),
                                      ([1000] * 7, 65536), ([5, 0, 0, 9], 3)])
def test_plan_launch_water_fills(rem, cap):
    take = plan_launch(rem, cap)
    assert sum(take) == min(cap, sum(rem))
    assert all(0 <= t <= r for t, r in zip(take, rem))
    live = [i for i, r in enumerate(rem) if r > 0]
    if cap >= len(live):          # every active request advances in every launch
        assert all(take[i] > 0 for i in live)
    # shares differ by at most one unless a request runs out
    unsat = [take[i] for i in live if take[i] < rem[i]]
    if unsat:
        assert max(unsat) - min(unsat) <= 1 or len(unsat) == 1


def test_admission_defers_when_over_budget():
    a = Admission(budget_bytes=1000, max_active=3)
    assert a.try_admit(400) and a.try_admit(400)
    assert not a.try_admit(400)          # over budget -> deferred
    a.release(400)
    assert a.try_admit(400)
    assert a.try_admit(150)              # 950 of 1000 used, third slot
    assert not a.try_admit(1)            # slots exhausted
    assert a.used == 950 and a.active == 3
    b = Admission(budget_bytes=10 ** 9, max_active=2)
    assert b.try_admit(1) and b.try_admit(1) and not b.try_admit(1)


def test_restart_table():
    t = restart_table(10)
    assert t.shape == (10, 4)
    np.testing.assert_allclose(t[0], [1, 0, 0, 0])
    np.testing.assert_allclose(t[1:], quats_from_fibo(9), atol=1e-7)
    np.testing.assert_allclose(np.linalg.norm(t, axis=1), 1.0, atol=1e-6)
