"""Admission by memory (no GPU): the screening service's query-row pools, their sizing from an
injected budget, and the FIFO admission step, driven through the same native host functions the
service's worker calls (fss_plan_query_pools, fss_rowpool_*, fss_admit_fifo)."""
import itertools

import numpy as np
import pytest

from fast_shepherd_score_amd import _lib
from fast_shepherd_score_amd.screen import RowPool, admit_fifo, plan_query_pools, query_row_bytes

BYTES = _lib.QPOOL_ROW_BYTES


def test_pool_sizes_without_a_budget_hold_max_active_queries():
    assert plan_query_pools(_lib.MODE_SHAPE, 8, 100, 50) == [800, 0, 0, 0]
    assert plan_query_pools(_lib.MODE_PHARM, 8, 100, 50) == [0, 400, 0, 0]
    assert plan_query_pools(_lib.MODE_SHAPE_PHARM, 8, 100, 50) == [800, 400, 0, 0]
    assert plan_query_pools(_lib.MODE_ESP_COMBO, 4, 100, 50, 300, 60) == [400, 0, 1200, 240]
    # the ShaEP shape term on the surface rows reads no shape rows
    assert plan_query_pools(_lib.MODE_ESP_COMBO, 4, 100, 50, 300, 60, shape_from_surface=True) == [0, 0, 1200, 240]
    # a budget of max_active maximum queries or more changes nothing
    one = query_row_bytes(100, 50)
    assert plan_query_pools(_lib.MODE_SHAPE_PHARM, 8, 100, 50, query_budget_bytes=8 * one) == [800, 400, 0, 0]
    assert plan_query_pools(_lib.MODE_SHAPE_PHARM, 8, 100, 50, query_budget_bytes=100 * one) == [800, 400, 0, 0]


@pytest.mark.parametrize("m", [1.0, 1.5, 2.0, 3.7, 7.99])
def test_a_budget_is_split_in_proportion_to_one_query(m):
    mq = (100, 50, 0, 0)
    one = query_row_bytes(*mq)
    budget = int(m * one)
    rows = plan_query_pools(_lib.MODE_SHAPE_PHARM, 8, *mq, query_budget_bytes=budget)
    assert sum(r * b for r, b in zip(rows, BYTES)) <= budget
    for r, n in zip(rows, mq):
        assert n <= r <= 8 * n
        assert r == int(budget * n // one)     # m maximum queries' worth of each pool


def test_a_budget_below_one_query_is_refused():
    one = query_row_bytes(100, 50)
    with pytest.raises(_lib.FssError, match="does not hold one query"):
        plan_query_pools(_lib.MODE_SHAPE_PHARM, 8, 100, 50, query_budget_bytes=one - 1)
    with pytest.raises(_lib.FssError, match=">= 0"):
        plan_query_pools(_lib.MODE_SHAPE, 8, 100, query_budget_bytes=-1)
    with pytest.raises(_lib.FssError, match="2\\^31"):
        plan_query_pools(_lib.MODE_SHAPE, 4096, 1 << 20)


def _first_fit_bitmap(free, n):
    """First-fit on a bitmap of free rows: the lowest start of n consecutive free rows."""
    run = 0
    for i, f in enumerate(free):
        run = run + 1 if f else 0
        if run == n:
            return i - n + 1
    return None


def test_row_pool_matches_first_fit_on_a_bitmap():
    """Every alloc and free of a random sequence against a bitmap restatement of first-fit (a
    missed merge of neighbouring free ranges, or a wrong split, changes a later answer)."""
    rng = np.random.default_rng(3)
    for trial in range(20):
        cap = int(rng.integers(1, 40))
        pool = RowPool(cap)
        free = [True] * cap
        held = []
        for _ in range(200):
            if held and rng.random() < 0.45:
                first, n = held.pop(int(rng.integers(len(held))))
                pool.free(first, n)
                for i in range(first, first + n):
                    assert not free[i]
                    free[i] = True
            else:
                n = int(rng.integers(1, max(2, cap // 2)))
                got, want = pool.alloc(n), _first_fit_bitmap(free, n)
                assert got == want, (trial, cap, n)
                if got is not None:
                    held.append((got, n))
                    for i in range(got, got + n):
                        free[i] = False
            f, largest = pool.stats()
            assert f == sum(free)
            assert largest == max((len(list(g)) for k, g in itertools.groupby(free) if k), default=0)
        for first, n in held:
            pool.free(first, n)
        assert pool.stats() == (cap, cap)      # every range merged back into one


def test_row_pool_rejects_ranges_it_did_not_hand_out():
    pool = RowPool(10)
    a = pool.alloc(4)
    assert a == 0 and pool.alloc(0) == 0
    with pytest.raises(_lib.FssError, match="not an allocated range"):
        pool.free(2, 4)          # half free
    with pytest.raises(_lib.FssError, match="not an allocated range"):
        pool.free(8, 4)          # past the end
    pool.free(a, 4)
    with pytest.raises(_lib.FssError, match="not an allocated range"):
        pool.free(a, 4)          # twice
    assert pool.alloc(11) is None and pool.alloc(10) == 0


def _simulate(budget, sizes, max_active, finish_order, mq=(64, 16)):
    """Drives the admission step as the service's worker does: admit what fits, then finish one
    admitted query (per finish_order) and release its ranges; returns the admission order, the
    steps at which a query waited for memory with a slot free, and the most held at once."""
    rows = plan_query_pools(_lib.MODE_SHAPE_PHARM, max_active, *mq, query_budget_bytes=budget)
    pools = [RowPool(r) for r in rows]
    pending = list(range(len(sizes)))
    active = {}                   # query -> (first rows, rows)
    order, waits, most = [], [], 0
    step = 0
    while pending or active:
        need = [list(sizes[q]) + [0, 0] for q in pending[:max_active - len(active)]]
        firsts = admit_fifo(pools, need, max_active - len(active))
        for q, first in zip(list(pending), firsts):
            active[q] = (first, list(sizes[q]) + [0, 0])
            order.append(q)
            pending.remove(q)
        if pending and len(active) < max_active:
            waits.append(step)
        most = max(most, len(active))
        # the ranges held in each pool are disjoint and inside it
        for c, cap in enumerate(rows):
            spans = sorted((f[c], f[c] + n[c]) for f, n in active.values() if n[c])
            assert all(0 <= a < b <= cap for a, b in spans)
            assert all(spans[i][1] <= spans[i + 1][0] for i in range(len(spans) - 1))
        assert active, "nothing admitted with every slot and row free"
        done = finish_order(sorted(active), step)
        first, n = active.pop(done)
        for c in range(len(pools)):
            pools[c].free(first[c], n[c])
        step += 1
    for p, r in zip(pools, rows):
        assert p.stats() == (r, r)
    return order, waits, most


@pytest.mark.parametrize("seed", range(4))
def test_an_injected_budget_defers_admission_fifo(seed):
    """A budget of 2.5 maximum queries with 8 slots: queries wait for memory with slots free,
    are admitted strictly in submission order, and every one is admitted in the end."""
    rng = np.random.default_rng(seed)
    mq = (64, 16)
    sizes = [(int(rng.integers(8, 65)), int(rng.integers(1, 17))) for _ in range(40)]
    budget = int(2.5 * query_row_bytes(*mq))
    pick = lambda act, step: act[int(rng.integers(len(act)))]       # queries finish in any order
    order, waits, most = _simulate(budget, sizes, 8, pick, mq)
    assert order == list(range(len(sizes)))
    assert waits, "the budget never held a query back"
    assert most < 8
    # without a budget the same run is bound by slots alone
    order, waits, most = _simulate(0, sizes, 8, pick, mq)
    assert order == list(range(len(sizes))) and waits == [] and most == 8


def test_small_queries_share_the_room_of_one_large_one():
    """Admission counts each query's own rows: a budget of two maximum queries admits many
    small ones at once."""
    mq = (64, 16)
    budget = 2 * query_row_bytes(*mq)
    order, waits, most = _simulate(budget, [(8, 2)] * 16, 16, lambda act, step: act[0], mq)
    assert most == 16 and waits == []
    order, waits, most = _simulate(budget, [(64, 16)] * 16, 16, lambda act, step: act[0], mq)
    assert most == 2 and waits


def test_a_waiting_query_is_not_overtaken():
    """A large query at the head of the queue holds back the small ones behind it."""
    pools = [RowPool(100)]
    assert admit_fifo(pools, [[60]], 4) == [[0]]
    assert admit_fifo(pools, [[50], [10], [10]], 3) == []            # 50 does not fit: nobody enters
    pools[0].free(0, 60)
    assert admit_fifo(pools, [[50], [10], [10]], 3) == [[0], [50], [60]]
    assert admit_fifo(pools, [[10], [10]], 1) == [[70]]              # one slot free: one admitted
    # a query that does not fit in a later pool gives back what it took in the earlier ones
    a, b = RowPool(10), RowPool(5)
    assert admit_fifo([a, b], [[4, 6]], 2) == []
    assert a.stats() == (10, 10) and b.stats() == (5, 5)
