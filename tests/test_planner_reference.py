"""The planner against the reference's own planner, compiled here from its
sources (oracle/_ref, test infrastructure).

The product's planner prices plans with GPU-measured constants; its
structure -- the construct-entry walk and its statistics, the
card_per_bit cardinalities (double, and the uint64 variant of the RRS
enumerator with its per-entry conversion and wrapping products), the group
counts, f(g), the per-plan cost composition, the greedy cut placement and
the RRS search -- is shared with fmcs_cost_primitives entry points, which
are run here with the reference's own static estimators as the primitives
and compared with the reference's private members on the same inputs:

* statistics, group counts, u64 cardinalities: exact;
* costs (fp64): relative 1e-12 (the reference's build may fuse a multiply
  and an add where this one does not);
* greedy plans and RRS visited plans: identical lists / sets.

The RRS engine's seed is the one the reference took from time(0)
(src/rrs.h:68).  The RRS bowls of tests/rrs_test.cpp:78-100 pin the
generic search (fmcs_rrs.h) on its own.
"""
import itertools
import os
import subprocess

import numpy as np
import pytest

import oracle_bind as ob
from fastmulticolsort_amd import learned, planner

pytestmark = pytest.mark.skipif(ob.ref() is None, reason="the reference library is built in this container only")

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REL = 1e-12


def workloads():
    """(n, bws, cards): configs 1, 2, 4, a tiny table-sized n, and seeded
    config-5 requests (cards kept below 2^32, the reference RRS's uint32)."""
    out = [(1 << 24, [17, 33], [8192, 1 << 24]),
           (1 << 26, [12, 20, 30], [4096, 1 << 20, 60000000]),
           (1 << 25, [2, 1, 12, 27], [3, 2, 2526, 30000000]),
           (200, [5, 9, 20], [20, 150, 190])]
    rng = np.random.default_rng(424242)
    for _ in range(6):
        r = learned.make_request(rng, max_log2=24)
        out.append((r["n"], r["bws"], [min(c, (1 << 32) - 1) for c in r["cards"]]))
    return out


def orders(k):
    ids = list(range(k))
    return [ids, ids[::-1]] if k > 1 else [ids]


def some_plans(S, seed):
    """Splits of S: rounds <= 2 exhaustive (or a sample), plus random splits of 3..6 rounds."""
    plans = []
    for r in (1, 2):
        plans += planner.split_exhaustive(S, r, 64)
    rng = np.random.default_rng(seed)
    for _ in range(40):
        r = int(rng.integers(max(1, -(-S // 64)), 7))
        if r * 64 < S or r > S:
            continue
        while True:
            cuts = sorted(rng.choice(np.arange(1, S), size=r - 1, replace=False)) if r > 1 else []
            sp = list(np.diff([0] + list(cuts) + [S]).astype(int))
            if all(1 <= w <= 64 for w in sp):
                plans.append([int(w) for w in sp])
                break
    return [list(map(int, p)) for p in plans]


def close(a, b, rel=REL):
    return abs(a - b) <= rel * max(abs(a), abs(b), 1e-300)


@pytest.fixture(scope="module")
def prims():
    return planner.CostPrimitives.from_addresses(*ob.ref_cost_callbacks())


@pytest.mark.parametrize("wi", range(10))
def test_statistics_group_counts_and_costs(prims, wi):
    n, bws, cards = workloads()[wi]
    S = sum(bws)
    ref = ob.RefPlanner(n, bws, cards)
    max_round = max(1, -(-(2 * S - 2) // 16))
    checked = 0
    for order in orders(len(bws)):
        for b in range(1, S + 1):
            assert close(planner.estimate_group_num(b, bws, cards, order), ref.group_num(b, order), 1e-14), b
        for sp in some_plans(S, wi):
            # getCardinalityFromStatistics: cardinalities and the stitch statistics
            c, a, s = planner.plan_statistics(bws, cards, order, sp)
            rc, ra, rs = ref.card_stats(order, sp)
            assert (a, s) == (ra, rs), (order, sp)
            assert all(close(x, y, 1e-14) for x, y in zip(c, rc)), (order, sp, c, rc)
            # PlanEnumRRS's uint64 cardinalities and its GetCost
            cu, _ = planner.plan_statistics_u64(bws, cards, order, sp)
            rcost, rcu, ra2, rs2 = ref.rrs_cost(order, sp, max_round)
            assert cu == rcu and (ra2, rs2) == (ra, rs), (order, sp, cu, rcu)
            t = planner.estimate_plan_prim(prims, n, bws, cards, order, sp, truncate=True)
            assert close(t, rcost), (order, sp, t, rcost)
            # estimate_plan_collection_greedy's costing
            e = planner.estimate_plan_prim(prims, n, bws, cards, order, sp, truncate=False)
            assert close(e, ref.plan_cost(order, sp)), (order, sp)
            checked += 1
    ref.close()
    assert checked >= 20


def test_column_cost_and_its_special_cases(prims):
    n = 1 << 25
    ref = ob.RefPlanner(n, [20, 20], [1 << 20, 1 << 20])
    for g in [0.25, 1.0, 1.0 + 1e-7, 1.0 + 1e-5, 1.5, 2.0, 2.49, 2.51, 3.0, 100.0, 255.6, 256.4, 511.5, 512.5, 1000.7,
              4096.0, 1e6, 3.3e7]:
        for bt in (0, 1, 2):
            mine = planner.column_cost_prim(prims, n, g, bt)
            assert close(mine, ref.column_cost(g, bt)), (g, bt)
            if g - 1.0 < 1e-6:
                assert mine == 0.0  # groups of one row need no sort
    ref.close()


def test_uint64_products_wrap_like_the_reference():
    # 4 columns of 2^31 distinct 31-bit values: the groups before round 4
    # exceed 2^64, and the reference's uint64 products wrap (mod 2^64)
    bws, cards = [31, 31, 31, 31], [(1 << 31) - 1] * 4
    order = [0, 1, 2, 3]
    sp = [31, 31, 31, 31]
    cu, groups = planner.plan_statistics_u64(bws, cards, order, sp)
    assert groups[0] == 1 and groups[1] == cu[0]
    assert groups[2] == (cu[0] * cu[1]) % (1 << 64)
    assert groups[3] == (cu[0] * cu[1] * cu[2]) % (1 << 64)
    ref = ob.RefPlanner(1 << 24, bws, cards)
    _, rcu, _, _ = ref.rrs_cost(order, sp, 8)
    assert cu == rcu
    ref.close()


@pytest.mark.parametrize("wi", range(10))
def test_greedy_plans_match_the_reference(prims, wi):
    n, bws, cards = workloads()[wi]
    ref = ob.RefPlanner(n, bws, cards)
    for order in orders(len(bws)):
        for rounds in range(3, 7):
            mine = planner.greedy_plans_prim(prims, n, bws, cards, order, rounds)
            assert mine == ref.greedy_plans(rounds, order), (order, rounds)
    ref.close()


@pytest.mark.parametrize("wi", [0, 1, 2, 4, 6, 8])
def test_rrs_visits_the_reference_plans(prims, wi):
    n, bws, cards = workloads()[wi]
    S = sum(bws)
    order = list(range(len(bws)))
    max_round = max(1, -(-(2 * S - 2) // 16))
    ref = ob.RefPlanner(n, bws, cards)
    seed, visited = ref.plan_enum_rrs(order, max_round)
    mine = planner.plan_enum_rrs_prim(prims, n, bws, cards, order, max_round, seed)
    ref.close()
    assert len(visited) >= 44
    assert sorted(map(tuple, mine)) == sorted(map(tuple, visited))


@pytest.fixture(scope="module")
def bowl_lib(tmp_path_factory):
    import ctypes as C
    out = tmp_path_factory.mktemp("rrs") / "librrs_bowl.so"
    subprocess.check_call(["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-I",
                           os.path.join(ROOT, "fastmulticolsort_amd", "csrc"),
                           os.path.join(ROOT, "tests", "rrs_bowl.cpp"), "-o", str(out)])
    lib = C.CDLL(str(out))
    lib.rrs_bowl.restype = C.c_double
    lib.rrs_bowl.argtypes = [C.c_int, C.c_uint64, C.POINTER(C.c_double), C.POINTER(C.c_double)]
    return lib


@pytest.mark.parametrize("seed", [1, 7, 2015, 123456789])
def test_rrs_bowls(bowl_lib, seed):
    # tests/rrs_test.cpp:78-100: (x-1)^2 + (y-2)^2 has its minimum 0 at (1, 2);
    # the two-bowl function's lower bowl has its minimum -1 at (-5, -1)
    import ctypes as C
    x, y = C.c_double(), C.c_double()
    c = bowl_lib.rrs_bowl(1, seed, C.byref(x), C.byref(y))
    assert abs(c) < 1e-3 and abs(x.value - 1) < 0.05 and abs(y.value - 2) < 0.05
    c = bowl_lib.rrs_bowl(2, seed, C.byref(x), C.byref(y))
    assert abs(c + 1) < 1e-3 and abs(x.value + 5) < 0.05 and abs(y.value + 1) < 0.05
