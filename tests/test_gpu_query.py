"""Queries over column files on the GPU, end to end (the reference's
experiments/main_tpch_* drivers through fastmulticolsort_amd.query): a
TPC-H layout text file of a lineitem-like shape (flag 2 bits, status 1,
date 12, key 27; duplicated tuples) is read, sorted or grouped on the GPU,
and the result file is checked against a host sort of the same rows."""
import numpy as np
import pytest

from fastmulticolsort_amd import columnar as cf
from fastmulticolsort_amd import query

pytestmark = pytest.mark.gpu

N = 1 << 16
WIDTHS = [2, 1, 12, 27]
WSTR = "2B1B12B27"


@pytest.fixture(scope="module")
def lineitem(tmp_path_factory):
    d = tmp_path_factory.mktemp("query")
    cols = query.generate(N, WIDTHS, group_num=N // 3, seed=11)
    cols[0][:] = cols[0] % np.uint16(3)  # flag takes 3 values, like l_returnflag
    p = d / "lineitem.tbl"
    cf.write_text(p, cols, WIDTHS, tpch=True)
    return d, p, cols


def host_order(cols, desc):
    """Sorted row ids: tuples compared column by column (descending columns
    negated), ties by row id."""
    keys = [[(-int(v) if d else int(v)) for v in c] for c, d in zip(cols, desc)]
    return np.array(sorted(range(len(cols[0])), key=lambda i: tuple(k[i] for k in keys) + (i,)), np.uint32)


def tuples(cols):
    return np.stack([c.astype(np.uint64) for c in cols], axis=1)


@pytest.mark.parametrize("comptype,asc_desc", [("ch", "1111"), ("st", "1111"), ("st", "12"), ("ch", "2121")])
def test_order_by_result_file_matches_host_sort(lineitem, comptype, asc_desc):
    d, p, cols = lineitem
    out = d / f"res_{comptype}_{asc_desc}.fmcs"
    info = query.main(["run", "-i", str(p), "-w", WSTR, "-a", asc_desc, "-p", comptype, "--out", str(out)])
    desc = query.parse_asc_desc(asc_desc, len(WIDTHS))
    exp = host_order(cols, desc)
    res = cf.read_columns(out)
    oid, gid, sorted_cols = res[0], res[1], res[2:]
    assert np.array_equal(oid, exp)
    for c, s in zip(cols, sorted_cols):
        assert np.array_equal(c[oid], s)
    # group ids: dense, starting at 0, changing exactly where the tuple changes
    t = tuples(sorted_cols)
    change = np.any(t[1:] != t[:-1], axis=1)
    assert gid[0] == 0 and np.array_equal(np.diff(gid.astype(np.int64)), change.astype(np.int64))
    assert info["ngroups"] == int(change.sum()) + 1
    meta = cf.columns_info(out)["columns"]
    assert [m["descending"] for m in meta[2:]] == desc
    cf.verify(out)


def test_group_by_groups_are_the_distinct_tuples(lineitem):
    d, p, cols = lineitem
    out = d / "groups.fmcs"
    info = query.main(["run", "-i", str(p), "-w", WSTR, "-o", "0", "--out", str(out)])
    ntup = np.unique(tuples(cols), axis=0).shape[0]
    assert info["ngroups"] == ntup
    res = cf.read_columns(out)
    oid, gid = res[0], res[1]
    assert np.array_equal(np.sort(oid), np.arange(N, dtype=np.uint32))
    t = tuples(cols)[oid]
    # every group holds one tuple, and different groups different tuples
    first = np.r_[True, gid[1:] != gid[:-1]]
    assert int(first.sum()) == ntup == int(np.unique(gid).size)
    assert np.all(np.all(t[1:] == t[:-1], axis=1) == (gid[1:] == gid[:-1]))


def test_binary_input_gives_the_same_result(lineitem):
    d, p, cols = lineitem
    b = d / "lineitem.fmcs"
    cf.write_columns(b, cols, WIDTHS, names=["flag", "status", "date", "key"])
    o1 = d / "from_text.fmcs"
    o2 = d / "from_binary.fmcs"
    query.main(["run", "-i", str(p), "-w", WSTR, "-p", "st", "--out", str(o1)])
    info = query.main(["run", "-i", str(b), "-p", "st", "--out", str(o2)])
    assert info["widths"] == WIDTHS
    assert np.array_equal(cf.read_columns(o1, [0])[0], cf.read_columns(o2, [0])[0])
    assert [m["name"] for m in cf.columns_info(o2)["columns"]] == ["oid", "group", "flag", "status", "date", "key"]


def test_exhaustive_search_runs_every_plan_to_the_same_order(lineitem):
    d, p, cols = lineitem
    summary, runs = query.main(["exhaustive", "-i", str(p), "-w", WSTR, "--max-plans", "24", "--max-rounds", "2",
                                "--warmup", "1", "--check"])
    # rounds 1 and 2 of S = 42 (1 + 41 plans); the stop comes after the round that passes 24
    assert summary["plans"] == len(runs) == 1 + 41
    assert sorted(tuple(r["split"]) for r in runs if r["rounds"] == 2) == [(a, 42 - a) for a in range(1, 42)]
    assert summary["all_same_order"] and all(r["same_order"] for r in runs)
    assert summary["best"]["gpu_ms"] == min(r["gpu_ms"] for r in runs)
