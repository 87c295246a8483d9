"""Pins the oracle's scan / compaction consumers (Unique, Reduce_By_Key,
Remove / Copy_If, Partition, transform scans, Count) before any GPU result is
checked against them: the reference tests' inputs with their results
(tests/golden/consumer_known_answers.json), and independent restatements
(itertools.groupby, numpy masks and cumulative sums) chosen so that a
plausible mistake fails -- comparing with the previous *kept* key instead of
the previous key, emitting a run's first key instead of its tail key, losing
order in the false half of a partition, applying init after instead of before
the first item."""
import itertools
import json
import os

import numpy as np
import pytest

from helpers import Oracle, workloads

KA = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "consumer_known_answers.json")))


@pytest.fixture(scope="module")
def orc():
    return Oracle()


@pytest.fixture(scope="module")
def wl():
    return workloads()


def test_known_unique(orc):
    for c in KA["unique"]:
        assert orc.unique(c["in"], "int32_t")[0].tolist() == c["out"]
    for c in KA["unique_by_key"]:
        k, v = orc.unique(c["keys"], "int32_t", np.array(c["values"], dtype=np.int32))
        assert k.tolist() == c["keys_out"] and v.tolist() == c["values_out"]


def test_known_reduce_by_key(orc):
    for c in KA["reduce_by_key"]:
        k, v = orc.reduce_by_key(c["keys"], c["values"], "int32_t", "int32_t")
        assert k.tolist() == c["keys_out"] and v.tolist() == c["values_out"]


def test_known_partition_remove(orc):
    for c in KA["partition"]:
        t, f = orc.partition(c["in"], "int32_t", c["pred"], c.get("stencil"))
        assert len(t) == c["count"] and t.tolist() + f.tolist() == c["out"]
    for c in KA["partition_copy"]:
        t, f = orc.partition(c["in"], "int32_t", c["pred"], c.get("stencil"))
        assert t.tolist() == c["true"] and f.tolist() == c["false"]
    for c in KA["remove"]:
        pred = ("eq", c["value"]) if "value" in c else c["pred"]
        assert orc.select(c["in"], "int32_t", pred, c.get("stencil"), negate=True).tolist() == c["out"], c["fn"]


def test_known_transform_scan_count(orc):
    for c in KA["transform_scan"]:
        out = orc.transform_scan(c["in"], "int32_t", c["unary"], c["op"], c["exclusive"], c.get("init", 0))
        assert out.tolist() == c["out"]
    x = np.arange(2000, dtype=np.int32) % 100
    assert orc.count(x, "int32_t", ("eq", 47)) == KA["count"][0]["out"]


def _runs(keys, div):
    """Runs of consecutive equivalent keys (itertools.groupby)."""
    if div > 1:
        return [list(g) for _, g in itertools.groupby(range(len(keys)), key=lambda i: int(keys[i]) // div
                if keys[i] >= 0 else -((-int(keys[i])) // div))]
    return [list(g) for _, g in itertools.groupby(range(len(keys)), key=lambda i: keys[i])]


@pytest.mark.parametrize("div", [1, 10])
def test_unique_matches_groupby(orc, wl, div):
    keys = wl.i32_range(5000, 11, -40, 40)
    keys = np.sort(keys)[::-1].copy() if div == 10 else keys  # long runs under div
    vals = np.arange(len(keys), dtype=np.int32)
    k, v = orc.unique(keys, "int32_t", vals, div=div)
    runs = _runs(keys, div)
    assert v.tolist() == [r[0] for r in runs]  # the first item of each run
    assert k.tolist() == [int(keys[r[0]]) for r in runs]


@pytest.mark.parametrize("div", [1, 10])
def test_reduce_by_key_matches_groupby(orc, wl, div):
    keys = np.sort(wl.i32_range(6000, 12, -300, 300))
    vals = wl.i32_range(6000, 13, -1000000000, 1000000000)
    k, v = orc.reduce_by_key(keys, vals, "int32_t", "int32_t", div=div)
    runs = _runs(keys, div)
    # the tail key of each run, and the run's sum wrapped to 32 bits
    assert k.tolist() == [int(keys[r[-1]]) for r in runs]
    sums = [((sum(int(vals[i]) for i in r) + 2 ** 31) % 2 ** 32) - 2 ** 31 for r in runs]
    assert v.tolist() == sums
    k2, m = orc.reduce_by_key(keys, vals, "int32_t", "int32_t", op="Maximum", div=div)
    assert m.tolist() == [max(int(vals[i]) for i in r) for r in runs]


def test_reduce_by_key_float_values(orc, wl):
    keys = np.repeat(np.arange(300, dtype=np.int32), 7)
    vals = wl.f32_unit(len(keys), 14)
    k, v = orc.reduce_by_key(keys, vals, "int32_t", "float")
    assert k.tolist() == list(range(300))
    ref = vals.astype(np.float64).reshape(300, 7).sum(axis=1)
    assert np.allclose(v, ref, rtol=1e-6)


def test_select_and_partition_match_numpy(orc, wl):
    x = wl.i32_range(7000, 15, -500, 500)
    st = wl.i32_range(7000, 16, 0, 3)
    mask = np.fmod(x, 3) == 1  # C++ truncating %
    assert orc.select(x, "int32_t", ("mod", 3, 1)).tolist() == x[mask].tolist()
    assert orc.select(x, "int32_t", ("mod", 3, 1), negate=True).tolist() == x[~mask].tolist()
    assert orc.select(x, "int32_t", ("nonzero",), stencil=st).tolist() == x[st != 0].tolist()
    assert orc.select(x, "int32_t", ("eq", 7), negate=True).tolist() == x[x != 7].tolist()
    t, f = orc.partition(x, "int32_t", ("mod", 3, 1))
    assert t.tolist() == x[mask].tolist() and f.tolist() == x[~mask].tolist()
    t, f = orc.partition(x, "int32_t", ("nonzero",), stencil=st)
    assert t.tolist() == x[st != 0].tolist() and f.tolist() == x[st == 0].tolist()


def test_transform_scan_matches_numpy(orc, wl):
    x = wl.i32_range(4000, 17, -1000, 1000)
    neg = (-x.astype(np.int64))
    assert orc.transform_scan(x, "int32_t", "Negate").tolist() == np.cumsum(neg).tolist()
    ex = orc.transform_scan(x, "int32_t", "Negate", exclusive=True, init=4)
    assert ex.tolist() == (4 + np.concatenate([[0], np.cumsum(neg)[:-1]])).tolist()
    sq = orc.transform_scan(x, "int64_t", "Square")
    assert sq.tolist() == np.cumsum(x.astype(np.int64) ** 2).tolist()


def test_count_matches_numpy(orc, wl):
    x = wl.i32_range(9000, 18, -50, 50)
    assert orc.count(x, "int32_t", ("mod", 7, 0)) == int(np.count_nonzero(np.fmod(x, 7) == 0))
    assert orc.count(x, "int32_t", ("eq", -3)) == int(np.count_nonzero(x == -3))
