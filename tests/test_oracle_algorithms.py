"""Pins oracle/algorithms.py (transform, replace, gather / scatter, sequence,
adjacent_difference, find / mismatch, extrema, inner_product, searches,
merge, is_sorted_until, partition_point): the reference tests' inputs with
their results (tests/golden/algorithms_known_answers.json), and every loop
form against an independent textbook routine on random inputs with many
ties (numpy searchsorted, argmin / argmax first occurrence, a stable argsort
of the concatenated inputs, flatnonzero)."""
import json
import os

import numpy as np
import pytest

from helpers import oracle_algorithms

KA = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "algorithms_known_answers.json")))


@pytest.fixture(scope="module")
def oa():
    return oracle_algorithms()


def i32(x):
    return np.array(x, dtype=np.int32)


def test_known_elementwise(oa):
    c = KA["transform_negate"]
    assert oa.transform(i32(c["in"]), "Negate").tolist() == c["out"]
    c = KA["transform_if_negate_odd"]
    x = i32(c["in"])
    assert oa.transform_if(x, "Negate", x % 2 != 0, x).tolist() == c["out"]
    c = KA["transform_if_stencil"]
    x = i32(c["in"])
    assert oa.transform_if(x, "Negate", i32(c["stencil"]) != 0, x).tolist() == c["out"]
    c = KA["replace"]
    assert oa.replace_copy(i32(c["in"]), c["old"], c["new"]).tolist() == c["out"]
    c = KA["replace_if_negative"]
    x = i32(c["in"])
    assert oa.replace_copy_mask(x, x < 0, c["new"]).tolist() == c["out"]
    c = KA["gather"]
    assert oa.gather(c["map"], i32(c["values"])).tolist() == c["out"]
    c = KA["gather_if"]
    assert oa.gather(c["map"], i32(c["values"]), i32([c["init"]] * 10), i32(c["stencil"]) != 0).tolist() == c["out"]
    c = KA["gather_if_even"]
    assert oa.gather(c["map"], i32(c["values"]), i32([c["init"]] * 10), i32(c["stencil"]) % 2 == 0).tolist() == c["out"]
    c = KA["scatter"]
    assert oa.scatter(i32(c["values"]), c["map"], np.zeros(10, np.int32)).tolist() == c["out"]
    c = KA["scatter_if"]
    assert oa.scatter(i32(c["values"]), c["map"], np.zeros(8, np.int32), i32(c["stencil"]) != 0).tolist() == c["out"]
    c = KA["sequence"]
    assert oa.sequence(10, np.int32).tolist() == c["out"]
    assert oa.sequence(10, np.int32, 1).tolist() == c["out_init1"]
    assert oa.sequence(10, np.int32, 1, 3).tolist() == c["out_init1_step3"]
    c = KA["adjacent_difference"]
    assert oa.adjacent_difference(i32(c["in"])).tolist() == c["minus"]
    assert oa.adjacent_difference(i32(c["in"]), "Plus").tolist() == c["plus"]


def test_known_reductions_searches_merge(oa):
    c = KA["extrema"]
    assert oa.min_element(c["in"]) == c["min"] and oa.max_element(c["in"]) == c["max"]
    c = KA["inner_product"]
    assert oa.inner_product(c["a"], c["b"], c["init"]) == c["out"]
    c = KA["find"]
    assert oa.find_first([v == 3 for v in c["in"]]) == c["find3"]
    assert oa.find_first([v == 9 for v in c["in"]]) is c["find9"]
    c = KA["mismatch"]
    assert oa.find_first([u != v for u, v in zip(c["a"], c["b"])]) == c["out"]
    c = KA["bounds_v"]
    assert [oa.lower_bound(c["input"], v) for v in c["values"]] == c["lower"]
    assert [oa.upper_bound(c["input"], v) for v in c["values"]] == c["upper"]
    assert [int(oa.binary_search(c["input"], v)) for v in c["values"]] == c["found"]
    c = KA["merge"]
    assert oa.merge(c["a"], c["b"])[0] == c["out"]
    for name, g in (("merge_by_key", False), ("merge_by_key_greater", True)):
        c = KA[name]
        k, v = oa.merge(c["ka"], c["kb"], g, c["va"], c["vb"])
        assert k == c["keys"] and v == c["values"], name


def test_loops_match_textbook_routines(oa):
    rng = np.random.default_rng(7)
    for trial in range(200):
        n = int(rng.integers(0, 60))
        a = np.sort(rng.integers(0, 12, n)).astype(np.int32)
        b = np.sort(rng.integers(0, 12, int(rng.integers(0, 60)))).astype(np.int32)
        va, vb = np.arange(len(a)), np.arange(len(a), len(a) + len(b))
        k1, v1 = oa.merge(a.tolist(), b.tolist(), False, va.tolist(), vb.tolist())
        k2, v2 = oa.merge_np(a, b, False, va, vb)
        assert k1 == k2.tolist() and v1 == v2.tolist()  # stable: payload pins tie order
        k1, v1 = oa.merge(a[::-1].tolist(), b[::-1].tolist(), True, va.tolist(), vb.tolist())
        k2, v2 = oa.merge_np(a[::-1], b[::-1], True, va, vb)
        assert k1 == k2.tolist() and v1 == v2.tolist()
        for q in range(-1, 14):
            assert oa.lower_bound(a.tolist(), q) == int(np.searchsorted(a, q, "left"))
            assert oa.upper_bound(a.tolist(), q) == int(np.searchsorted(a, q, "right"))
        x = rng.integers(-5, 5, max(1, n)).astype(np.int32)
        assert oa.min_element(x) == int(np.argmin(x)) and oa.max_element(x) == int(np.argmax(x))
        assert (oa.adjacent_difference(x) == oa.adjacent_difference_np(x)).all()
        m = x > 2
        assert oa.find_first(m.tolist()) == oa.find_first_np(m)
        srt = oa.is_sorted_until(x.tolist())
        assert srt == next((i for i in range(1, len(x)) if x[i] < x[i - 1]), len(x))
        pm = np.sort(x)[::-1] > 0  # a partitioned mask
        assert oa.is_partitioned(pm.tolist()) and oa.partition_point(pm.tolist()) == int(pm.sum())
        if m.any() and (~m).any():
            assert oa.is_partitioned(m.tolist()) == bool((np.flatnonzero(~m).min() > np.flatnonzero(m).max()))
