"""Pins the oracle (oracle/oracle.cpp) before it is trusted as the checker.

Each check is chosen so that a plausible mistake in the oracle fails it:
the reference's own known answers (docs/QuickStartGuide.md, cpp/test/*),
brute force against an independent definition (Python's stable `sorted`),
closed forms, and the -0.0/+0.0 and Greater-tie cases that a textbook float
flip or a reversed ascending sort would get wrong.
"""
import itertools
import json
import math
import os

import numpy as np
import pytest

from helpers import Oracle, workloads, NP

GOLDEN = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "known_answers.json")))


@pytest.fixture(scope="module")
def orc():
    return Oracle()


@pytest.fixture(scope="module")
def wl():
    return workloads()


# ---------------------------------------------------------------- known answers
def test_known_sort(orc):
    for c in GOLDEN["sort"]:
        out = orc.sort(c["in"], c["elem"], descending=c["comp"] == "Greater")
        assert out.tolist() == c["out"]


def test_known_sort_by_key(orc):
    for c in GOLDEN["sort_by_key"]:
        k, v = orc.sort_by_key(c["keys"], np.array(c["values"], dtype=NP[c["velem"]]), c["kelem"],
                               descending=c["comp"] == "Greater")
        assert k.tolist() == c["keys_out"]
        assert v.tolist() == c["values_out"]


def test_known_scans(orc):
    for c in GOLDEN["inclusive_scan"]:
        assert orc.inclusive_scan(c["in"], c["elem"], c["op"]).tolist() == c["out"]
    for c in GOLDEN["exclusive_scan"]:
        assert orc.exclusive_scan(c["in"], c["elem"], c["init"], c["op"]).tolist() == c["out"]


def test_known_scan_by_key(orc):
    for c in GOLDEN["scan_by_key"]:
        out = orc.scan_by_key(c["keys"], c["values"], "int32_t", "int32_t", c["exclusive"], c["init"])
        assert out.tolist() == c["out"]


def test_known_reduce(orc):
    for c in GOLDEN["reduce"]:
        assert orc.reduce(c["in"], c["elem"], c["init"], c["op"]) == c["out"]


def test_known_copy_if(orc):
    for c in GOLDEN["copy_if"]:
        assert orc.copy_if_mod(c["in"], c["modulus"], c["remainder"]).tolist() == c["out"]
    for c in GOLDEN["copy_if_stencil"]:
        assert orc.copy_if_stencil_mod(c["in"], c["stencil"], c["modulus"], c["remainder"]).tolist() == c["out"]


# ---------------------------------------------------------------- sort: brute force + special cases
def _py_stable_perm(keys, descending):
    # independent definition: Python's stable sort on the numeric value
    # (for floats -0.0 == +0.0 compares equal, exactly the reference preorder)
    idx = list(range(len(keys)))
    return sorted(idx, key=lambda i: keys[i], reverse=descending)


@pytest.mark.parametrize("elem", ["int32_t", "uint32_t", "float", "int8_t", "uint16_t", "int64_t", "double"])
@pytest.mark.parametrize("descending", [False, True])
def test_sort_perm_brute_force(orc, elem, descending):
    vals = {"int32_t": [-7, 0, 3, 3, -7, 2**31 - 1, -2**31],
            "uint32_t": [0, 5, 5, 2**32 - 1, 1, 2**31, 7],
            "float": [0.0, -0.0, 1.5, -2.25, -0.0, 1.5, float("inf")],
            "int8_t": [-128, 127, 0, -1, 0, 5, -128],
            "uint16_t": [65535, 0, 3, 3, 65535, 1, 2],
            "int64_t": [-2**63, 2**63 - 1, 0, -1, -1, 5, 0],
            "double": [-0.0, 0.0, -1e300, 1e-300, -0.0, 3.0, -float("inf")]}[elem]
    for n in range(1, 7):
        for perm in itertools.permutations(vals[:n]):
            keys = np.array(perm, dtype=NP[elem])
            got = orc.sort_perm(keys, elem, descending).tolist()
            want = _py_stable_perm([k.item() for k in keys], descending)
            assert got == want, (elem, descending, perm)


def test_sort_signed_zero_order(orc):
    # -0.0 and +0.0 are one key (built_in.h:436-443): a stable sort keeps
    # their input order.  The textbook flip (sign ? ~u : u|0x80000000) would
    # put every -0.0 before every +0.0 and fail this.
    keys = np.array([0.0, -0.0, 1.0, -0.0, 0.0, -1.0], dtype=np.float32)
    out = orc.sort(keys, "float")
    assert out.view(np.uint32).tolist() == np.array([-1.0, 0.0, -0.0, -0.0, 0.0, 1.0],
                                                    dtype=np.float32).view(np.uint32).tolist()
    out = orc.sort(keys, "float", descending=True)
    assert out.view(np.uint32).tolist() == np.array([1.0, 0.0, -0.0, -0.0, 0.0, -1.0],
                                                    dtype=np.float32).view(np.uint32).tolist()


def test_sort_greater_ties_keep_input_order(orc):
    # Greater is a stable descending sort; reversing an ascending sort would
    # put equal keys in reverse input order.
    k, v = orc.sort_by_key([1, 2, 1, 2], np.array([0, 1, 2, 3], dtype=np.int32), "int32_t", descending=True)
    assert k.tolist() == [2, 2, 1, 1]
    assert v.tolist() == [1, 3, 0, 2]


def test_sort_shuffled_identity(orc):
    # cpp/test/test_sort.cpp:82-168: shuffled 1..10000 sorts to the identity,
    # payload 10000-i follows its key
    rng = np.random.default_rng(1)
    keys = np.arange(1, 10001, dtype=np.int32)
    vals = (10000 - np.arange(10000)).astype(np.int32)
    p = rng.permutation(10000)
    k, v = orc.sort_by_key(keys[p], vals[p], "int32_t")
    assert (k == keys).all() and (v == vals).all()


def test_sort_matches_numpy_stable(orc, wl):
    keys = wl.u32_low_entropy(1 << 14, 2)
    perm = orc.sort_perm(keys, "uint32_t")
    assert (perm == np.argsort(keys, kind="stable")).all()
    f = wl.f32_normal_signed_zeros(1 << 14, 3, 4)
    perm = orc.sort_perm(f, "float")
    # numpy's stable argsort compares floats with <, so -0.0 == +0.0 as well
    assert (perm == np.argsort(f, kind="stable")).all()


# ---------------------------------------------------------------- scans
def test_scan_closed_forms(orc):
    ones = np.ones(1000, dtype=np.int32)
    i = np.arange(1000)
    # Counter(1) x 1000 (test_scan.cpp:9-19): inclusive of 1..1000 = (i+1)(i+2)/2
    assert (orc.inclusive_scan(i + 1, "int32_t") == (i + 1) * (i + 2) // 2).all()
    assert (orc.exclusive_scan(ones, "int32_t", 7) == 7 + i).all()


def test_scan_wraps_like_uint32(orc, wl):
    x = wl.i32_range(1 << 12, 9, -2**31, 2**31 - 1)
    want = np.cumsum(x.view(np.uint32), dtype=np.uint32).view(np.int32)
    assert (orc.inclusive_scan(x, "int32_t") == want).all()
    ex = orc.exclusive_scan(x, "int32_t", 3)
    assert ex[0] == 3 and (ex[1:] == (want[:-1].view(np.uint32) + np.uint32(3)).view(np.int32)).all()


def test_scan_max_min(orc, wl):
    x = wl.i32_range(5000, 11, -1000, 1000)
    assert (orc.inclusive_scan(x, "int32_t", "Maximum") == np.maximum.accumulate(x)).all()
    assert (orc.inclusive_scan(x, "int32_t", "Minimum") == np.minimum.accumulate(x)).all()


def test_scan_by_key_closed_form(orc):
    # test_scan_by_key.cpp:8-24: keys i/300, values i%300 -> triangular numbers
    i = np.arange(3000)
    out = orc.scan_by_key(i // 300, i % 300, "int32_t", "int32_t")
    r = i % 300
    assert (out == r * (r + 1) // 2).all()
    ex = orc.scan_by_key(i // 300, i % 300, "int32_t", "int32_t", exclusive=True, init=2)
    assert (ex == 2 + (r - 1) * r // 2).all()


def test_scan_by_key_per_segment(orc, wl):
    keys = np.repeat(np.arange(50), wl.segment_sizes(50, 7, 1, 100)[0].astype(np.int64))
    vals = wl.i32_range(len(keys), 5, -1000, 1000)
    out = orc.scan_by_key(keys, vals, "int32_t", "int32_t")
    starts = np.flatnonzero(np.r_[True, keys[1:] != keys[:-1]])
    ends = np.r_[starts[1:], len(keys)]
    for s, e in zip(starts, ends):
        assert (out[s:e] == np.cumsum(vals[s:e])).all()


def test_scan_by_key_float_keys_signed_zero(orc):
    # EqualTo on floats: -0.0 == +0.0 continues a segment, NaN never does
    out = orc.scan_by_key([0.0, -0.0, float("nan"), float("nan"), 1.0], [1, 1, 1, 1, 1], "float", "int32_t")
    assert out.tolist() == [1, 2, 1, 1, 1]


# ---------------------------------------------------------------- reduce
def test_reduce_closed_form(orc):
    n = 100000
    assert orc.reduce(np.arange(1, n + 1, dtype=np.int64), "int64_t") == n * (n + 1) // 2
    assert orc.reduce(np.arange(10, dtype=np.int32), "int32_t", 5, "Maximum") == 9
    assert orc.reduce([], "int32_t", 4) == 4


def test_reduce_fp64_and_ref_tree(orc, wl):
    x = wl.f32_unit(1 << 16, 6)
    exact = math.fsum(x.astype(np.float64))
    assert abs(orc.sum_f32_fp64(x) - exact) <= 1e-9 * exact
    # the reference tree order is exact on integer-valued floats below 2^24
    ints = np.arange(1, 5001, dtype=np.float32)
    assert orc.reduce_ref_tree_f32(ints) == 5000 * 5001 / 2
    assert orc.reduce_ref_tree_f32(ints, init=7.0) == 5000 * 5001 / 2 + 7
    assert abs(orc.reduce_ref_tree_f32(x) - exact) <= 1e-5 * exact


def test_reduce_ref_tree_tiny_brute_force(orc):
    # n <= 256: one block; the halving tree adds buf[t+s] into buf[t]
    rng = np.random.default_rng(3)
    for n in [1, 2, 3, 5, 17, 255, 256, 257]:
        x = rng.standard_normal(n).astype(np.float32)
        buf = list(x) + [np.float32(0)] * (256 * ((n + 255) // 256) - n)
        if n <= 256:
            s = 128
            while s:
                for t in range(s):
                    if t + s < n:
                        buf[t] = np.float32(buf[t] + buf[t + s])
                s >>= 1
            assert orc.reduce_ref_tree_f32(x) == buf[0]


# ---------------------------------------------------------------- copy_if
def test_copy_if_negative_multiples(orc, wl):
    x = wl.i32_range(1 << 14, 5, -1000, 1000)
    got = orc.copy_if_mod(x, 3, 0)
    assert (got == x[np.fmod(x, 3) == 0]).all()
    assert (got < 0).any()  # negative multiples of 3 are kept (C++ truncating %)


# ---------------------------------------------------------------- generators
def test_generators_pinned(wl):
    # The standard fixes mt19937: the 10000th output of seed 5489 is 4123659995
    assert int(wl.u32_uniform(10000, 5489)[-1]) == 4123659995
    # ... and mt19937_64: the 10000th output of seed 5489 is 9981545732273789042
    u = wl.unit_f64(10000, 5489)[-1]
    assert int(u * 2**53) == 9981545732273789042 >> 11
    le = wl.u32_low_entropy(1 << 16, 2)
    ones = np.unpackbits(le.view(np.uint8)).mean()
    assert abs(ones - 1 / 8) < 0.01
    f = wl.f32_normal_signed_zeros(1 << 16, 3, 4)
    bits = f.view(np.uint32)
    assert 0.007 < (bits == 0x80000000).mean() < 0.013
    assert 0.007 < (bits == 0).mean() < 0.013
    assert abs(f.mean()) < 0.02 and abs(f.std() - 1) < 0.02
    x = wl.i32_range(1 << 16, 5, -1000, 1000)
    assert x.min() == -1000 and x.max() == 1000
    sizes, total = wl.segment_sizes(4096, 7, 16, 1 << 20)
    assert sizes.min() >= 16 and sizes.max() <= 1 << 20 and total == int(sizes.sum())


def test_host_baseline_segmented_matches_oracle(orc, wl):
    """The segmented host baseline computes per-segment stable sorts and scans."""
    sizes = np.array([0, 1, 5, 1000, 37, 4096, 2], dtype=np.uint64)
    offs = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint64)
    n = int(offs[-1])
    keys = wl.i32_range(n, 3, 0, 20).astype(np.uint32)  # many ties: stability matters
    vals = np.arange(n, dtype=np.uint32)
    k2, v2 = keys.copy(), vals.copy()
    orc.L.bl_segmented_sort_pairs_u32(k2.ctypes.data, v2.ctypes.data, offs.ctypes.data, len(sizes))
    x = wl.i32_range(n, 4, -1000, 1000)
    y = np.empty_like(x)
    orc.L.bl_segmented_inclusive_scan_i32(x.ctypes.data, y.ctypes.data, offs.ctypes.data, len(sizes))
    for s in range(len(sizes)):
        b, e = int(offs[s]), int(offs[s + 1])
        perm = orc.sort_perm(keys[b:e], "uint32_t")
        assert (k2[b:e] == keys[b:e][perm]).all() and (v2[b:e] == perm + b).all()
        assert (y[b:e] == orc.inclusive_scan(x[b:e], "int32_t")).all()
