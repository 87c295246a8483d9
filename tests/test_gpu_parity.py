"""GPU parity: the HIP path (through libPyThrustRTC.so's C ABI, driven by the
reference-shaped Python API) against the oracle on the same seeded inputs.

Integer / index / byte results are compared bit-exactly; float reductions and
scans against the FP64 oracle within the tolerance stated in each test.
"""
import json
import os
import sys
import threading

import numpy as np
import pytest

from helpers import Oracle, workloads, NP

pytestmark = pytest.mark.gpu

GOLDEN = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "known_answers.json")))


@pytest.fixture(scope="module")
def trtc():
    import thrustrtc_amd
    return thrustrtc_amd


@pytest.fixture(scope="module")
def orc():
    return Oracle()


@pytest.fixture(scope="module")
def wl():
    return workloads()


def say(*a):
    print(*a, flush=True)


# ---------------------------------------------------------------- known answers
def test_known_answers(trtc):
    for c in GOLDEN["sort"]:
        d = trtc.device_vector_from_list(c["in"], c["elem"])
        trtc.Sort(d, trtc.Greater() if c["comp"] == "Greater" else trtc.Less())
        assert d.to_host().tolist() == c["out"]
    for c in GOLDEN["sort_by_key"]:
        k = trtc.device_vector_from_list(c["keys"], c["kelem"])
        v = trtc.device_vector_from_list(c["values"], c["velem"])
        trtc.Sort_By_Key(k, v, trtc.Greater() if c["comp"] == "Greater" else None)
        assert k.to_host().tolist() == c["keys_out"] and v.to_host().tolist() == c["values_out"]
    ops = {"Plus": trtc.Plus, "Maximum": trtc.Maximum}
    for c in GOLDEN["inclusive_scan"]:
        d = trtc.device_vector_from_list(c["in"], c["elem"])
        trtc.Inclusive_Scan(d, d, ops[c["op"]]())
        assert d.to_host().tolist() == c["out"]
    for c in GOLDEN["exclusive_scan"]:
        d = trtc.device_vector_from_list(c["in"], c["elem"])
        if c["init"] == 0 and c["op"] == "Plus":
            trtc.Exclusive_Scan(d, d)
        else:
            trtc.Exclusive_Scan(d, d, trtc.DVInt32(c["init"]), ops[c["op"]]())
        assert d.to_host().tolist() == c["out"]
    for c in GOLDEN["scan_by_key"]:
        k = trtc.device_vector_from_list(c["keys"], "int32_t")
        v = trtc.device_vector_from_list(c["values"], "int32_t")
        if not c["exclusive"]:
            trtc.Inclusive_Scan_By_Key(k, v, v)
        elif c["init"] == 0:
            trtc.Exclusive_Scan_By_Key(k, v, v)
        else:
            trtc.Exclusive_Scan_By_Key(k, v, v, trtc.DVInt32(c["init"]))
        assert v.to_host().tolist() == c["out"]
    for c in GOLDEN["reduce"]:
        d = trtc.device_vector_from_list(c["in"], c["elem"])
        if c["init"] is None:
            r = trtc.Reduce(d)
        elif c["op"] == "Plus":
            r = trtc.Reduce(d, trtc.DVInt32(c["init"]))
        else:
            r = trtc.Reduce(d, trtc.DVInt32(c["init"]), ops[c["op"]]())
        assert r == c["out"]
    is_even = trtc.Functor({}, ["x"], "         return x % 2 == 0;\n")
    for c in GOLDEN["copy_if"]:
        d = trtc.device_vector_from_list(c["in"], "int32_t")
        o = trtc.device_vector("int32_t", len(c["in"]))
        cnt = trtc.Copy_If(d, o, is_even)
        assert o.to_host(0, cnt).tolist() == c["out"]
    for c in GOLDEN["copy_if_stencil"]:
        d = trtc.device_vector_from_list(c["in"], "int32_t")
        s = trtc.device_vector_from_list(c["stencil"], "int32_t")
        o = trtc.device_vector("int32_t", len(c["in"]))
        cnt = trtc.Copy_If_Stencil(d, s, o, is_even)
        assert o.to_host(0, cnt).tolist() == c["out"]


# ---------------------------------------------------------------- sort
SORT_SIZES = [1, 2, 3, 255, 4095, 4096, 4097, 65537, 1 << 20]


def _keys(wl, kind, n, seed):
    if kind == "u32":
        return wl.u32_uniform(n, seed), "uint32_t"
    if kind == "u32_low":
        return wl.u32_low_entropy(n, seed), "uint32_t"
    if kind == "f32":
        return wl.f32_normal_signed_zeros(n, seed, seed + 1), "float"
    if kind == "i32":
        return wl.i32_range(n, seed, -2**31, 2**31 - 1), "int32_t"
    raise ValueError(kind)


@pytest.mark.parametrize("kind", ["u32", "u32_low", "f32", "i32"])
@pytest.mark.parametrize("n", SORT_SIZES)
@pytest.mark.parametrize("greater", [False, True])
def test_sort_by_key_parity(trtc, orc, wl, kind, n, greater):
    say("sort_by_key", kind, n, greater)
    keys, elem = _keys(wl, kind, n, 2)
    idx = np.arange(n, dtype=np.uint32)
    dk = trtc.device_vector_from_numpy(keys)
    dv = trtc.device_vector_from_numpy(idx)
    trtc.Sort_By_Key(dk, dv, trtc.Greater() if greater else trtc.Less())
    perm = orc.sort_perm(keys, elem, greater)
    gk, gv = dk.to_host(), dv.to_host()
    # keys bit-exact (-0.0 stays -0.0), payload = the oracle's stable permutation
    assert (gk.view(np.uint32) == keys[perm].view(np.uint32)).all()
    assert (gv == perm.astype(np.uint32)).all()


# Multi-million-key inputs: thousands of tiles per pass, long lookback walks,
# a partial last tile, skewed digits.
@pytest.mark.parametrize("kind", ["u32", "u32_low", "f32", "i32"])
@pytest.mark.parametrize("n", [1 << 22, (1 << 22) + 12345, 3 << 21])
@pytest.mark.parametrize("greater", [False, True])
def test_sort_by_key_large(trtc, orc, wl, kind, n, greater):
    say("sort_by_key large", kind, n, greater)
    keys, elem = _keys(wl, kind, n, 5)
    idx = np.arange(n, dtype=np.uint32)
    dk = trtc.device_vector_from_numpy(keys)
    dv = trtc.device_vector_from_numpy(idx)
    trtc.Sort_By_Key(dk, dv, trtc.Greater() if greater else trtc.Less())
    perm = orc.sort_perm(keys, elem, greater)
    gk, gv = dk.to_host(), dv.to_host()
    assert (gk.view(np.uint32) == keys[perm].view(np.uint32)).all()
    assert (gv == perm.astype(np.uint32)).all()


def test_sort_large_keys_only_and_16bit(trtc, orc, wl):
    say("sort large keys-only u32 and 16-bit keys")
    n = (1 << 22) + 777
    keys, elem = _keys(wl, "u32_low", n, 9)
    dk = trtc.device_vector_from_numpy(keys)
    trtc.Sort(dk)
    assert (dk.to_host() == orc.sort(keys, elem)).all()
    k16 = (wl.u32_uniform(n, 10) & 0xFFFF).astype(np.uint16)
    d16 = trtc.device_vector_from_numpy(k16)
    dv = trtc.device_vector_from_numpy(np.arange(n, dtype=np.uint32))
    trtc.Sort_By_Key(d16, dv)
    perm = orc.sort_perm(k16, "uint16_t")
    assert (d16.to_host() == k16[perm]).all() and (dv.to_host() == perm.astype(np.uint32)).all()


@pytest.mark.parametrize("elem", ["int8_t", "uint8_t", "int16_t", "uint16_t", "int64_t", "uint64_t", "double",
                                  "float", "int32_t", "uint32_t"])
def test_sort_types(trtc, orc, wl, elem):
    say("sort types", elem)
    n = 70001
    raw = wl.u32_uniform(2 * n, 5).view(np.uint64) if NP[elem]().itemsize == 8 else wl.u32_uniform(n, 5)
    keys = raw.view(np.uint8)[: n * NP[elem]().itemsize].view(NP[elem]).copy()
    if elem in ("float", "double"):
        keys = np.where(np.isnan(keys), 0, keys).astype(NP[elem])
        keys[::97] = -0.0
    for greater in (False, True):
        d = trtc.device_vector_from_numpy(keys)
        trtc.Sort(d, trtc.Greater() if greater else None)
        want = keys[orc.sort_perm(keys, elem, greater)]
        assert (d.to_host().view(np.uint8) == want.view(np.uint8)).all(), (elem, greater)


@pytest.mark.parametrize("velem", ["int8_t", "int16_t", "int64_t", "double"])
def test_sort_by_key_value_widths(trtc, orc, wl, velem):
    n = 30000
    keys = wl.u32_low_entropy(n, 8)
    vals = (np.arange(n) * 7 + 3).astype(NP[velem])
    dk = trtc.device_vector_from_numpy(keys)
    dv = trtc.device_vector_from_numpy(vals)
    trtc.Sort_By_Key(dk, dv)
    perm = orc.sort_perm(keys, "uint32_t")
    assert (dk.to_host() == keys[perm]).all()
    assert (dv.to_host() == vals[perm]).all()


def test_sort_values_longer_than_keys(trtc, orc, wl):
    # only the first keys.size() values are permuted (radix_sort.cpp:328-370)
    keys = wl.u32_uniform(1000, 3)
    vals = np.arange(1500, dtype=np.int32)
    dk = trtc.device_vector_from_numpy(keys)
    dv = trtc.device_vector_from_numpy(vals)
    trtc.Sort_By_Key(dk, dv)
    perm = orc.sort_perm(keys, "uint32_t")
    got = dv.to_host()
    assert (got[:1000] == perm).all() and (got[1000:] == vals[1000:]).all()


def test_sort_all_equal_and_sorted(trtc):
    for arr in [np.full(100000, 7, np.uint32), np.arange(100000, dtype=np.uint32),
                np.arange(100000, 0, -1).astype(np.uint32)]:
        dk = trtc.device_vector_from_numpy(arr)
        dv = trtc.device_vector_from_numpy(np.arange(len(arr), dtype=np.uint32))
        trtc.Sort_By_Key(dk, dv)
        perm = np.argsort(arr, kind="stable")
        assert (dk.to_host() == arr[perm]).all() and (dv.to_host() == perm).all()


def constant_digit_keys(wl, case, n, seed):
    """Keys some of whose 8-bit digits are the same in every key (the passes
    over those digits are skipped; reference radix_sort.cpp:323-325, 351-353)."""
    x = wl.u32_uniform(n, seed)
    if case == "below_2^16":        # digits 2, 3 constant: two passes run
        return x & np.uint32(0xFFFF), "uint32_t"
    if case == "middle_only":       # digits 0 and 3 constant
        return (x & np.uint32(0x00FFFF00)) | np.uint32(0x5A000077), "uint32_t"
    if case == "one_digit":         # digits 0, 2, 3 constant: one pass runs (odd: result copied back)
        return (x & np.uint32(0xFF00)) | np.uint32(0x13000000), "uint32_t"
    if case == "three_digits":      # digit 1 constant: three passes run
        return (x & np.uint32(0xFFFF00FF)) | np.uint32(0x4200), "uint32_t"
    if case == "all_equal":         # no pass runs
        return np.full(n, 0xDEADBEEF, dtype=np.uint32), "uint32_t"
    if case == "i32_small":         # signed keys in [-100, 100]: top digits 0x00 / 0xFF vary, so all run
        return wl.i32_range(n, seed, -100, 100), "int32_t"
    if case == "f32_const_sign":    # floats in [1, 2): sign/exponent byte constant
        return (np.float32(1.0) + (x >> np.uint32(9)).astype(np.float32) * np.float32(2.0 ** -23)), "float"
    if case == "u64_below_2^40":    # 8-byte keys, digits 5..7 constant
        return (x.astype(np.uint64) << np.uint64(8)) | np.uint64(0x11), "uint64_t"
    raise ValueError(case)


CONST_CASES = ["below_2^16", "middle_only", "one_digit", "three_digits", "all_equal", "i32_small",
               "f32_const_sign", "u64_below_2^40"]


@pytest.mark.parametrize("case", CONST_CASES)
@pytest.mark.parametrize("n", [100003, (1 << 20) + 3, 1 << 22])
@pytest.mark.parametrize("greater", [False, True])
def test_sort_by_key_constant_digits(trtc, orc, wl, case, n, greater):
    say(f"constant digits {case} n={n} greater={greater}")
    keys, elem = constant_digit_keys(wl, case, n, 13)
    dk = trtc.device_vector_from_numpy(keys)
    dv = trtc.device_vector_from_numpy(np.arange(n, dtype=np.uint32))
    trtc.Sort_By_Key(dk, dv, trtc.Greater() if greater else None)
    perm = orc.sort_perm(keys, elem, greater)
    assert np.array_equal(dk.to_host().view(f"u{keys.itemsize}"), keys[perm].view(f"u{keys.itemsize}"))
    assert np.array_equal(dv.to_host(), perm.astype(np.uint32))


@pytest.mark.parametrize("case", ["below_2^16", "one_digit", "all_equal"])
def test_sort_keys_only_constant_digits(trtc, orc, wl, case):
    say(f"keys-only constant digits {case}")
    keys, elem = constant_digit_keys(wl, case, (1 << 21) + 5, 17)
    dk = trtc.device_vector_from_numpy(keys)
    trtc.Sort(dk)
    assert np.array_equal(dk.to_host(), orc.sort(keys, elem))


@pytest.mark.parametrize("case", ["one_digit", "three_digits", "all_equal", "below_2^16"])
def test_long_pass_constant_digits(trtc, orc, wl, case):
    """In-place sorts on the long-tile pass (2^23 + 5 keys) whose constant
    digits leave an odd number of sorting passes (one_digit, three_digits): the
    parity copy runs in the launch of a constant digit's pass."""
    n = (1 << 23) + 5
    say(f"long-tile pass, constant digits {case} n={n}")
    keys, elem = constant_digit_keys(wl, case, n, 19)
    dk = trtc.device_vector_from_numpy(keys)
    dv = trtc.device_vector_from_numpy(np.arange(n, dtype=np.uint32))
    trtc.Sort_By_Key(dk, dv)
    perm = orc.sort_perm(keys, elem, False)
    assert np.array_equal(dk.to_host(), keys[perm])
    assert np.array_equal(dv.to_host(), perm.astype(np.uint32))
    dk2 = trtc.device_vector_from_numpy(keys)
    trtc.Sort(dk2, trtc.Greater())
    assert np.array_equal(dk2.to_host(), orc.sort(keys, elem, True))


@pytest.mark.parametrize("low_const", [True, False])
@pytest.mark.parametrize("n", [100003, (1 << 22) + 9])
def test_sort_u16_one_constant_digit(trtc, orc, wl, low_const, n):
    """2-byte keys with one constant digit: one sorting pass in place, the
    parity copy in the other digit's launch (before or after the pass)."""
    say(f"u16 keys, one constant digit (low={low_const}) n={n}")
    x = wl.u32_uniform(n, 23)
    keys = ((x & np.uint32(0xFF00)) | np.uint32(0x5A)) if low_const else ((x & np.uint32(0xFF)) | np.uint32(0xA500))
    keys = keys.astype(np.uint16)
    dk = trtc.device_vector_from_numpy(keys)
    dv = trtc.device_vector_from_numpy(np.arange(n, dtype=np.uint32))
    trtc.Sort_By_Key(dk, dv)
    perm = orc.sort_perm(keys, "uint16_t", False)
    assert np.array_equal(dk.to_host(), keys[perm])
    assert np.array_equal(dv.to_host(), perm.astype(np.uint32))


# ---------------------------------------------------------------- scan
SCAN_SIZES = [1, 2, 1000, 4095, 4096, 4097, 100003, 1 << 20]


@pytest.mark.parametrize("n", SCAN_SIZES)
def test_scan_i32_parity(trtc, orc, wl, n):
    say("scan i32", n)
    x = wl.i32_range(n, 0, -100, 100)
    d = trtc.device_vector_from_numpy(x)
    o = trtc.device_vector("int32_t", n)
    trtc.Inclusive_Scan(d, o)
    assert (o.to_host() == orc.inclusive_scan(x, "int32_t")).all()
    trtc.Exclusive_Scan(d, o, trtc.DVInt32(17))
    assert (o.to_host() == orc.exclusive_scan(x, "int32_t", 17)).all()
    trtc.Inclusive_Scan(d, o, trtc.Maximum())
    assert (o.to_host() == orc.inclusive_scan(x, "int32_t", "Maximum")).all()
    # in place
    trtc.Exclusive_Scan(d, d)
    assert (d.to_host() == orc.exclusive_scan(x, "int32_t", 0)).all()


@pytest.mark.parametrize("elem", ["int64_t", "uint32_t", "double", "float", "int8_t", "uint16_t"])
def test_scan_types(trtc, orc, wl, elem):
    n = 50001
    x = wl.i32_range(n, 4, -50, 50).astype(NP[elem])
    d = trtc.device_vector_from_numpy(x)
    o = trtc.device_vector(elem, n)
    trtc.Inclusive_Scan(d, o)
    want = orc.inclusive_scan(x, elem)
    if elem in ("float", "double"):
        # integer-valued, |partial sums| < 2^24: exact in any association
        assert (o.to_host() == want).all()
    else:
        assert (o.to_host() == want).all()
    trtc.Inclusive_Scan(d, o, trtc.Minimum())
    assert (o.to_host() == orc.inclusive_scan(x, elem, "Minimum")).all()


def test_scan_float_tolerance(trtc, wl):
    n = 1 << 20
    x = wl.f32_unit(n, 6)
    d = trtc.device_vector_from_numpy(x)
    trtc.Inclusive_Scan(d, d)
    ref = np.cumsum(x.astype(np.float64))
    got = d.to_host().astype(np.float64)
    # relative error bound for float partial sums of positive terms
    assert (np.abs(got - ref) <= 1e-5 * ref + 1e-6).all()


def test_scan_user_functor_and_type_mix(trtc, orc, wl):
    n = 20000
    x = wl.i32_range(n, 1, -100, 100).astype(np.int8)
    d = trtc.device_vector_from_numpy(x)
    o = trtc.device_vector("int32_t", n)
    # int8 input scanned in int32 (the output type), op from a code string
    op = trtc.Functor({}, ["a", "b"], "        return a + b;\n")
    trtc.Inclusive_Scan(d, o, op)
    assert (o.to_host() == np.cumsum(x.astype(np.int64)).astype(np.int32)).all()


def test_scan_non_commutative_op(trtc):
    # Pair-free non-commutative associative op: keep the first non-zero
    n = 10000
    x = np.zeros(n, dtype=np.int32)
    x[[5000, 7000]] = [3, 9]
    d = trtc.device_vector_from_numpy(x)
    op = trtc.Functor({}, ["a", "b"], "        return a != 0 ? a : b;\n")
    trtc.Inclusive_Scan(d, d, op)
    want = np.zeros(n, dtype=np.int32)
    want[5000:] = 3
    assert (d.to_host() == want).all()


# ---------------------------------------------------------------- reduce
@pytest.mark.parametrize("n", [1, 2, 1023, 1024, 1025, 100003, 1 << 22])
def test_reduce_parity(trtc, orc, wl, n):
    say("reduce", n)
    x = wl.i32_range(n, 3, -1000, 1000)
    d = trtc.device_vector_from_numpy(x)
    assert trtc.Reduce(d) == orc.reduce(x, "int32_t")
    assert trtc.Reduce(d, trtc.DVInt32(5)) == orc.reduce(x, "int32_t", 5)
    assert trtc.Reduce(d, trtc.DVInt32(-5000), trtc.Maximum()) == orc.reduce(x, "int32_t", -5000, "Maximum")
    f = wl.f32_unit(n, 6)
    df = trtc.device_vector_from_numpy(f)
    r = trtc.Reduce(df)
    ref = orc.sum_f32_fp64(f)
    assert abs(r - ref) <= 1e-5 * abs(ref) + 1e-6  # stated tolerance: rel 1e-5 vs FP64


def test_reduce_deterministic_and_user_op(trtc, wl):
    f = wl.f32_normal_signed_zeros(1 << 20, 9, 10)
    d = trtc.device_vector_from_numpy(f)
    r1, r2 = trtc.Reduce(d), trtc.Reduce(d)
    assert np.float32(r1).view(np.uint32) == np.float32(r2).view(np.uint32)
    op = trtc.Functor({}, ["a", "b"], "        return a > b ? a : b;\n")
    assert trtc.Reduce(d, trtc.DVFloat(-1e30), op) == f.max()


# ---------------------------------------------------------------- copy_if
@pytest.mark.parametrize("n", [1, 7, 4095, 4096, 4097, 100003, 1 << 21])
def test_copy_if_parity(trtc, orc, wl, n):
    say("copy_if", n)
    x = wl.i32_range(n, 5, -1000, 1000)
    d = trtc.device_vector_from_numpy(x)
    o = trtc.device_vector_from_numpy(np.full(n, -7777, dtype=np.int32))
    pred = trtc.Functor({}, ["x"], "        return x % 3 == 0;\n")
    cnt = trtc.Copy_If(d, o, pred)
    want = orc.copy_if_mod(x, 3, 0)
    assert cnt == len(want)
    got = o.to_host()
    assert (got[:cnt] == want).all()
    assert (got[cnt:] == -7777).all()  # untouched past the count


def test_copy_if_captured_vector(trtc, wl):
    # a predicate capturing a device vector and a scalar (functor with members)
    n = 50000
    x = wl.i32_range(n, 8, 0, 99)
    lut = np.zeros(100, dtype=np.int32)
    lut[[3, 50, 77]] = 1
    d = trtc.device_vector_from_numpy(x)
    dl = trtc.device_vector_from_numpy(lut)
    o = trtc.device_vector("int32_t", n)
    pred = trtc.Functor({"lut": dl, "bias": trtc.DVInt32(0)}, ["x"], "        return lut[x] + bias != 0;\n")
    cnt = trtc.Copy_If(d, o, pred)
    want = x[lut[x] != 0]
    assert cnt == len(want) and (o.to_host(0, cnt) == want).all()


# ---------------------------------------------------------------- scan_by_key
@pytest.mark.parametrize("n", [1, 5, 4096, 4097, 100003, 1 << 20])
def test_scan_by_key_parity(trtc, orc, wl, n):
    say("scan_by_key", n)
    sizes, _ = wl.segment_sizes(n, 11, 1, 5000)
    keys = np.repeat(np.arange(len(sizes), dtype=np.int32), sizes.astype(np.int64))[:n]
    vals = wl.i32_range(n, 12, -1000, 1000)
    dk = trtc.device_vector_from_numpy(keys)
    dv = trtc.device_vector_from_numpy(vals)
    o = trtc.device_vector("int32_t", n)
    trtc.Inclusive_Scan_By_Key(dk, dv, o)
    assert (o.to_host() == orc.scan_by_key(keys, vals, "int32_t", "int32_t")).all()
    trtc.Exclusive_Scan_By_Key(dk, dv, o, trtc.DVInt32(3))
    assert (o.to_host() == orc.scan_by_key(keys, vals, "int32_t", "int32_t", True, 3)).all()
    trtc.Exclusive_Scan_By_Key(dk, dv, dv)  # in place
    assert (dv.to_host() == orc.scan_by_key(keys, vals, "int32_t", "int32_t", True, 0)).all()


# 8-byte keys or values run 1024 x 16 = 16384-item tiles: edges of those tiles
@pytest.mark.parametrize("n", [16383, 16384, 16385, 3 * 16384 + 7, 200003])
@pytest.mark.parametrize("kelem,velem", [("int64_t", "int64_t"), ("int32_t", "double"), ("int64_t", "int32_t")])
def test_scan_by_key_8byte_tiles(trtc, orc, wl, n, kelem, velem):
    say("scan_by_key 8-byte", kelem, velem, n)
    sizes, _ = wl.segment_sizes(n, 13, 1, 3000)
    keys = np.repeat(np.arange(len(sizes)), sizes.astype(np.int64))[:n].astype(NP_OF[kelem])
    vals = wl.i32_range(n, 14, -1000, 1000).astype(NP_OF[velem])  # integers: exact in double too
    dk = trtc.device_vector_from_numpy(keys)
    dv = trtc.device_vector_from_numpy(vals)
    o = trtc.device_vector(velem, n)
    trtc.Inclusive_Scan_By_Key(dk, dv, o)
    assert (o.to_host() == orc.scan_by_key(keys, vals, kelem, velem)).all()
    trtc.Exclusive_Scan_By_Key(dk, dv, o)
    assert (o.to_host() == orc.scan_by_key(keys, vals, kelem, velem, True, 0)).all()


NP_OF = {"int32_t": np.int32, "int64_t": np.int64, "double": np.float64}


def test_scan_by_key_float_keys_and_max(trtc, orc):
    keys = np.array([0.0, -0.0, 1.0, 1.0, 2.0, -0.0] * 3000, dtype=np.float32)
    vals = np.arange(len(keys), dtype=np.int32) % 13
    dk = trtc.device_vector_from_numpy(keys)
    dv = trtc.device_vector_from_numpy(vals)
    o = trtc.device_vector("int32_t", len(keys))
    trtc.Inclusive_Scan_By_Key(dk, dv, o, trtc.EqualTo(), trtc.Maximum())
    assert (o.to_host() == orc.scan_by_key(keys, vals, "float", "int32_t", op="Maximum")).all()


# ---------------------------------------------------------------- user kernels / context
def test_kernel_and_for(trtc):
    harr = np.array([1.0, 2.0, 3.0, 4.0, 5.0], dtype=np.float32)
    darr = trtc.device_vector_from_numpy(harr)
    kernel = trtc.Kernel(["arr_in", "arr_out", "k"], """
    size_t idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= arr_in.size()) return;
    arr_out[idx] = arr_in[idx]*k;
    """)
    out = trtc.device_vector("float", 5)
    kernel.launch(1, 128, [darr, out, trtc.DVFloat(10.0)])
    assert out.to_host().tolist() == [10, 20, 30, 40, 50]
    darr2 = trtc.device_vector_from_list([6, 7, 8, 9, 10], "int32_t")
    out2 = trtc.device_vector("int32_t", 5)
    kernel.launch(1, 128, [darr2, out2, trtc.DVInt32(5)])
    assert out2.to_host().tolist() == [30, 35, 40, 45, 50]
    loop = trtc.For(["arr_in", "arr_out", "k"], "idx", "    arr_out[idx] = arr_in[idx]*k;\n")
    loop.launch_n(5, [darr, out, trtc.DVFloat(2.0)])
    assert out.to_host().tolist() == [2, 4, 6, 8, 10]
    assert trtc.Get_PTX_Arch() == 950


def test_multithread_for(trtc):
    # python/test/test_multithread.py: one For launched from three threads
    darr = trtc.device_vector_from_numpy(np.array([1.0, 2.0, 3.0, 4.0, 5.0], dtype=np.float32))
    loop = trtc.For(["arr_in", "arr_out", "k"], "idx", "    arr_out[idx] = arr_in[idx]*k;\n")
    results, errors = [], []

    def work():
        try:
            o = trtc.device_vector("float", 5)
            loop.launch_n(5, [darr, o, trtc.DVFloat(10.0)])
            results.append(o.to_host().tolist())
        except Exception as e:  # pragma: no cover
            errors.append(e)

    ts = [threading.Thread(target=work) for _ in range(3)]
    [t.start() for t in ts]
    [t.join() for t in ts]
    assert not errors and results == [[10, 20, 30, 40, 50]] * 3


def test_bad_functor_is_an_error_not_a_crash(trtc):
    d = trtc.device_vector_from_list([1, 2, 3], "int32_t")
    o = trtc.device_vector("int32_t", 3)
    bad = trtc.Functor({}, ["x"], "        return x %% ;\n")
    with pytest.raises(SystemError):
        trtc.Copy_If(d, o, bad)
    good = trtc.Functor({}, ["x"], "        return x > 1;\n")
    assert trtc.Copy_If(d, o, good) == 2


# ---------------------------------------------------------------- comparison sort
MSORT_SIZES = [1, 2, 100, 2047, 2048, 2049, 10000, 100003, 1 << 20]


@pytest.mark.parametrize("n", MSORT_SIZES)
def test_sort_user_comparator(trtc, wl, n):
    """A user comparator takes the stable merge path: low byte of the key only,
    so most keys tie and stability decides the order."""
    say("merge sort low-byte comparator", n)
    keys = wl.u32_uniform(n, 13)
    comp = trtc.Functor({}, ["x", "y"], "        return (x & 255) < (y & 255);\n")
    d = trtc.device_vector_from_numpy(keys)
    trtc.Sort(d, comp)
    exp = keys[np.argsort(keys & 255, kind="stable")]
    assert (d.to_host() == exp).all()


@pytest.mark.parametrize("n", [5, 2049, 100003, 1 << 20])
def test_sort_by_key_user_comparator_values(trtc, orc, wl, n):
    """Descending by a user functor (not the built-in Greater) with int32 keys,
    and float keys by magnitude with int16 values (payload gathered)."""
    say("merge sort_by_key user comparators", n)
    keys = wl.i32_range(n, 14, -500, 500)
    idx = np.arange(n, dtype=np.uint32)
    dk = trtc.device_vector_from_numpy(keys)
    dv = trtc.device_vector_from_numpy(idx)
    trtc.Sort_By_Key(dk, dv, trtc.Functor({}, ["x", "y"], "        return x > y;\n"))
    perm = orc.sort_perm(keys, "int32_t", True)
    assert (dk.to_host() == keys[perm]).all() and (dv.to_host() == perm.astype(np.uint32)).all()
    fk = wl.f32_normal_signed_zeros(n, 15, 16)
    vals = (np.arange(n) % 30000).astype(np.int16)
    dfk = trtc.device_vector_from_numpy(fk)
    dvv = trtc.device_vector_from_numpy(vals)
    trtc.Sort_By_Key(dfk, dvv, trtc.Functor({}, ["x", "y"],
                                            "        return (x < 0 ? -x : x) < (y < 0 ? -y : y);\n"))
    p2 = np.argsort(np.abs(fk), kind="stable")
    assert (dfk.to_host().view(np.uint32) == fk[p2].view(np.uint32)).all()
    assert (dvv.to_host() == vals[p2]).all()


def test_sort_user_comparator_known_answer(trtc):
    # docs/QuickStartGuide.md (Sort with a comparator): {1,4,2,8,5,7} descending
    d = trtc.device_vector_from_list([1, 4, 2, 8, 5, 7], "int32_t")
    trtc.Sort(d, trtc.Functor({}, ["x", "y"], "        return x > y;\n"))
    assert d.to_host().tolist() == [8, 7, 5, 4, 2, 1]


# ---------------------------------------------------------------- For
def test_for_per_index_semantics(trtc):
    """The For kernel takes several indices per thread, but the body keeps
    the semantics of one index per thread (impl_context.inl:851-898): a
    `return` ends that index only, a write to a parameter does not carry over
    to the next index, begin / end bound every index."""
    n = 100003
    out = trtc.device_vector_from_numpy(np.full(n + 10, -1, dtype=np.int64))
    loop = trtc.For(["out", "k"], "i", """
    if (i % 3 == 0) return;
    k = k + (long long)i;
    out[i] = k;
""")
    loop.launch(5, n, [out, trtc.DVInt64(1000)])
    got = out.to_host()
    idx = np.arange(n + 10)
    exp = np.where((idx >= 5) & (idx < n) & (idx % 3 != 0), 1000 + idx, -1)
    assert np.array_equal(got, exp)
    # the same For again (the resolved kernel is reused) and after a code block
    loop.launch(0, 7, [out, trtc.DVInt64(0)])
    assert out.to_host()[:7].tolist() == [-1, 1, 2, -1, 4, 5, -1]
    trtc.Add_Code_Block("__device__ inline long long trtc_test_twice(long long x) { return 2 * x; }\n")
    loop2 = trtc.For(["out"], "i", "    out[i] = trtc_test_twice((long long)i);\n")
    loop2.launch_n(9, [out])
    assert out.to_host()[:9].tolist() == [2 * i for i in range(9)]


def test_replace_copy_reads_each_element_once(trtc):
    """Replace_Copy / Replace_Copy_If evaluate a computed input once per
    element (replace.cpp:27-28, 39-40): a custom vector that counts its
    evaluations is read n times."""
    n = 70001
    cnt = trtc.device_vector_from_numpy(np.zeros(1, dtype=np.uint64))
    src = trtc.DVCustomVector({"cnt": cnt}, "idx", """
        atomicAdd((unsigned long long*)&cnt[0], 1ull);
        return (int)(idx % 7);
""", "int32_t", n)
    out = trtc.device_vector("int32_t", n)
    trtc.Replace_Copy(src, out, trtc.DVInt32(3), trtc.DVInt32(-1))
    exp = np.arange(n) % 7
    assert np.array_equal(out.to_host(), np.where(exp == 3, -1, exp))
    assert int(cnt.to_host()[0]) == n
    odd = trtc.Functor({}, ["x"], "        return x % 2 == 1;\n")
    trtc.Replace_Copy_If(src, out, odd, trtc.DVInt32(100))
    assert np.array_equal(out.to_host(), np.where(exp % 2 == 1, 100, exp))
    assert int(cnt.to_host()[0]) == 2 * n
