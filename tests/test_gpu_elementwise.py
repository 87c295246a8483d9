"""GPU parity at the edges of the element-wise map kernel (Copy, Transform,
Transform_Binary: 8 indices per thread in blocks of 2048), of the merge with
parallel merge-path splits (Merge / Merge_By_Key, 2048-output tiles) and of
the (index, value) pair reduction behind Min/Max_Element.  Expected values
come from oracle/algorithms.py (merge: a stable sort of the concatenation,
ties from the first input first) and numpy element-wise arithmetic; every
comparison is exact."""
import numpy as np
import pytest

from helpers import oracle_algorithms

pytestmark = pytest.mark.gpu

# around the map block (2048 indices) and the merge tile (2048 outputs)
SIZES = [1, 7, 255, 256, 2047, 2048, 2049, 6143, 100003, (1 << 22) + 7]


@pytest.fixture(scope="module")
def trtc():
    import thrustrtc_amd
    return thrustrtc_amd


@pytest.fixture(scope="module")
def oa():
    return oracle_algorithms()


def say(*a):
    print(*a, flush=True)


@pytest.mark.parametrize("n", SIZES)
def test_copy_and_transform_sizes(trtc, n):
    say(f"copy / transform n={n}")
    rng = np.random.default_rng(n)
    x = rng.integers(-2**31, 2**31 - 1, size=n, dtype=np.int32)
    d = trtc.device_vector_from_numpy(x)
    o = trtc.device_vector("int32_t", n)
    trtc.Copy(d, o)
    assert np.array_equal(o.to_host(), x)
    of = trtc.device_vector("float", n)
    trtc.Copy(d, of)
    assert np.array_equal(of.to_host(), x.astype(np.float32))
    trtc.Transform(d, o, trtc.Negate())
    assert np.array_equal(o.to_host(), (np.zeros_like(x) - x))
    y = rng.integers(-1000, 1000, size=n, dtype=np.int32)
    trtc.Transform_Binary(d, trtc.device_vector_from_numpy(y), o, trtc.Plus())
    assert np.array_equal(o.to_host(), (x + y).astype(np.int32))
    # in place: the output is the input
    trtc.Transform(d, d, trtc.Negate())
    assert np.array_equal(d.to_host(), (np.zeros_like(x) - x))


@pytest.mark.parametrize("n", [5, 4097, (1 << 20) + 3])
def test_copy_widths(trtc, n):
    say(f"copy widths n={n}")
    rng = np.random.default_rng(7 + n)
    b = rng.integers(0, 255, size=n, dtype=np.uint8)  # same-type 1-byte copy (hipMemcpyAsync)
    o8 = trtc.device_vector("uint8_t", n)
    trtc.Copy(trtc.device_vector_from_numpy(b), o8)
    assert np.array_equal(o8.to_host(), b)
    q = rng.integers(-2**40, 2**40, size=n, dtype=np.int64)
    od = trtc.device_vector("double", n)
    trtc.Copy(trtc.device_vector_from_numpy(q), od)
    assert np.array_equal(od.to_host(), q.astype(np.float64))
    o64 = trtc.device_vector("int64_t", n)
    trtc.Copy(trtc.device_vector_from_numpy(q), o64)
    assert np.array_equal(o64.to_host(), q)


# (the reference's Python layer refuses 0-size vectors, so every input holds one item at least)
# tiles: 8192 outputs for Merge (512 x 16), 2048 for Merge_By_Key (256 x 8)
@pytest.mark.parametrize("na,nb", [(1, 5), (5, 1), (1, 2048), (2047, 2049), (3000, 17), (8191, 1), (4096, 4097),
                                   (12289, 20000), (100003, 65536), ((1 << 21) + 1, (1 << 20) - 3)])
def test_merge_shapes(trtc, oa, na, nb):
    say(f"merge na={na} nb={nb}")
    rng = np.random.default_rng(na * 31 + nb)
    # few distinct keys: many ties across the two inputs and across tiles
    a = np.sort(rng.integers(-50, 50, size=na, dtype=np.int32))
    b = np.sort(rng.integers(-50, 50, size=nb, dtype=np.int32))
    va = np.arange(na, dtype=np.int32)
    vb = np.arange(na, na + nb, dtype=np.int32)
    wk, wv = oa.merge_np(a, b, False, va, vb)
    da, db = trtc.device_vector_from_numpy(a), trtc.device_vector_from_numpy(b)
    o = trtc.device_vector("int32_t", na + nb)
    trtc.Merge(da, db, o)
    assert np.array_equal(o.to_host(), wk)
    ko, vo = trtc.device_vector("int32_t", na + nb), trtc.device_vector("int32_t", na + nb)
    trtc.Merge_By_Key(da, db, trtc.device_vector_from_numpy(va), trtc.device_vector_from_numpy(vb), ko, vo)
    assert np.array_equal(ko.to_host(), wk) and np.array_equal(vo.to_host(), wv)  # payload pins tie order
    # descending under Greater
    ad, bd = a[::-1].copy(), b[::-1].copy()
    gk, gv = oa.merge_np(ad, bd, True, va, vb)
    trtc.Merge_By_Key(trtc.device_vector_from_numpy(ad), trtc.device_vector_from_numpy(bd),
                      trtc.device_vector_from_numpy(va), trtc.device_vector_from_numpy(vb), ko, vo, trtc.Greater())
    assert np.array_equal(ko.to_host(), gk) and np.array_equal(vo.to_host(), gv)


@pytest.mark.parametrize("n", [1, 2, 2049, 100003, (1 << 22) + 5])
def test_extrema_ties(trtc, n):
    say(f"extrema ties n={n}")
    rng = np.random.default_rng(11 + n)
    x = rng.integers(-3, 4, size=n, dtype=np.int32)  # the extremes occur many times
    d = trtc.device_vector_from_numpy(x)
    assert trtc.Min_Element(d) == int(np.argmin(x))  # numpy: the first occurrence
    assert trtc.Max_Element(d) == int(np.argmax(x))
    assert tuple(trtc.MinMax_Element(d)) == (int(np.argmin(x)), int(np.argmax(x)))
    assert trtc.Min_Element(d, trtc.Greater()) == int(np.argmax(x))
    f = x.astype(np.float64) * 0.5
    df = trtc.device_vector_from_numpy(f)
    assert trtc.Min_Element(df) == int(np.argmin(f)) and trtc.Max_Element(df) == int(np.argmax(f))
    allsame = trtc.device_vector_from_numpy(np.full(n, 9, dtype=np.int32))
    assert trtc.Min_Element(allsame) == 0 and trtc.Max_Element(allsame) == 0


@pytest.mark.parametrize("dtype,ctype", [("int8", "int8_t"), ("uint16", "uint16_t"), ("int64", "int64_t"),
                                         ("uint64", "uint64_t"), ("float32", "float")])
@pytest.mark.parametrize("n", [3, 17, 4099, (1 << 20) + 3])
def test_extrema_widths(trtc, dtype, ctype, n):
    """The streaming extrema kernel reads 16/sizeof(T) elements per load:
    every width, sizes with a scalar tail, few distinct values (ties)."""
    say(f"extrema widths {dtype} n={n}")
    rng = np.random.default_rng(n * 7 + len(dtype))
    x = rng.integers(0, 5, size=n).astype(dtype)
    d = trtc.device_vector_from_numpy(x)
    assert trtc.Min_Element(d) == int(np.argmin(x)) and trtc.Max_Element(d) == int(np.argmax(x))
    assert trtc.Max_Element(d, trtc.Greater()) == int(np.argmin(x))
    # the extremes once, at the very end and at the very start
    y = np.full(n, 2, dtype=dtype)
    y[-1] = 0
    y[0] = 4
    dy = trtc.device_vector_from_numpy(y)
    assert tuple(trtc.MinMax_Element(dy)) == (n - 1, 0)


@pytest.mark.parametrize("na,nb", [(3, 8190), (8193, 8191), (70001, 123457)])
def test_merge_wide_types(trtc, oa, na, nb):
    """8-byte keys on the 8192-output Merge tiles (67 KB of LDS) and 8-byte
    values staged beside the keys in Merge_By_Key."""
    say(f"merge wide na={na} nb={nb}")
    rng = np.random.default_rng(na + 3 * nb)
    a = np.sort(rng.integers(-40, 40, size=na).astype(np.float64) * 0.25)
    b = np.sort(rng.integers(-40, 40, size=nb).astype(np.float64) * 0.25)
    va = np.arange(na, dtype=np.int64) * 3
    vb = np.arange(na, na + nb, dtype=np.int64) * 3
    wk, wv = oa.merge_np(a, b, False, va, vb)
    da, db = trtc.device_vector_from_numpy(a), trtc.device_vector_from_numpy(b)
    o = trtc.device_vector("double", na + nb)
    trtc.Merge(da, db, o)
    assert np.array_equal(o.to_host(), wk)
    vo = trtc.device_vector("int64_t", na + nb)
    trtc.Merge_By_Key(da, db, trtc.device_vector_from_numpy(va), trtc.device_vector_from_numpy(vb), o, vo)
    assert np.array_equal(o.to_host(), wk) and np.array_equal(vo.to_host(), wv)


@pytest.mark.parametrize("n", [1, 16383, 16384, 16385, 3 * 16384 + 1, (1 << 22) + 3])
def test_scan_8byte_tiles(trtc, n):
    """Whole-input scans of 8-byte elements run 1024 x 16 = 16384-element
    tiles: tile edges, inclusive and exclusive, int64 and float64 holding
    integers (every partial sum exact, so the comparison is exact)."""
    say(f"8-byte scan n={n}")
    rng = np.random.default_rng(n + 99)
    x = rng.integers(-2**40, 2**40, size=n, dtype=np.int64)
    d = trtc.device_vector_from_numpy(x)
    o = trtc.device_vector("int64_t", n)
    trtc.Inclusive_Scan(d, o)
    assert np.array_equal(o.to_host(), np.cumsum(x))
    trtc.Exclusive_Scan(d, o)
    ex = np.concatenate([[0], np.cumsum(x)[:-1]]).astype(np.int64)
    assert np.array_equal(o.to_host(), ex)
    f = rng.integers(-1000, 1000, size=n).astype(np.float64)
    df = trtc.device_vector_from_numpy(f)
    of = trtc.device_vector("double", n)
    trtc.Inclusive_Scan(df, of)
    assert np.array_equal(of.to_host(), np.cumsum(f))


@pytest.mark.parametrize("n", [3, 17, 4099, (1 << 20) + 5])
def test_copy_offsets_and_short_output(trtc, n):
    """Same-type contiguous copies: 16-byte aligned ends run the 16-byte copy
    kernel (with a byte tail), ranges starting mid-vector the map kernel (4 and
    8 bytes) or hipMemcpyAsync (1 byte); an output shorter than the input is
    an error for every path, with nothing written."""
    say(f"copy offsets n={n}")
    rng = np.random.default_rng(n + 5)
    for dt, ct in [(np.int32, "int32_t"), (np.uint8, "uint8_t"), (np.float64, "double")]:
        x = rng.integers(0, 100, size=n + 1).astype(dt)
        d = trtc.device_vector_from_numpy(x)
        o = trtc.device_vector(ct, n + 1)
        trtc.Copy(trtc.DVRange(d, 1, n + 1), trtc.DVRange(o, 0, n))  # source off by one item
        assert np.array_equal(o.to_host(0, n), x[1:])
        trtc.Copy(trtc.DVRange(d, 0, n), trtc.DVRange(o, 1, n + 1))  # destination off by one item
        assert np.array_equal(o.to_host(1, n + 1), x[:n])
        keep = o.to_host()
        with pytest.raises(SystemError):
            trtc.Copy(d, trtc.DVRange(o, 0, n))
        assert np.array_equal(o.to_host(), keep)


@pytest.mark.parametrize("n", [1, 2047, 2049, 100003, (1 << 22) + 7])
def test_fill_sequence_tabulate_sizes(trtc, n):
    """Fill, Sequence (three forms) and Tabulate on the element-wise map kernel
    (8 indices per thread in blocks of 2048): sizes around the block, int8
    wrap-around and doubles."""
    say(f"fill/sequence/tabulate n={n}")
    d = trtc.device_vector("int32_t", n)
    trtc.Fill(d, trtc.DVInt32(-4))
    assert (d.to_host() == -4).all()
    trtc.Sequence(d)
    assert np.array_equal(d.to_host(), np.arange(n, dtype=np.int32))
    trtc.Sequence(d, trtc.DVInt32(5), trtc.DVInt32(-3))
    want = (5 - 3 * np.arange(n, dtype=np.int64)).astype(np.int32)
    assert np.array_equal(d.to_host(), want)
    trtc.Tabulate(d, trtc.Negate())
    assert np.array_equal(d.to_host(), (-want.astype(np.int64)).astype(np.int32))
    b = trtc.device_vector("int8_t", n)
    trtc.Sequence(b, trtc.DVInt8(100))
    assert np.array_equal(b.to_host(), (100 + np.arange(n)).astype(np.int8))  # wraps as the reference's C++ does
    f = trtc.device_vector("double", n)
    trtc.Sequence(f, trtc.DVDouble(0.5), trtc.DVDouble(0.25))
    assert np.array_equal(f.to_host(), 0.5 + np.arange(n, dtype=np.float64) * 0.25)


@pytest.mark.parametrize("n", [1, 2, 2048, 2049, 100003])
def test_replace_copy_and_adjacent_difference_sizes(trtc, oa, n):
    """Replace_Copy / Replace_Copy_If / Adjacent_Difference on the element-wise
    map kernel: sizes around the 2048-index block, the in-place difference
    (which reads a copy of the input) and a user operator."""
    say(f"replace_copy / adjacent_difference n={n}")
    rng = np.random.default_rng(n + 17)
    x = rng.integers(-5, 5, size=n, dtype=np.int32)
    d = trtc.device_vector_from_numpy(x)
    o = trtc.device_vector("int32_t", n)
    trtc.Replace_Copy(d, o, trtc.DVInt32(3), trtc.DVInt32(-100))
    assert np.array_equal(o.to_host(), np.where(x == 3, -100, x))
    neg = trtc.Functor({}, ["x"], "        return x < 0;\n")
    trtc.Replace_Copy_If(d, o, neg, trtc.DVInt32(0))
    assert np.array_equal(o.to_host(), np.where(x < 0, 0, x))
    trtc.Adjacent_Difference(d, o)
    assert np.array_equal(o.to_host(), oa.adjacent_difference_np(x))
    trtc.Adjacent_Difference(d, o, trtc.Plus())
    want = x.copy()
    want[1:] = x[1:] + x[:-1]
    assert np.array_equal(o.to_host(), want)
    trtc.Adjacent_Difference(d, d)  # in place
    assert np.array_equal(d.to_host(), oa.adjacent_difference_np(x))
