"""GPU parity of the tile kernels on contiguous views that start past a
16-byte boundary: a DVVectorAdaptor 1..3 elements into a device allocation
(as segments of a batch, request inputs of the server and caller-owned
buffers at an offset are).  Full tiles of such views are loaded as the
16-byte pieces that cover them, shifted into place in LDS; partial tiles
element by element.  Every result is compared with the oracle, bit-exactly.
"""
import numpy as np
import pytest

from helpers import NP, Oracle, workloads

pytestmark = pytest.mark.gpu

# three full 16384-element tiles and a partial one; a size with many tiles
SIZES = [3 * 16384 + 5, (1 << 20) + 3]


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


def shifted(trtc, x, elem, shift, sentinel=0):
    """x as a DVVectorAdaptor starting `shift` elements into its allocation
    (16-byte aligned), with sentinel elements before and after it."""
    from thrustrtc_amd.native import native
    dt = NP[elem]
    buf = np.concatenate([np.full(shift, sentinel, dtype=dt), np.asarray(x, dtype=dt),
                          np.full(4, sentinel, dtype=dt)])
    base = trtc.device_vector_from_numpy(buf)
    ptr = native.n_dvvectorlike_data(base.m_cptr)
    assert ptr % 16 == 0, "device allocations are 16-byte aligned"
    view = trtc.DVVectorAdaptor(elem, len(x), ptr + shift * dt().itemsize, owner=base)
    view.base, view.shift = base, shift
    return view, base


def host(view, begin=0, end=None):
    """Elements [begin, end) of an adaptor made by shifted(), read through
    the allocation that owns them (adaptors have no to_host)."""
    n = view.size()
    end = n if end is None else end
    return view.base.to_host(view.shift + begin, view.shift + end)


@pytest.mark.parametrize("elem,shift", [("int32_t", 1), ("int32_t", 2), ("int32_t", 3), ("uint32_t", 1),
                                        ("float", 3), ("int64_t", 1), ("double", 1)])
@pytest.mark.parametrize("n", SIZES)
def test_scan_shifted(trtc, orc, wl, elem, shift, n):
    say(f"shifted scan {elem} shift={shift} n={n}")
    lo, hi = (-1, 1) if elem == "float" else (-1000, 1000)  # float prefixes stay exact integers
    if elem == "uint32_t":
        lo, hi = 0, 1000
    x = wl.i32_range(n, 61, lo, hi).astype(NP[elem])
    dx, bx = shifted(trtc, x, elem, shift)
    dy, by = shifted(trtc, np.zeros(n, NP[elem]), elem, (shift + 1) % 4, sentinel=7)
    trtc.Inclusive_Scan(dx, dy)
    assert np.array_equal(host(dy), orc.inclusive_scan(x, elem))
    trtc.Exclusive_Scan(dx, dy)
    assert np.array_equal(host(dy), orc.exclusive_scan(x, elem, 0))
    # neither the input nor the output's neighbours change
    hb = by.to_host()
    s2 = (shift + 1) % 4
    assert (hb[:s2] == 7).all() and (hb[s2 + n:] == 7).all()
    assert np.array_equal(bx.to_host()[shift:shift + n], x)
    # in place
    trtc.Inclusive_Scan(dx, dx)
    assert np.array_equal(host(dx), orc.inclusive_scan(x, elem))
    hb = bx.to_host()
    assert (hb[:shift] == 0).all() and (hb[shift + n:] == 0).all()


@pytest.mark.parametrize("shift,sshift", [(1, 0), (2, 3), (3, 1)])
@pytest.mark.parametrize("n", SIZES)
def test_copy_if_shifted(trtc, orc, wl, shift, sshift, n):
    say(f"shifted copy_if shift={shift} stencil shift={sshift} n={n}")
    x = wl.i32_range(n, 62, -1000, 1000)
    st = wl.i32_range(n, 63, 0, 3)
    dx, _ = shifted(trtc, x, "int32_t", shift)
    ds, _ = shifted(trtc, st, "int32_t", sshift)
    out, _ = shifted(trtc, np.zeros(n, np.int32), "int32_t", 2)
    mod3 = trtc.Functor({}, ["x"], "        return x % 3 == 0;\n")
    want = orc.select(x, "int32_t", ("mod", 3, 0))
    c = trtc.Copy_If(dx, out, mod3)
    assert c == len(want) and np.array_equal(host(out, 0, c), want)
    want = orc.select(x, "int32_t", ("nonzero",), stencil=st)
    c = trtc.Copy_If_Stencil(dx, ds, out, trtc.Identity())
    assert c == len(want) and np.array_equal(host(out, 0, c), want)


@pytest.mark.parametrize("kshift,vshift", [(1, 2), (3, 0), (2, 2)])
@pytest.mark.parametrize("n", SIZES)
def test_by_key_and_selections_shifted(trtc, orc, wl, kshift, vshift, n):
    say(f"shifted by-key kernels key shift={kshift} value shift={vshift} n={n}")
    sizes, _ = wl.segment_sizes(n, 64, 1, 40)
    k = int(np.searchsorted(np.cumsum(sizes.astype(np.int64)), n)) + 1
    keys = np.repeat(np.arange(k, dtype=np.int32), sizes[:k].astype(np.int64))[:n]
    vals = wl.i32_range(n, 65, -1000, 1000)
    dk, _ = shifted(trtc, keys, "int32_t", kshift)
    dv, _ = shifted(trtc, vals, "int32_t", vshift)
    o1, _ = shifted(trtc, np.zeros(n, np.int32), "int32_t", 1)
    o2, _ = shifted(trtc, np.zeros(n, np.int32), "int32_t", 3)
    trtc.Inclusive_Scan_By_Key(dk, dv, o1)
    assert np.array_equal(host(o1), orc.scan_by_key(keys, vals, "int32_t", "int32_t"))
    c = trtc.Reduce_By_Key(dk, dv, o1, o2)
    wk, wv = orc.reduce_by_key(keys, vals, "int32_t", "int32_t")
    assert c == len(wk) and np.array_equal(host(o1, 0, c), wk) and np.array_equal(host(o2, 0, c), wv)
    c = trtc.Unique_Copy(dk, o1)
    (wu,) = orc.unique(keys, "int32_t")
    assert c == len(wu) and np.array_equal(host(o1, 0, c), wu)
    p3 = trtc.Functor({}, ["x"], "        return x % 3 == 1;\n")
    c = trtc.Partition_Copy(dv, o1, o2, p3)
    wt, wf = orc.partition(vals, "int32_t", ("mod", 3, 1))
    assert c == len(wt) and np.array_equal(host(o1, 0, c), wt) and np.array_equal(host(o2, 0, n - c), wf)
