"""n_inclusive_scan_noalloc (the allocation-free Inclusive_Scan) against the
allocating n_inclusive_scan through the C ABI on two calls in a row, the
second on a smaller input, for 4- and 8-byte, integer and float elements:
integer results bitwise equal to each other and to the oracle, float results
both within 1e-5 of the FP64 running sum; Plus by default and a built-in
Maximum; a user functor only after an allocating call compiled it; inputs
beyond the fixed scratch rejected before anything runs."""
import numpy as np
import pytest

from helpers import Oracle, workloads

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def env():
    import thrustrtc_amd as trtc
    from thrustrtc_amd.native import native
    return trtc, native, Oracle(), workloads()


def say(*a):
    print(*a, flush=True)


def data(wl, elem, n, seed):
    if elem == "float":
        return wl.f32_unit(n, seed)
    x = wl.i32_range(n, seed, -1000, 1000)
    return x.astype({"int32_t": np.int32, "uint32_t": np.uint32, "int64_t": np.int64, "double": np.float64}[elem])


@pytest.mark.parametrize("elem", ["int32_t", "uint32_t", "int64_t", "float", "double"])
def test_noalloc_equals_allocating_scan(env, elem):
    trtc, native, orc, wl = env
    for n in (5_000_003, 70_001):  # the second call is smaller: state left by the first must not leak
        say(f"noalloc scan {elem} n={n}")
        x = data(wl, elem, n, 5)
        din = trtc.device_vector_from_numpy(x)
        a = trtc.device_vector(elem, n)
        b = trtc.device_vector_from_numpy(np.full(n, 7, dtype=x.dtype))
        assert native.n_inclusive_scan(din.m_cptr, a.m_cptr, None) == 0
        assert native.n_inclusive_scan_noalloc(din.m_cptr, b.m_cptr, None) == 0
        ha, hb = a.to_host(), b.to_host()
        if elem in ("int32_t", "uint32_t", "int64_t"):
            # integer scans are exact: bitwise equal to each other and the oracle
            assert np.array_equal(ha, hb)
            assert np.array_equal(hb, orc.inclusive_scan(x, elem))
        else:
            # float scans fold the lookback's predecessors in an order that
            # depends on timing, so two runs agree only to rounding: both are
            # checked against the FP64 running sum (relative 1e-5)
            ref = np.cumsum(x.astype(np.float64))
            for got in (ha, hb):
                assert bool(np.all(np.abs(got - ref) <= 1e-5 * np.abs(ref) + 1e-6))


def test_noalloc_maximum_and_in_place(env):
    trtc, native, orc, wl = env
    for n in (3_000_017, 1_000):
        say(f"noalloc scan Maximum in place n={n}")
        x = wl.i32_range(n, 9, -100000, 100000)
        d = trtc.device_vector_from_numpy(x)
        mx = trtc.Maximum()
        assert native.n_inclusive_scan_noalloc(d.m_cptr, d.m_cptr, mx.m_cptr) == 0
        assert np.array_equal(d.to_host(), orc.inclusive_scan(x, "int32_t", "Maximum"))


def test_noalloc_user_functor_needs_an_earlier_compile(env):
    trtc, native, orc, wl = env
    say("noalloc scan with a user functor")
    x = wl.i32_range(200_000, 3, -50, 50)
    d = trtc.device_vector_from_numpy(x)
    o1 = trtc.device_vector("int32_t", len(x))
    o2 = trtc.device_vector("int32_t", len(x))
    op = trtc.Functor({}, ["a", "b"], "        return a + b + (a > 1000000 ? 1 : 0);\n")
    assert native.n_inclusive_scan_noalloc(d.m_cptr, o2.m_cptr, op.m_cptr) == -1  # nothing compiled yet
    assert native.n_inclusive_scan(d.m_cptr, o1.m_cptr, op.m_cptr) == 0
    assert native.n_inclusive_scan_noalloc(d.m_cptr, o2.m_cptr, op.m_cptr) == 0
    assert np.array_equal(o1.to_host(), o2.to_host())


def test_noalloc_size_limit(env):
    trtc, native, orc, wl = env
    tile = 16384
    # 8-byte elements: at most 32767 tiles fit the 1 MiB of lookback state
    say("noalloc scan int64 at its maximum size and one past it")
    n = 32767 * tile
    d = trtc.device_vector_from_numpy(np.ones(n + 1, dtype=np.int64))
    big = d
    fit = trtc.DVVectorAdaptor("int64_t", n, native.n_dvvectorlike_data(d.m_cptr))
    o = trtc.device_vector("int64_t", n + 1)
    assert native.n_inclusive_scan_noalloc(big.m_cptr, o.m_cptr, None) == -1
    assert native.n_inclusive_scan_noalloc(fit.m_cptr, o.m_cptr, None) == 0
    h = o.to_host()
    assert h[0] == 1 and h[n - 1] == n and bool(np.array_equal(h[:n], np.arange(1, n + 1)))
