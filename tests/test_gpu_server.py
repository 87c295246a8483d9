"""GPU tests of the request server: results of mixed requests against the
oracle, every result identical alone and in a mix (floats bit-identical
too: a request's reduction order depends only on its length), the per-call
mode against the batched mode, HBM-budget admission, and a short load replay.
"""
import numpy as np
import pytest

from helpers import Oracle, workloads

pytestmark = pytest.mark.gpu


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


ARENA = 1 << 22


@pytest.fixture(scope="module")
def arenas(trtc, wl):
    """Device-resident inputs: requests read windows of these."""
    h = {"keys": wl.u32_uniform(ARENA, 2), "vals": np.arange(ARENA, dtype=np.uint32),
         "x": wl.i32_range(ARENA, 5, -1000, 1000), "f": wl.f32_unit(ARENA, 6),
         "fk": wl.f32_normal_signed_zeros(ARENA, 3, 4)}
    d = {k: trtc.device_vector_from_numpy(v) for k, v in h.items()}
    ptr = {k: trtc.native.n_dvvectorlike_data(v.m_cptr) for k, v in d.items()}
    return h, d, ptr


def make_requests(rng, count, sizes):
    reqs = []
    for i in range(count):
        n = int(sizes[i % len(sizes)])
        off = int(rng.integers(0, ARENA - n + 1))
        op = int(rng.choice([0, 0, 0, 1, 2, 2, 3]))
        reqs.append((op, n, off))
    return reqs


def submit(trtc, srv, arenas, req):
    h, d, ptr = arenas
    op, n, off = req
    if op == 0:
        return srv.submit(trtc.REQ_SORT_BY_KEY, "uint32_t", ptr["keys"] + 4 * off, n, "uint32_t", ptr["vals"] + 4 * off)
    if op == 1:
        return srv.submit(trtc.REQ_SORT, "float", ptr["fk"] + 4 * off, n)
    if op == 2:
        return srv.submit(trtc.REQ_INCLUSIVE_SCAN, "int32_t", ptr["x"] + 4 * off, n)
    return srv.submit(trtc.REQ_REDUCE, "float", ptr["f"] + 4 * off, n)


def read(trtc, srv, t, req):
    op, n, off = req
    def get(which, elem, cnt):
        a = trtc.DVVectorAdaptor(elem, cnt, srv.output(t, which))
        dst = trtc.device_vector(elem, cnt)
        trtc.Copy(a, dst)
        return dst.to_host()
    if op == 0:
        return get(0, "uint32_t", n), get(1, "uint32_t", n)
    if op == 1:
        return (get(0, "float", n),)
    if op == 2:
        return (get(0, "int32_t", n),)
    return (get(0, "float", 1),)


def check_oracle(orc, arenas, req, out):
    h = arenas[0]
    op, n, off = req
    if op == 0:
        k = h["keys"][off:off + n]
        perm = orc.sort_perm(k, "uint32_t")
        assert (out[0] == k[perm]).all() and (out[1] == (perm + off).astype(np.uint32)).all()
    elif op == 1:
        k = h["fk"][off:off + n]
        assert (out[0].view(np.uint32) == orc.sort(k, "float").view(np.uint32)).all()
    elif op == 2:
        assert (out[0] == orc.inclusive_scan(h["x"][off:off + n], "int32_t")).all()
    else:
        ref = orc.sum_f32_fp64(h["f"][off:off + n])
        assert abs(float(out[0][0]) - ref) <= 1e-5 * abs(ref)


SIZES = [1, 2, 100, 1024, 5000, 8192, 8193, 20000, 100003, 1 << 20]


def run_all(trtc, srv, arenas, reqs):
    ts = [submit(trtc, srv, arenas, r) for r in reqs]
    outs = []
    for t, r in zip(ts, reqs):
        srv.wait(t)
        outs.append(read(trtc, srv, t, r))
        assert srv.latency_us(t) > 0
        srv.release(t)
    return outs


def test_server_mixed_requests_match_oracle(trtc, orc, arenas):
    say("server: 60 mixed requests vs oracle")
    rng = np.random.default_rng(1)
    reqs = make_requests(rng, 60, SIZES)
    srv = trtc.Server(64 << 30, 3, True)
    outs = run_all(trtc, srv, arenas, reqs)
    for r, o in zip(reqs, outs):
        check_oracle(orc, arenas, r, o)
    st = srv.stats()
    assert st["served"] == 60 and st["failed"] == 0


def test_server_alone_equals_in_mix(trtc, arenas):
    say("server: alone == in mix")
    rng = np.random.default_rng(2)
    probes = [(0, 100003, 17), (1, 50000, 999), (2, 1 << 20, 5), (3, 1 << 20, 12345), (3, 777, 3), (0, 4000, 0)]
    srv = trtc.Server(64 << 30, 3, True)
    alone = [run_all(trtc, srv, arenas, [p])[0] for p in probes]
    mix = make_requests(rng, 200, SIZES)
    pos = sorted(rng.choice(len(mix) + len(probes), len(probes), replace=False).tolist())
    allreq = list(mix)
    for p, r in zip(pos, probes):
        allreq.insert(p, r)
    outs = run_all(trtc, srv, arenas, allreq)
    for p, r, a in zip(pos, probes, alone):
        got = outs[p]
        assert all(x.tobytes() == y.tobytes() for x, y in zip(got, a)), r
    assert srv.stats()["mean_batch"] > 1.0


def test_server_per_call_mode_matches_batched(trtc, arenas):
    say("server: per-call mode == batched mode")
    rng = np.random.default_rng(3)
    reqs = make_requests(rng, 40, SIZES)
    a = run_all(trtc, trtc.Server(64 << 30, 3, True), arenas, reqs)
    b = run_all(trtc, trtc.Server(64 << 30, 1, False), arenas, reqs)
    for r, x, y in zip(reqs, a, b):
        if r[0] == 3:  # reductions: same tolerance, different fixed orders
            assert abs(float(x[0][0]) - float(y[0][0])) <= 2e-5 * abs(float(y[0][0]))
        else:
            assert all(u.tobytes() == v.tobytes() for u, v in zip(x, y)), r


@pytest.mark.parametrize("n", [100003, 1 << 20, (1 << 23) + 5])
def test_server_per_call_constant_digits(trtc, orc, wl, n):
    """Per-call sorts write a separate output (out of place): with every digit
    constant the source is copied in pass 0's launch; with one or three
    sorting passes the first pass reads the source and the last writes the
    output."""
    say(f"server per-call: constant-digit keys, n = {n}")
    x = wl.u32_uniform(n, 29)
    cases = {"all_equal": np.full(n, 0xDEADBEEF, dtype=np.uint32),
             "one_digit": (x & np.uint32(0xFF00)) | np.uint32(0x13000000),
             "three_digits": (x & np.uint32(0xFFFF00FF)) | np.uint32(0x4200)}
    srv = trtc.Server(64 << 30, 1, False)
    vals = np.arange(n, dtype=np.uint32)
    dvals = trtc.device_vector_from_numpy(vals)
    for name, keys in cases.items():
        dk = trtc.device_vector_from_numpy(keys)
        pk = trtc.native.n_dvvectorlike_data(dk.m_cptr)
        pv = trtc.native.n_dvvectorlike_data(dvals.m_cptr)
        t = srv.submit(trtc.REQ_SORT_BY_KEY, "uint32_t", pk, n, "uint32_t", pv)
        t2 = srv.submit(trtc.REQ_SORT, "uint32_t", pk, n)
        srv.wait(t)
        srv.wait(t2)
        got = []
        for tt, which in ((t, 0), (t, 1), (t2, 0)):
            a = trtc.DVVectorAdaptor("uint32_t", n, srv.output(tt, which))
            dst = trtc.device_vector("uint32_t", n)
            trtc.Copy(a, dst)
            got.append(dst.to_host())
        srv.release(t)
        srv.release(t2)
        perm = orc.sort_perm(keys, "uint32_t")
        assert np.array_equal(got[0], keys[perm]), name
        assert np.array_equal(got[1], perm.astype(np.uint32)), name
        assert np.array_equal(got[2], keys[perm]), name
        # the input is left as it was
        assert np.array_equal(dk.to_host(), keys), name


def test_server_batched_source_offsets_and_empty_tiles(trtc, orc, arenas):
    """Batched sorts whose sources start at every offset from a 128-byte line
    (pass 0 gives up that offset in tile 0, later passes the output's), with
    sizes that leave the last tile of some passes empty (tile counts allow for
    the largest offset): 16374 / 40957 keys on 8192-key tiles, 294907 /
    294949 keys on 32768-key tiles, and a one-tile 4096."""
    say("server: batched sorts, source line offsets 0..31, empty last tiles")
    sizes = [16374, 16385, 40957, 294907, 294949, 4096]
    reqs = []
    for off in range(32):
        reqs.append((0, sizes[off % len(sizes)], 1000 + off))
    srv = trtc.Server(64 << 30, 3, True)
    outs = run_all(trtc, srv, arenas, reqs)
    for r, o in zip(reqs, outs):
        check_oracle(orc, arenas, r, o)
    assert srv.stats()["failed"] == 0


def test_server_small_budget_admission(trtc, orc, arenas):
    say("server: 32 MiB budget, 80 requests")
    rng = np.random.default_rng(4)
    reqs = make_requests(rng, 80, [1000, 30000, 200000, 1 << 20])
    budget = 32 << 20
    srv = trtc.Server(budget, 3, True)
    outs = run_all(trtc, srv, arenas, reqs)
    for r, o in zip(reqs, outs):
        check_oracle(orc, arenas, r, o)
    st = srv.stats()
    assert st["served"] == 80
    # one request larger than the budget is admitted only onto an idle server
    assert st["peak_held_bytes"] <= max(budget, 40 << 20)


def test_server_replay_short_trace(trtc, wl, arenas):
    say("server: replay of a 300-request Poisson trace")
    h, d, ptr = arenas
    cnt = 300
    gaps = wl.exponential(cnt, 8, 1.0 / 200.0)  # mean gap 200 us
    arrival = np.cumsum(gaps)
    u = wl.unit_f64(2 * cnt, 9)
    ns = np.floor(2.0 ** (10 + 12 * u[:cnt])).astype(np.uint64)
    ops, ke, ve, pa, pb = [], [], [], [], []
    for i in range(cnt):
        n = int(ns[i])
        off = int((ARENA - n) * u[cnt + i])
        m = u[cnt + i]
        if m < 0.6:
            ops.append(0); ke.append("uint32_t"); ve.append("uint32_t")
            pa.append(ptr["keys"] + 4 * off); pb.append(ptr["vals"] + 4 * off)
        elif m < 0.9:
            ops.append(2); ke.append("int32_t"); ve.append(None); pa.append(ptr["x"] + 4 * off); pb.append(None)
        else:
            ops.append(3); ke.append("float"); ve.append(None); pa.append(ptr["f"] + 4 * off); pb.append(None)
    srv = trtc.Server(64 << 30, 3, True)
    span, lat = srv.replay(arrival, ops, ke, ve, pa, pb, ns)
    assert (lat > 0).all()
    assert span > 0
    assert srv.stats()["served"] == cnt
