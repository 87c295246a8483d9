"""GPU parity of the batched segmented primitives and the graph-captured
Batch: every segment must equal the oracle's result for that segment alone
(sorts and integer scans bit-exact, stability pinned by an index payload;
float reductions within 1e-5 relative of the FP64 sum), for every tile
schedule the autotuner can choose, and a segment's result must not depend on
what else is in the batch.
"""
import numpy as np
import pytest

from helpers import Oracle, workloads, NP

pytestmark = pytest.mark.gpu

# every schedule the batch autotuner may pick for 4-byte and 8-byte keys
# (test_tuner_candidates_are_all_tested checks this against the tuner's own
# candidate list)
SCHEDULES_4 = ["rest:512x16", "rest:512x8", "rest:256x16", "<=2048:256x8 rest:512x16", "<=4096:256x16 rest:512x16",
               "<=2048:256x8 <=262144:512x16 rest:1024x16", "<=4096:256x16 <=262144:512x16 rest:1024x16",
               "<=2048:256x8 <=262144:512x16 rest:1024x32", "<=4096:256x16 <=262144:512x16 rest:1024x32"]
SCHEDULES_8 = ["rest:512x8", "rest:256x8", "<=2048:256x8 rest:512x8", "<=2048:256x8 <=262144:512x8 rest:1024x8"]


@pytest.fixture(scope="module")
def trtc():
    import thrustrtc_amd
    yield thrustrtc_amd
    thrustrtc_amd.Set_Batch_Schedule(None)


@pytest.fixture(scope="module")
def orc():
    return Oracle()


@pytest.fixture(scope="module")
def wl():
    return workloads()


def say(*a):
    print(*a, flush=True)


def edge_sizes():
    # empty, single, tile boundaries of every schedule, multi-tile segments
    return [0, 1, 2, 3, 63, 64, 65, 511, 512, 1000, 2047, 2048, 2049, 4095, 4096, 4097, 8191, 8192, 8193, 0,
            16383, 16384, 16385, 30000, 100003, 5, 262144 + 17]


def large_edge_sizes():
    # segments on both sides of the 2^18 class boundary of the 1024-thread
    # schedules, and 1024x32 (32768-key) / 1024x16 (16384-key) / 1024x8
    # (8192-key) tile edges
    return [262143, 262144, 262145, 262144 + 8191, 262144 + 8193, 262144 + 16383, 262144 + 16385,
            262144 + 32767, 262144 + 32769, 10 * 32768, 10 * 32768 + 1, 1 << 20, (1 << 20) + 1]


def sizes_for(schedule):
    return edge_sizes() + (large_edge_sizes() if "1024x" in schedule else [])


def offsets_of(sizes):
    return np.concatenate([[0], np.cumsum(np.asarray(sizes, dtype=np.uint64))]).astype(np.uint64)


def keys_for(wl, kind, n, seed):
    if kind == "u32":
        return wl.u32_uniform(n, seed), "uint32_t"
    if kind == "u32_low":
        return wl.u32_low_entropy(n, seed), "uint32_t"
    if kind == "f32":
        return wl.f32_normal_signed_zeros(n, seed, seed + 1), "float"
    if kind == "i32":
        return wl.i32_range(n, seed, -50, 50), "int32_t"  # many ties
    if kind == "u64":
        a = wl.u32_uniform(n, seed).astype(np.uint64) << np.uint64(32)
        return a | (wl.u32_low_entropy(n, seed + 1).astype(np.uint64) & np.uint64(0xFF)), "uint64_t"
    if kind == "f64":
        return wl.f32_normal_signed_zeros(n, seed, seed + 1).astype(np.float64), "double"
    raise ValueError(kind)


def check_segments_sorted(orc, keys, out_k, out_v, offs, elem, greater, vals_dtype=np.uint32):
    for s in range(len(offs) - 1):
        b, e = int(offs[s]), int(offs[s + 1])
        if e == b:
            continue
        perm = orc.sort_perm(keys[b:e], elem, greater)
        exp_k = keys[b:e][perm]
        assert (out_k[b:e].view(f"u{keys.itemsize}") == exp_k.view(f"u{keys.itemsize}")).all(), f"segment {s} keys"
        if out_v is not None:
            assert (out_v[b:e] == (perm + b).astype(vals_dtype)).all(), f"segment {s} payload (stability)"


@pytest.mark.parametrize("schedule", SCHEDULES_4)
@pytest.mark.parametrize("kind", ["u32", "f32", "i32"])
@pytest.mark.parametrize("greater", [False, True])
def test_segmented_sort_by_key_4byte(trtc, orc, wl, schedule, kind, greater):
    say(f"segmented sort_by_key {kind} greater={greater} schedule={schedule}")
    sizes = sizes_for(schedule)
    offs = offsets_of(sizes)
    n = int(offs[-1])
    keys, elem = keys_for(wl, kind, n, 11)
    dk = trtc.device_vector_from_numpy(keys)
    dv = trtc.device_vector_from_numpy(np.arange(n, dtype=np.uint32))
    trtc.Set_Batch_Schedule(schedule)
    trtc.Segmented_Sort_By_Key(dk, dv, offs, trtc.Greater() if greater else None)
    check_segments_sorted(orc, keys, dk.to_host(), dv.to_host(), offs, elem, greater)


@pytest.mark.parametrize("schedule", SCHEDULES_8)
@pytest.mark.parametrize("kind", ["u64", "f64"])
def test_segmented_sort_by_key_8byte(trtc, orc, wl, schedule, kind):
    say(f"segmented sort_by_key {kind} schedule={schedule}")
    sizes = sizes_for(schedule)
    offs = offsets_of(sizes)
    n = int(offs[-1])
    keys, elem = keys_for(wl, kind, n, 21)
    dk = trtc.device_vector_from_numpy(keys)
    dv = trtc.device_vector_from_numpy(np.arange(n, dtype=np.uint64))
    trtc.Set_Batch_Schedule(schedule)
    trtc.Segmented_Sort_By_Key(dk, dv, offs)
    check_segments_sorted(orc, keys, dk.to_host(), dv.to_host(), offs, elem, False, np.uint64)


@pytest.mark.parametrize("schedule", ["rest:512x16", "<=2048:256x8 rest:512x16",
                                      "<=2048:256x8 <=262144:512x16 rest:1024x16",
                                      "<=2048:256x8 <=262144:512x16 rest:1024x32", "rest:1024x32"])
def test_segmented_sort_keys_only(trtc, orc, wl, schedule):
    say(f"segmented keys-only sort schedule={schedule}")
    offs = offsets_of(sizes_for(schedule))
    n = int(offs[-1])
    keys, elem = keys_for(wl, "u32_low", n, 31)
    dk = trtc.device_vector_from_numpy(keys)
    trtc.Set_Batch_Schedule(schedule)
    trtc.Segmented_Sort(dk, offs)
    check_segments_sorted(orc, keys, dk.to_host(), None, offs, elem, False)


@pytest.mark.parametrize("elem", ["uint8_t", "int16_t"])
@pytest.mark.parametrize("schedule", ["rest:512x16", "rest:256x8"])
def test_segmented_sort_narrow_keys(trtc, orc, wl, elem, schedule):
    """1- and 2-byte keys: multi-tile segments start up to 15 / 7 keys past a
    16-byte boundary, and tile 0 of each gives up that many keys so that the
    later tiles start aligned."""
    say(f"segmented sort_by_key {elem} keys schedule={schedule}")
    trtc.Set_Batch_Schedule(schedule)
    sizes = [1, 9000, 3, 20000, 7, 8193, 13, 16385, 5, 40000, 2, 2049, 11, 4097]
    offs = offsets_of(sizes)
    n = int(offs[-1])
    raw = wl.u32_uniform(n, 47)
    keys = raw.astype(np.uint8) if elem == "uint8_t" else (raw >> np.uint32(16)).astype(np.uint16).view(np.int16)
    dk = trtc.device_vector_from_numpy(keys)
    dv = trtc.device_vector_from_numpy(np.arange(n, dtype=np.uint32))
    trtc.Segmented_Sort_By_Key(dk, dv, offs)
    check_segments_sorted(orc, keys, dk.to_host(), dv.to_host(), offs, elem, False)


def test_segmented_sort_narrow_values(trtc, orc, wl):
    say("segmented sort_by_key u32 keys, int16 values, odd widths")
    trtc.Set_Batch_Schedule(None)
    offs = offsets_of([5000, 9000, 17, 40000])
    n = int(offs[-1])
    keys, elem = keys_for(wl, "u32", n, 41)
    vals = (np.arange(n) % 30000).astype(np.int16)
    dk = trtc.device_vector_from_numpy(keys)
    dv = trtc.device_vector_from_numpy(vals)
    trtc.Segmented_Sort_By_Key(dk, dv, offs)
    ok, ov = dk.to_host(), dv.to_host()
    for s in range(len(offs) - 1):
        b, e = int(offs[s]), int(offs[s + 1])
        perm = orc.sort_perm(keys[b:e], elem)
        assert (ok[b:e] == keys[b:e][perm]).all()
        assert (ov[b:e] == vals[b:e][perm]).all()


def test_segmented_sort_rejects_bad_offsets(trtc):
    dk = trtc.device_vector("uint32_t", 100)
    dv = trtc.device_vector("uint32_t", 100)
    with pytest.raises(SystemError):
        trtc.Segmented_Sort_By_Key(dk, dv, [0, 50, 40])
    with pytest.raises(SystemError):
        trtc.Segmented_Sort_By_Key(dk, dv, [0, 101])
    with pytest.raises(ValueError):
        trtc.Set_Batch_Schedule("rest:300x16")
    with pytest.raises(ValueError):
        trtc.Set_Batch_Schedule("rest:512x32")  # 32 keys per thread only in the 1024-thread long tiles
    trtc.Set_Batch_Schedule("rest:1024x32")  # long tiles: 4-byte keys with 4-byte or no values only
    d8 = trtc.device_vector("uint64_t", 100)
    with pytest.raises(SystemError):
        trtc.Segmented_Sort_By_Key(d8, dv, [0, 50, 100])
    trtc.Set_Batch_Schedule(None)


def test_segmented_log_uniform_autotuned(trtc, orc, wl):
    """The batched-segments workload shape at a reduced size: 256 segments,
    sizes log-uniform in [16, 2^17], autotuned schedule."""
    say("segmented log-uniform 256 segments, autotuned")
    trtc.Set_Batch_Schedule(None)
    trtc.Set_Batch_Autotune(True)
    sizes, total = wl.segment_sizes(256, 7, 16, 1 << 17)
    offs = offsets_of(sizes)
    keys = wl.u32_uniform(total, 2)
    dk = trtc.device_vector_from_numpy(keys)
    dv = trtc.device_vector_from_numpy(np.arange(total, dtype=np.uint32))
    trtc.Segmented_Sort_By_Key(dk, dv, offs)
    check_segments_sorted(orc, keys, dk.to_host(), dv.to_host(), offs, "uint32_t", False)
    rep = trtc.Batch_Tuning_Report()
    assert rep and rep[-1]["chosen"] in [c["schedule"] for c in rep[-1]["candidates"]]


@pytest.mark.parametrize("vals", ["uint32_t", "uint64_t"])
def test_tuner_candidates_are_all_tested(trtc, wl, vals):
    """Every schedule the autotuner may choose is one of the schedules the
    parity tests above run (a new candidate must come with its parity test)."""
    say(f"autotuner candidates, {vals} values")
    trtc.Set_Batch_Schedule(None)
    trtc.Set_Batch_Autotune(True)
    sizes, total = wl.segment_sizes(96, 23, 16, 1 << 19)
    offs = offsets_of(sizes)
    dk = trtc.device_vector_from_numpy(wl.u32_uniform(total, 3))
    dv = trtc.device_vector(vals, total)
    trtc.Segmented_Sort_By_Key(dk, dv, offs)
    rep = trtc.Batch_Tuning_Report()
    assert rep
    tested = SCHEDULES_4 if vals == "uint32_t" else SCHEDULES_8
    for c in rep[-1]["candidates"]:
        assert c["schedule"] in tested, c["schedule"]


# ---------------------------------------------------------------- scan
@pytest.mark.parametrize("elem", ["int32_t", "uint32_t", "int64_t", "float", "double"])
@pytest.mark.parametrize("exclusive", [False, True])
def test_segmented_scan(trtc, orc, wl, elem, exclusive):
    say(f"segmented scan {elem} exclusive={exclusive}")
    # edges of the 16384-element tiles (512 x 32 for 4-byte, 1024 x 16 for 8-byte elements)
    offs = offsets_of(edge_sizes() + [32767, 32768, 32769, 3 * 16384, 3 * 16384 + 1])
    n = int(offs[-1])
    # float values in {-1, 0, 1}: every prefix stays an exact integer
    lo, hi = (-1, 1) if elem == "float" else (-1000, 1000)
    x = wl.i32_range(n, 5, lo, hi).astype(NP[elem])
    dx = trtc.device_vector_from_numpy(x)
    dy = trtc.device_vector(elem, n)
    if exclusive:
        trtc.Segmented_Exclusive_Scan(dx, dy, offs)
    else:
        trtc.Segmented_Inclusive_Scan(dx, dy, offs)
    y = dy.to_host()
    for s in range(len(offs) - 1):
        b, e = int(offs[s]), int(offs[s + 1])
        if e == b:
            continue
        exp = orc.exclusive_scan(x[b:e], elem) if exclusive else orc.inclusive_scan(x[b:e], elem)
        assert (y[b:e] == exp).all(), f"segment {s}"


def test_segmented_scan_in_place_and_equals_scan_by_key(trtc, orc, wl):
    say("segmented scan in place == scan_by_key over segment ids")
    sizes, total = wl.segment_sizes(64, 9, 16, 1 << 16)
    offs = offsets_of(sizes)
    x = wl.i32_range(total, 3, -1000, 1000)
    ids = np.repeat(np.arange(len(sizes), dtype=np.int32), sizes.astype(np.int64))
    dx = trtc.device_vector_from_numpy(x)
    trtc.Segmented_Inclusive_Scan(dx, dx, offs)
    dk = trtc.device_vector_from_numpy(ids)
    dv = trtc.device_vector_from_numpy(x)
    dz = trtc.device_vector("int32_t", total)
    trtc.Inclusive_Scan_By_Key(dk, dv, dz)
    assert (dx.to_host() == dz.to_host()).all()
    assert (dz.to_host() == orc.scan_by_key(ids, x, "int32_t", "int32_t")).all()


# ---------------------------------------------------------------- reduce
def test_segmented_reduce(trtc, orc, wl):
    say("segmented reduce f32 / i32")
    sizes = [s for s in edge_sizes() if s > 0] + [1 << 20, 16384 * 3 + 1]
    offs = offsets_of(sizes)
    n = int(offs[-1])
    f = wl.f32_unit(n, 6)
    df = trtc.device_vector_from_numpy(f)
    res = trtc.device_vector("float", len(sizes))
    trtc.Segmented_Reduce(df, offs, res)
    r = res.to_host()
    for s in range(len(sizes)):
        b, e = int(offs[s]), int(offs[s + 1])
        ref = orc.sum_f32_fp64(f[b:e])
        assert abs(float(r[s]) - ref) <= 1e-5 * abs(ref), f"segment {s}"
    x = wl.i32_range(n, 8, -1000, 1000)
    dx = trtc.device_vector_from_numpy(x)
    ri = trtc.device_vector("int32_t", len(sizes))
    trtc.Segmented_Reduce(dx, offs, ri, trtc.Maximum())
    got = ri.to_host()
    for s in range(len(sizes)):
        b, e = int(offs[s]), int(offs[s + 1])
        assert got[s] == x[b:e].max()


def test_segmented_reduce_alone_equals_in_batch(trtc, wl):
    """A segment's float result is bit-identical alone and in any batch."""
    say("segmented reduce alone == in batch")
    sizes = [100003, 7, 1 << 18, 16385]
    offs = offsets_of(sizes)
    n = int(offs[-1])
    f = wl.f32_unit(n, 12)
    df = trtc.device_vector_from_numpy(f)
    res = trtc.device_vector("float", len(sizes))
    trtc.Segmented_Reduce(df, offs, res)
    together = res.to_host()
    for s in range(len(sizes)):
        b, e = int(offs[s]), int(offs[s + 1])
        one = trtc.device_vector_from_numpy(f[b:e].copy())
        r1 = trtc.device_vector("float", 1)
        trtc.Segmented_Reduce(one, [0, e - b], r1)
        assert r1.to_host()[0].tobytes() == together[s].tobytes()


# ---------------------------------------------------------------- graph batch
def test_batch_graph_replays(trtc, orc, wl):
    """Sort_By_Key + Inclusive_Scan + Reduce over the same segments captured in
    one graph; every replay (with eager work and synchronisation between)
    matches the oracle."""
    say("batch graph: sort + scan + reduce, 3 replays")
    trtc.Set_Batch_Schedule(None)
    sizes, total = wl.segment_sizes(128, 17, 16, 1 << 16)
    sizes = np.concatenate([sizes, np.array([1, 0, 20000], dtype=np.uint64)])
    offs = offsets_of(sizes)
    total = int(offs[-1])
    keys = wl.u32_uniform(total, 4)
    x = wl.i32_range(total, 5, -1000, 1000)
    f = wl.f32_unit(total, 6)
    dk0 = trtc.device_vector_from_numpy(keys)
    dk = trtc.device_vector("uint32_t", total)
    dv = trtc.device_vector("uint32_t", total)
    dx = trtc.device_vector_from_numpy(x)
    dy = trtc.device_vector("int32_t", total)
    df = trtc.device_vector_from_numpy(f)
    nz = [s for s in range(len(sizes)) if sizes[s] > 0]
    roffs = offsets_of([sizes[s] for s in nz])
    df_nz = trtc.device_vector_from_numpy(np.concatenate([f[int(offs[s]):int(offs[s + 1])] for s in nz]))
    res = trtc.device_vector("float", len(nz))
    b = trtc.Batch()
    b.add_sort_by_key(dk, dv, offs)
    b.add_inclusive_scan(dx, dy, offs)
    b.add_reduce(df_nz, roffs, res)
    b.instantiate()
    iota = trtc.device_vector_from_numpy(np.arange(total, dtype=np.uint32))
    first = None
    for rep in range(3):
        trtc.Copy(dk0, dk)
        trtc.Copy(iota, dv)
        trtc.Fill(dy, trtc.DVInt32(0))
        trtc.Wait()
        b.launch()
        trtc.Wait()
        check_segments_sorted(orc, keys, dk.to_host(), dv.to_host(), offs, "uint32_t", False)
        y = dy.to_host()
        for s in range(len(sizes)):
            lo, hi = int(offs[s]), int(offs[s + 1])
            if hi > lo:
                assert (y[lo:hi] == orc.inclusive_scan(x[lo:hi], "int32_t")).all()
        r = res.to_host()
        if first is None:
            first = r.copy()
            for i, s in enumerate(nz):
                lo, hi = int(offs[s]), int(offs[s + 1])
                ref = orc.sum_f32_fp64(f[lo:hi])
                assert abs(float(r[i]) - ref) <= 1e-5 * abs(ref)
        else:
            assert r.tobytes() == first.tobytes()
    d = b.describe()
    assert [o["op"] for o in d] == ["sort", "inclusive_scan", "reduce"]
