"""Headline benchmark: TRTC_Sort_By_Key on 2^28 uint32 keys + uint32 payload
on one MI355X, with the HBM-roofline fraction of the dominant kernel and the
host baseline of the same call.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n LOG2] [--no-cpu] [--no-variants]

One step = one Sort_By_Key of N = 2^28 pairs.  Every step sorts its own
pristine copy of the input, all copies resident in HBM before the timed region
starts (no restore inside it).  Prints exactly one JSON line on stdout.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def golden():
    """Hashes of the oracle's outputs at the full workload sizes (written by
    tests/golden/make_large_hashes.py); None when absent."""
    p = os.path.join(ROOT, "tests", "golden", "large_hashes.json")
    return json.load(open(p)) if os.path.exists(p) else None


def kernel_ms(prof, prefix):
    names = [k for k in prof if k.startswith(prefix)]
    if not names:
        return None, None
    k = names[0]
    return k, prof[k]["total_ms"] / prof[k]["count"]


def profiled(native, fn, reps):
    """Runs fn() reps times between n_profile_begin/end; returns (wall seconds
    per call, per-kernel launch times)."""
    import ctypes
    native.n_profile_begin()
    t0 = time.perf_counter()
    for _ in range(reps):
        fn()
    native.n_wait()
    dt = (time.perf_counter() - t0) / reps
    buf = ctypes.create_string_buffer(1 << 16)
    native.n_profile_end(buf, 1 << 16)
    return dt, json.loads(buf.value.decode())


def leg_copy(trtc, n, reps, warmup):
    """Device copy of n uint32 (4 B read + 4 B written per element), two ways:
    TRTC_Copy between plain vectors (the 16-byte-per-thread copy kernel
    trtc_copy_bytes16: the measured HBM copy ceiling beside the 8 TB/s
    nominal peak) and a JIT For kernel with the body `dst[i] = src[i]` (what
    a user-written element-wise loop reaches)."""
    src = trtc.device_vector("uint32_t", n)
    dst = trtc.device_vector("uint32_t", n)
    loop = trtc.For(["src", "dst"], "i", "    dst[i] = src[i];\n")

    def timed(fn):
        for _ in range(warmup):
            fn()
        trtc.Wait()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        trtc.Wait()
        sec = (time.perf_counter() - t0) / reps
        return {"ms": round(sec * 1e3, 4), "gbs": round(8 * n / sec / 1e9, 1)}

    ceiling = timed(lambda: trtc.Copy(src, dst))
    for_copy = timed(lambda: loop.launch_n(n, [src, dst]))
    return {"n": n, "bytes_per_launch": 8 * n, "ms": ceiling["ms"], "gbs": ceiling["gbs"],
            "kernel": "trtc_copy_bytes16 (TRTC_Copy)", "for_copy": for_copy}


def leg_single_pass(trtc, native, wl, a, log):
    """Inclusive_Scan int32, Reduce float32 and Copy_If int32 (x % 3 == 0,
    hipRTC functor) on 2^29 elements, each checked against the oracle's
    full-size output hash."""
    import numpy as np
    n = 1 << 29
    reps = 10
    gold = golden()
    res = {}
    x = wl.i32_range(n, 5, -1000, 1000)
    dx = trtc.device_vector_from_numpy(x)
    dout = trtc.device_vector("int32_t", n)

    def h(arr):
        return "%016x" % wl._lib.wl_fnv1a64(arr.ctypes.data, arr.nbytes)

    # --- Inclusive_Scan
    trtc.Inclusive_Scan(dx, dout)
    trtc.Wait()
    dt, prof = profiled(native, lambda: trtc.Inclusive_Scan(dx, dout), reps)
    kname, kms = kernel_ms(prof, "trtc_scan_k<")
    ok = None
    if gold:
        ok = h(dout.to_host()) == gold["inclusive_scan_i32_2p29_seed5"]["out"]
    res["inclusive_scan_i32"] = {"ms": round(dt * 1e3, 4), "gbs_8N": round(8 * n / dt / 1e9, 1),
                                 "kernel": kname, "kernel_ms": round(kms, 4),
                                 "kernel_gbs": round(8 * n / kms / 1e6, 1),
                                 "frac_hbm": round(8 * n / kms / 1e6 / HBM_PEAK_GBS, 4), "bit_exact": ok}
    log(f"scan: {res['inclusive_scan_i32']}")

    # --- Copy_If with a hipRTC functor
    pred = trtc.Functor({}, ["x"], "        return x % 3 == 0;\n")
    t0 = time.perf_counter()
    cnt = trtc.Copy_If(dx, dout, pred)
    first = time.perf_counter() - t0
    dt, prof = profiled(native, lambda: trtc.Copy_If(dx, dout, pred), reps)
    kname, kms = kernel_ms(prof, "trtc_copy_if_k<")
    algo = 4 * n + 4 * cnt
    ok = None
    if gold:
        g = gold["copy_if_mod3_i32_2p29_seed5"]
        ok = cnt == g["count"] and h(dout.to_host(0, cnt)) == g["out"]
    res["copy_if_i32_mod3"] = {"ms": round(dt * 1e3, 4), "kept": int(cnt), "algorithmic_bytes": algo,
                               "kernel": kname, "kernel_ms": round(kms, 4),
                               "kernel_gbs": round(algo / kms / 1e6, 1),
                               "frac_hbm": round(algo / kms / 1e6 / HBM_PEAK_GBS, 4),
                               "first_call_ms_incl_jit": round(first * 1e3, 1), "bit_exact": ok}
    log(f"copy_if: {res['copy_if_i32_mod3']}")
    del dx, dout

    # --- Reduce float32
    f = wl.f32_unit(n, 6)
    df = trtc.device_vector_from_numpy(f)
    trtc.Reduce(df)
    r = [None]

    def red():
        r[0] = trtc.Reduce(df)
    dt, prof = profiled(native, red, reps)
    kname, kms = kernel_ms(prof, "trtc_reduce_k<")
    rel = None
    if gold:
        ref = gold["reduce_f32_2p29_seed6"]["sum_fp64"]
        rel = abs(float(r[0]) - ref) / abs(ref)
    res["reduce_f32"] = {"ms": round(dt * 1e3, 4), "gbs_4N": round(4 * n / dt / 1e9, 1),
                         "kernel": kname, "kernel_ms": round(kms, 4),
                         "kernel_gbs": round(4 * n / kms / 1e6, 1),
                         "frac_hbm": round(4 * n / kms / 1e6 / HBM_PEAK_GBS, 4),
                         "rel_err_vs_fp64": rel, "within_1e-5": None if rel is None else rel <= 1e-5}
    log(f"reduce: {res['reduce_f32']}")
    del df

    if not a.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        from helpers import Oracle
        L = Oracle().L
        out = np.empty_like(x)
        t0 = time.perf_counter()
        L.bl_inclusive_scan_i32(x.ctypes.data, out.ctypes.data, n)
        t_scan = time.perf_counter() - t0
        t0 = time.perf_counter()
        L.bl_copy_if_mod3_i32(x.ctypes.data, out.ctypes.data, n)
        t_cif = time.perf_counter() - t0
        t0 = time.perf_counter()
        L.bl_reduce_f32(f.ctypes.data, n)
        t_red = time.perf_counter() - t0
        res["cpu"] = {"cores": int(L.bl_threads()), "kind": "port",
                      "inclusive_scan_gbs_8N": round(8 * n / t_scan / 1e9, 2),
                      "copy_if_gbs": round((4 * n + 4 * res["copy_if_i32_mod3"]["kept"]) / t_cif / 1e9, 2),
                      "reduce_gbs_4N": round(4 * n / t_red / 1e9, 2),
                      "sample": "the full 2^29 inputs, one call each (__gnu_parallel::partial_sum, "
                                "OpenMP count/scan/write copy_if, __gnu_parallel::accumulate)"}
        log(f"cpu single-pass: {res['cpu']}")
    return res


def leg_consumers(trtc, native, wl, a, log):
    """The scan / compaction consumers and the comparison sort on 2^28 int32
    items: fused Reduce_By_Key, Inclusive_Scan_By_Key and Unique_Copy over keys in runs of 1..64
    (log-uniform), Partition_Copy / Remove_Copy_If with x % 3 == 0,
    Transform_Inclusive_Scan (Negate), Transform (Negate), Gather (random map),
    Copy (int32, and int32 -> float32), Min_Element, Merge, and Sort with a
    user comparator (merge path) on 2^24 uint32 keys.  Each GPU result is compared with the oracle's
    (which is also the timed single-core host baseline)."""
    import numpy as np
    n = 1 << 28
    reps = 5
    res = {}
    sizes, _ = wl.segment_sizes(n // 8, 51, 1, 64)
    k = int(np.searchsorted(np.cumsum(sizes.astype(np.int64)), n)) + 1
    keys = np.repeat(np.arange(k, dtype=np.int32), sizes[:k].astype(np.int64))[:n]
    vals = wl.i32_range(n, 52, -1000, 1000)
    x = wl.i32_range(n, 5, -1000, 1000)
    dk, dv, dx = (trtc.device_vector_from_numpy(v) for v in (keys, vals, x))
    o1, o2 = trtc.device_vector("int32_t", n), trtc.device_vector("int32_t", n)
    mod3 = trtc.Functor({}, ["x"], "        return x % 3 == 0;\n")
    orc = None
    if not a.no_cpu:
        from helpers import Oracle
        orc = Oracle()

    def entry(name, fn, prefix, algo_fn, check):
        fn()
        trtc.Wait()
        out = [None]

        def call():
            out[0] = fn()
        dt, prof = profiled(native, call, reps)
        if isinstance(prefix, tuple):  # several launches per call: their times add up
            found = [kernel_ms(prof, p_) for p_ in prefix]
            kname = " + ".join(k_ for k_, _ in found if k_)
            kms = sum(m_ for _, m_ in found if m_)
        else:
            kname, kms = kernel_ms(prof, prefix)
        if not kms:
            raise RuntimeError(f"{name}: no launch of {prefix} recorded during the timed calls")
        algo = algo_fn(out[0])
        r = {"ms": round(dt * 1e3, 4), "kernel": kname, "kernel_ms": round(kms, 4), "algorithmic_bytes": algo,
             "kernel_gbs": round(algo / kms / 1e6, 1), "frac_hbm": round(algo / kms / 1e6 / HBM_PEAK_GBS, 4)}
        if orc is not None:
            t0 = time.perf_counter()
            ok = check(out[0])
            r["cpu_oracle_s"] = round(time.perf_counter() - t0, 3)
            r["cpu_gbs"] = round(algo / r["cpu_oracle_s"] / 1e9, 3)
            r["bit_exact"] = bool(ok)
        res[name] = r
        log(f"{name}: {r}")

    def chk_rbk(c):
        wk, wvals = orc.reduce_by_key(keys, vals, "int32_t", "int32_t")
        return c == len(wk) and (o1.to_host(0, c) == wk).all() and (o2.to_host(0, c) == wvals).all()
    entry("reduce_by_key_i32", lambda: trtc.Reduce_By_Key(dk, dv, o1, o2), "trtc_reduce_by_key_k<",
          lambda c: 8 * n + 8 * c, chk_rbk)

    def chk_sbk(_):
        return (o1.to_host() == orc.scan_by_key(keys, vals, "int32_t", "int32_t")).all()
    entry("inclusive_scan_by_key_i32", lambda: trtc.Inclusive_Scan_By_Key(dk, dv, o1), "trtc_scan_by_key_k<",
          lambda _: 12 * n, chk_sbk)

    def chk_uniq(c):
        (w,) = orc.unique(keys, "int32_t")
        return c == len(w) and (o1.to_host(0, c) == w).all()
    entry("unique_copy_i32", lambda: trtc.Unique_Copy(dk, o1), "trtc_unique_k<", lambda c: 4 * n + 4 * c, chk_uniq)
    del dk, dv

    def chk_part(c):
        wt, wf = orc.partition(x, "int32_t", ("mod", 3, 0))
        return c == len(wt) and (o1.to_host(0, c) == wt).all() and (o2.to_host(0, n - c) == wf).all()
    entry("partition_copy_i32_mod3", lambda: trtc.Partition_Copy(dx, o1, o2, mod3), "trtc_partition_k<",
          lambda c: 8 * n, chk_part)

    def chk_rem(c):
        w = orc.select(x, "int32_t", ("mod", 3, 0), negate=True)
        return c == len(w) and (o1.to_host(0, c) == w).all()
    entry("remove_copy_if_i32_mod3", lambda: trtc.Remove_Copy_If(dx, o1, mod3), "trtc_copy_if_k<",
          lambda c: 4 * n + 4 * c, chk_rem)

    def chk_tscan(_):
        return (o1.to_host() == orc.transform_scan(x, "int32_t", "Negate")).all()
    entry("transform_inclusive_scan_i32_negate", lambda: trtc.Transform_Inclusive_Scan(dx, o1, trtc.Negate(),
          trtc.Plus()), "trtc_scan_k<", lambda _: 8 * n, chk_tscan)
    # the other algorithms: element-wise, random gather, index reduction, merge
    def chk_neg(_):
        return (o1.to_host() == -x).all()
    entry("transform_negate_i32", lambda: trtc.Transform(dx, o1, trtc.Negate()), "trtc_map", lambda _: 8 * n, chk_neg)
    gmap = (wl.u32_uniform(n, 53) % np.uint32(n)).astype(np.int32)
    dmap = trtc.device_vector_from_numpy(gmap)

    def chk_gather(_):
        return (o1.to_host() == x[gmap]).all()
    entry("gather_i32_random_map", lambda: trtc.Gather(dmap, dx, o1), "trtc_for", lambda _: 12 * n, chk_gather)
    del dmap

    def chk_copy(_):
        return (o1.to_host() == x).all()
    entry("copy_i32", lambda: trtc.Copy(dx, o1), "trtc_copy_bytes16", lambda _: 8 * n, chk_copy)
    of = trtc.device_vector("float", n)

    def chk_copy_f32(_):
        return (of.to_host() == x.astype(np.float32)).all()
    entry("copy_i32_to_f32", lambda: trtc.Copy(dx, of), "trtc_map", lambda _: 8 * n, chk_copy_f32)
    del of

    def chk_min(r):
        return r == int(np.argmin(x))
    entry("min_element_i32", lambda: trtc.Min_Element(dx), "trtc_extreme_k<", lambda _: 4 * n, chk_min)
    half = n // 2
    sa, sb = np.sort(x[:half]), np.sort(x[half:])
    da, db = trtc.device_vector_from_numpy(sa), trtc.device_vector_from_numpy(sb)

    def chk_merge(_):
        return (o1.to_host() == np.sort(x)).all()
    entry("merge_i32_2x2p27", lambda: trtc.Merge(da, db, o1), ("trtc_merge_split_k<", "trtc_merge_k<"), lambda _: 8 * n, chk_merge)
    del da, db, dx, o1, o2

    # comparison sort with a user comparator: LDS tile sort + merge passes
    ns = 1 << 24
    ks = wl.u32_uniform(ns, 2)
    src = trtc.device_vector_from_numpy(ks)
    work = trtc.device_vector("uint32_t", ns)
    less = trtc.Functor({}, ["x", "y"], "        return x < y;\n")
    trtc.Copy(src, work)
    trtc.Sort(work, less)
    trtc.Wait()

    def msort():
        trtc.Copy(src, work)
        trtc.Sort(work, less)
    dt, prof = profiled(native, msort, reps)
    mk = {kk: v for kk, v in prof.items() if kk.startswith("trtc_msort")}
    kms = sum(v["total_ms"] for v in mk.values()) / reps
    passes = 1 + max(0, int(np.ceil(np.log2(ns / 2048))))
    algo = passes * 8 * ns  # every pass reads and writes each 4-byte key once
    r = {"n": ns, "ms_incl_restore": round(dt * 1e3, 4), "kernels_ms": round(kms, 4), "passes": passes,
         "mkeys_per_s": round(ns / kms / 1e3, 1), "algorithmic_bytes": algo,
         "kernel_gbs": round(algo / kms / 1e6, 1), "frac_hbm": round(algo / kms / 1e6 / HBM_PEAK_GBS, 4),
         "kernels": mk}
    if orc is not None:
        got = work.to_host()
        ck = ks.copy()
        t0 = time.perf_counter()
        orc.L.bl_sort_u32(ck.ctypes.data, ns)
        r["cpu_sort_s"] = round(time.perf_counter() - t0, 3)
        r["cpu_cores"] = int(orc.L.bl_threads())
        r["bit_exact"] = bool((got == ck).all())
    res["merge_sort_u32_user_comparator_2p24"] = r
    log(f"merge sort: {r}")
    res["cpu"] = {"kind": "port", "cores": 1,
                  "sample": "the full inputs, one sequential oracle call each (the same call checks the GPU "
                            "result); the merge-sort row uses __gnu_parallel::sort on all cores"}
    return res


def leg_small_calls(trtc, native, wl, log, cpu=True):
    """Latency of single small calls: TRTC_Sort on 2^20 uint32 (seed 0) and
    TRTC_Inclusive_Scan on 2^20 int32 in [-100, 100] (seed 0); median over
    100 calls, input restored by a device copy outside the timed call."""
    import numpy as np
    n = 1 << 20
    gold = golden()
    k = wl.u32_uniform(n, 0)
    src = trtc.device_vector_from_numpy(k)
    dk = trtc.device_vector("uint32_t", n)
    lat = []
    for i in range(110):
        trtc.Copy(src, dk)
        trtc.Wait()
        t0 = time.perf_counter()
        trtc.Sort(dk)
        trtc.Wait()
        if i >= 10:
            lat.append(time.perf_counter() - t0)
    out = dk.to_host()
    ok = None
    if gold:
        ok = "%016x" % wl._lib.wl_fnv1a64(out.ctypes.data, out.nbytes) == gold["sort_u32_2p20_uniform_seed0"]["keys_out"]
    res = {"sort_u32": {"median_us": round(1e6 * float(np.median(lat)), 1),
                        "p99_us": round(1e6 * float(np.percentile(lat, 99)), 1), "bit_exact": ok}}
    x = wl.i32_range(n, 0, -100, 100)
    dx = trtc.device_vector_from_numpy(x)
    dy = trtc.device_vector("int32_t", n)
    lat = []
    for i in range(110):
        t0 = time.perf_counter()
        trtc.Inclusive_Scan(dx, dy)
        trtc.Wait()
        if i >= 10:
            lat.append(time.perf_counter() - t0)
    out = dy.to_host()
    ok = None
    if gold:
        ok = "%016x" % wl._lib.wl_fnv1a64(out.ctypes.data, out.nbytes) == gold["inclusive_scan_i32_2p20_seed0"]["out"]
    res["inclusive_scan_i32"] = {"median_us": round(1e6 * float(np.median(lat)), 1),
                                 "p99_us": round(1e6 * float(np.percentile(lat, 99)), 1), "bit_exact": ok}
    if cpu:
        from helpers import Oracle
        L = Oracle().L
        ts, tc = [], []
        for i in range(51):
            kk = k.copy()
            t0 = time.perf_counter()
            L.bl_sort_u32(kk.ctypes.data, n)
            ts.append(time.perf_counter() - t0)
            yy = np.empty_like(x)
            t0 = time.perf_counter()
            L.bl_inclusive_scan_i32(x.ctypes.data, yy.ctypes.data, n)
            tc.append(time.perf_counter() - t0)
        res["cpu"] = {"cores": int(L.bl_threads()), "kind": "port",
                      "sort_u32_median_us": round(1e6 * float(np.median(ts[1:])), 1),
                      "inclusive_scan_i32_median_us": round(1e6 * float(np.median(tc[1:])), 1),
                      "sample": "the same 2^20 inputs, median of 50 calls after one warm-up "
                                "(__gnu_parallel::stable_sort, __gnu_parallel::partial_sum)"}
    log(f"small calls: {res}")
    return res


def leg_segmented(trtc, native, wl, a, log):
    """Batched segments: 4096 segments with sizes log-uniform in [16, 2^20]
    (mt19937_64 seed 7), uint32 keys (uniform, seed 2) + uint32 index payload
    Sort_By_Key and int32 (uniform [-1000, 1000], seed 5) Inclusive_Scan over
    the same segments, recorded in one Batch (autotuned tile schedule) and
    replayed as one hipGraph launch; against a loop of one Sort_By_Key and one
    Inclusive_Scan call per segment through the C ABI."""
    import ctypes
    import numpy as np
    sizes, total = wl.segment_sizes(a.segments, 7, 16, 1 << 20)
    offs = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint64)
    keys = wl.u32_uniform(total, 2)
    x = wl.i32_range(total, 5, -1000, 1000)
    k0 = trtc.device_vector_from_numpy(keys)
    iota = trtc.device_vector_from_numpy(np.arange(total, dtype=np.uint32))
    dk = trtc.device_vector("uint32_t", total)
    dv = trtc.device_vector("uint32_t", total)
    dx = trtc.device_vector_from_numpy(x)
    dy = trtc.device_vector("int32_t", total)

    def restore():
        trtc.Copy(k0, dk)
        trtc.Copy(iota, dv)
        trtc.Wait()

    def hsh(v):
        arr = v.to_host()
        return "%016x" % wl._lib.wl_fnv1a64(arr.ctypes.data, arr.nbytes)

    t0 = time.perf_counter()
    b = trtc.Batch()
    b.add_sort_by_key(dk, dv, offs)
    b.add_inclusive_scan(dx, dy, offs)
    restore()
    b.instantiate()
    inst_s = time.perf_counter() - t0
    reps = 5
    walls, gpu = [], []
    for i in range(reps + 1):
        restore()
        t1 = time.perf_counter()
        dt, prof = profiled(native, b.launch, 1)
        if i:
            walls.append(dt)
            gpu.append(prof.get("trtc_batch_graph", {"total_ms": dt * 1e3})["total_ms"])
    batched_ms = float(np.median(gpu))
    hk, hv, hy = hsh(dk), hsh(dv), hsh(dy)

    # per-segment loop through the C ABI (views created outside the timing)
    kp, vp = native.n_dvvectorlike_data(dk.m_cptr), native.n_dvvectorlike_data(dv.m_cptr)
    xp, yp = native.n_dvvectorlike_data(dx.m_cptr), native.n_dvvectorlike_data(dy.m_cptr)
    views = []
    for s in range(len(sizes)):
        lo, n = int(offs[s]), int(sizes[s])
        views.append((trtc.DVVectorAdaptor("uint32_t", n, kp + 4 * lo), trtc.DVVectorAdaptor("uint32_t", n, vp + 4 * lo),
                      trtc.DVVectorAdaptor("int32_t", n, xp + 4 * lo), trtc.DVVectorAdaptor("int32_t", n, yp + 4 * lo)))
    loop_ms = []
    for i in range(3):
        restore()
        trtc.Fill(dy, trtc.DVInt32(0))
        trtc.Wait()
        t1 = time.perf_counter()
        for (ka, va, xa, ya) in views:
            native.n_sort_by_key(ka.m_cptr, va.m_cptr, None)
            native.n_inclusive_scan(xa.m_cptr, ya.m_cptr, None)
        trtc.Wait()
        if i:
            loop_ms.append((time.perf_counter() - t1) * 1e3)
    same = (hsh(dk), hsh(dv), hsh(dy)) == (hk, hv, hy)
    loop = float(np.median(loop_ms))

    # oracle spot check of 32 segments (smallest, largest and a spread)
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from helpers import Oracle
    orc = Oracle()
    gk, gv, gy = dk.to_host(), dv.to_host(), dy.to_host()
    order = np.argsort(sizes)
    pick = sorted(set(order[:8].tolist() + order[-8:].tolist() + order[:: max(1, len(order) // 16)].tolist()))
    spot = True
    for s in pick:
        lo, hi = int(offs[s]), int(offs[s + 1])
        perm = orc.sort_perm(keys[lo:hi], "uint32_t")
        spot = spot and bool((gk[lo:hi] == keys[lo:hi][perm]).all()) and bool((gv[lo:hi] == perm + lo).all())
        spot = spot and bool((gy[lo:hi] == orc.inclusive_scan(x[lo:hi], "int32_t")).all())
    cpu = None
    if not a.no_cpu:
        kk, vv, yy = keys.copy(), np.arange(total, dtype=np.uint32), np.empty_like(x)
        t0 = time.perf_counter()
        orc.L.bl_segmented_sort_pairs_u32(kk.ctypes.data, vv.ctypes.data, offs.ctypes.data, len(sizes))
        orc.L.bl_segmented_inclusive_scan_i32(x.ctypes.data, yy.ctypes.data, offs.ctypes.data, len(sizes))
        dt = time.perf_counter() - t0
        cpu = {"ms": round(dt * 1e3, 1), "cores": int(orc.L.bl_threads()), "kind": "port",
               "sample": "the full workload once: OpenMP over segments, std::stable_sort of (key, value) pairs "
                         "and std::inclusive_scan per segment"}
        log(f"segmented cpu baseline: {cpu}")
    res = {"segments": int(len(sizes)), "total_elems": int(total),
           "sizes": "log-uniform [16, 2^20], mt19937_64 seed 7",
           "batched_graph_ms": round(batched_ms, 3), "batched_wall_ms": round(1e3 * float(np.median(walls)), 3),
           "per_segment_loop_ms": round(loop, 3), "speedup_vs_loop": round(loop / batched_ms, 2),
           "sort_gkeys_per_s_incl_scan": round(total / batched_ms / 1e6, 2),
           "pct_hbm_68N_plus_8N": round(100 * (76.0 * total) / (batched_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 2),
           "instantiate_s_incl_tuning": round(inst_s, 2),
           "batch": b.describe(), "tuning": trtc.Batch_Tuning_Report(),
           "batched_equals_loop_bitwise": same, "oracle_spot_check_32_segments": spot, "cpu": cpu}
    log(f"segmented: {res}")
    return res


def leg_served(trtc, native, wl, a, log):
    """Served stream: Poisson arrivals, sizes 2^U(10,24), 60% Sort_By_Key
    uint32/uint32, 30% Inclusive_Scan int32, 10% Reduce float32 on
    device-resident inputs; 64 GB admission budget.  The per-call baseline
    serves one request at a time through the single-call functions; the
    batched server admits and coalesces continuously on 3 streams.  Offered
    loads are 0.5x, 1x and 2x the measured per-call saturation rate."""
    import numpy as np
    cnt = a.served_requests
    arena = 1 << 25
    keys = trtc.device_vector_from_numpy(wl.u32_uniform(arena, 2))
    vals = trtc.device_vector_from_numpy(np.arange(arena, dtype=np.uint32))
    xs = trtc.device_vector_from_numpy(wl.i32_range(arena, 5, -1000, 1000))
    fs = trtc.device_vector_from_numpy(wl.f32_unit(arena, 6))
    pk, pv = native.n_dvvectorlike_data(keys.m_cptr), native.n_dvvectorlike_data(vals.m_cptr)
    px, pf = native.n_dvvectorlike_data(xs.m_cptr), native.n_dvvectorlike_data(fs.m_cptr)
    u = wl.unit_f64(4 * cnt, 8)
    gaps_unit = -np.log1p(-u[:cnt])                    # exponential, mean 1
    ns = np.floor(2.0 ** (10.0 + 14.0 * u[cnt:2 * cnt])).astype(np.uint64)
    ops, ke, ve, pa, pb = [], [], [], [], []
    for i in range(cnt):
        n = int(ns[i])
        off = int((arena - n) * u[3 * cnt + i])
        m = u[2 * cnt + i]
        if m < 0.6:
            ops.append(0); ke.append("uint32_t"); ve.append("uint32_t"); pa.append(pk + 4 * off); pb.append(pv + 4 * off)
        elif m < 0.9:
            ops.append(2); ke.append("int32_t"); ve.append(None); pa.append(px + 4 * off); pb.append(None)
        else:
            ops.append(3); ke.append("float"); ve.append(None); pa.append(pf + 4 * off); pb.append(None)
    elems = float(ns.sum())
    budget = 64 * 10 ** 9

    def replay(batching, rate):
        srv = trtc.Server(budget, 3 if batching else 1, batching)
        # warm-up: kernels loaded, pool populated
        srv.replay(np.zeros(min(cnt, 64)), ops[:64], ke[:64], ve[:64], pa[:64], pb[:64], ns[:64])
        arr = np.zeros(cnt) if rate is None else np.cumsum(gaps_unit) / rate * 1e6
        span, lat = srv.replay(arr, ops, ke, ve, pa, pb, ns)
        st = srv.stats()
        ok = bool((lat > 0).all())
        return {"offered_req_per_s": None if rate is None else round(rate, 1),
                "throughput_req_per_s": round(cnt / span, 1), "throughput_gelem_per_s": round(elems / span / 1e9, 3),
                "p50_us": round(float(np.percentile(lat, 50)), 1), "p99_us": round(float(np.percentile(lat, 99)), 1),
                "mean_us": round(float(lat.mean()), 1), "span_s": round(span, 4), "all_ok": ok,
                "mean_batch": st["mean_batch"], "peak_held_bytes": st["peak_held_bytes"]}

    per_call_sat = replay(False, None)
    log(f"served per-call saturation: {per_call_sat}")
    batched_sat = replay(True, None)
    log(f"served batched saturation: {batched_sat}")
    base = per_call_sat["throughput_req_per_s"]
    loads = {}
    for f in (0.5, 1.0, 2.0):
        loads[f"{f}x"] = replay(True, f * base)
        log(f"served batched at {f}x per-call rate: {loads[f'{f}x']}")
    per_call_half = replay(False, 0.5 * base)
    log(f"served per-call at 0.5x: {per_call_half}")
    return {"requests": cnt, "mix": "60% Sort_By_Key u32/u32, 30% Inclusive_Scan i32, 10% Reduce f32",
            "sizes": "2^U(10,24)", "budget_bytes": budget, "slots": 3,
            "per_call_saturation": per_call_sat, "batched_saturation": batched_sat,
            "throughput_ratio_batched_vs_per_call": round(batched_sat["throughput_req_per_s"] / base, 2),
            "batched_at_load": loads, "per_call_at_0.5x": per_call_half}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=28, help="log2 of the number of pairs")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-variants", action="store_true")
    ap.add_argument("--no-single-pass", action="store_true", help="skip the 2^29 scan / reduce / copy_if leg")
    ap.add_argument("--no-small", action="store_true", help="skip the 2^20 latency leg")
    ap.add_argument("--no-segmented", action="store_true", help="skip the batched-segments leg")
    ap.add_argument("--segments", type=int, default=4096, help="segments of the batched-segments leg")
    ap.add_argument("--no-served", action="store_true", help="skip the served-stream leg")
    ap.add_argument("--served-requests", type=int, default=3000, help="requests per served-stream run")
    ap.add_argument("--no-consumers", action="store_true", help="skip the scan / compaction consumers leg")
    ap.add_argument("--only", choices=["sort", "single_pass", "small", "segmented", "served", "consumers"],
                    default=None,
                    help="run one leg only (the headline fields are then absent)")
    ap.add_argument("--mode", choices=["sort", "copy"], default="sort",
                    help="copy: a plain device copy of a known size (4 B/elem read + 4 B/elem write), "
                         "used to calibrate HBM byte counters")
    a = ap.parse_args()
    if a.gpus != 1:
        log("bench.py measures one GPU only: the primitive core runs on a single MI355X "
            "(multi-GPU sharding is a design note, not a measured path); use --gpus 1")
        sys.exit(2)

    if not os.path.exists(os.path.join(ROOT, "thrustrtc_amd", "libPyThrustRTC.so")):
        import __graft_entry__
        __graft_entry__.build()

    import numpy as np
    import thrustrtc_amd as trtc
    from thrustrtc_amd.native import native
    from thrustrtc_amd import workloads as wl

    N = 1 << a.n
    K, W = a.steps, a.warmup

    if a.mode == "copy":
        print(json.dumps(dict(mode="copy", **leg_copy(trtc, N, K, W))), flush=True)
        return

    if a.only and a.only != "sort":
        if a.only == "single_pass":
            res = leg_single_pass(trtc, native, wl, a, log)
        elif a.only == "small":
            res = leg_small_calls(trtc, native, wl, log, not a.no_cpu)
        elif a.only == "served":
            res = leg_served(trtc, native, wl, a, log)
        elif a.only == "consumers":
            res = leg_consumers(trtc, native, wl, a, log)
        else:
            res = leg_segmented(trtc, native, wl, a, log)
        print(json.dumps({"leg": a.only, "result": res}), flush=True)
        return

    def make_copies(keys, vals, count):
        src_k = trtc.device_vector_from_numpy(keys)
        src_v = trtc.device_vector_from_numpy(vals)
        ks, vs = [], []
        for _ in range(count):
            dk = trtc.device_vector("uint32_t" if keys.dtype == np.uint32 else "float", N)
            dv = trtc.device_vector("uint32_t", N)
            trtc.Copy(src_k, dk)
            trtc.Copy(src_v, dv)
            ks.append(dk)
            vs.append(dv)
        trtc.Wait()
        return ks, vs

    def run(keys, steps, warmup, check=False, profile=False):
        vals = np.arange(N, dtype=np.uint32)
        ks, vs = make_copies(keys, vals, steps + warmup)
        for i in range(warmup):
            trtc.Sort_By_Key(ks[i], vs[i])
        trtc.Wait()
        if profile:
            native.n_profile_begin()
        t0 = time.perf_counter()
        for i in range(warmup, warmup + steps):
            trtc.Sort_By_Key(ks[i], vs[i])
        trtc.Wait()
        t1 = time.perf_counter()
        prof = None
        if profile:
            buf = __import__("ctypes").create_string_buffer(1 << 16)
            native.n_profile_end(buf, 1 << 16)
            prof = json.loads(buf.value.decode())
        ok = None
        if check:
            gk = ks[-1].to_host()
            gv = vs[-1].to_host()
            ok = bool((gk.view(np.uint32) == keys[gv].view(np.uint32)).all())
            if keys.dtype == np.uint32:
                ok = ok and bool((gk[1:] >= gk[:-1]).all())
                # stable: equal keys keep increasing payload
                eq = gk[1:] == gk[:-1]
                ok = ok and bool((gv[1:][eq] > gv[:-1][eq]).all())
        del ks, vs
        return (t1 - t0) / steps, prof, ok

    log(f"generating 2^{a.n} uniform uint32 keys (mt19937 seed 2)")
    keys = wl.u32_uniform(N, 2)
    sec, prof, ok = run(keys, K, W, check=True, profile=True)
    if not ok:
        log("ERROR: sorted output failed the sortedness/stability/permutation check")
        sys.exit(1)
    ms = sec * 1e3
    gkeys = N / sec / 1e9
    log(f"Sort_By_Key 2^{a.n} u32 uniform: {ms:.3f} ms/step, {gkeys:.2f} Gkeys/s, kernels: {prof}")

    # dominant kernel: one onesweep digit pass (reads 8 B/pair, writes 8 B/pair);
    # from 2^23 keys (one tile per CU) the long-tile pass, below it the 1024x16 pass
    pass_name = max((k for k in prof if k.startswith(("trtc_radix_pass_long_k<4,", "trtc_radix_pass_k<4, 4,"))),
                    key=lambda k: prof[k]["total_ms"])
    pass_ms = prof[pass_name]["total_ms"] / prof[pass_name]["count"]
    pass_bytes = 16.0 * N
    achieved = pass_bytes / (pass_ms * 1e-3) / 1e9
    model_gbs = 68.0 * N / sec / 1e9

    # HBM bytes per launch of the same kernel, from the committed rocprofv3
    # FETCH_SIZE / WRITE_SIZE runs (scaled by a copy kernel of known size);
    # only valid for the workload they were measured on
    traffic = None
    tfile = os.path.join(ROOT, "profiles", "radix_pass_traffic.json")
    if os.path.exists(tfile):
        t = json.load(open(tfile))
        if t.get("n") == N and pass_name == t.get("kernel"):
            traffic = round(t["read_bytes"] + t["write_bytes"])

    # the per-GPU shape of an 8-way sharded 2^28 sort: 2^25 pairs
    shard = None
    if a.n == 28 and not a.no_variants:
        ks = wl.u32_uniform(1 << 25, 2)
        N_save = N
        N = 1 << 25
        s3, _, ok3 = run(ks, K, W, check=True)
        N = N_save
        shard = {"n": 1 << 25, "gkeys_per_s": round((1 << 25) / s3 / 1e9, 3), "ms_per_step": round(s3 * 1e3, 4),
                 "pct_hbm_68N": round(100 * 68.0 * (1 << 25) / s3 / 1e9 / HBM_PEAK_GBS, 2), "bit_exact_check": ok3}
        log(f"shard 2^25: {shard}")
        del ks

    variants = {}
    if not a.no_variants:
        for name, gen in [("u32_low_entropy", lambda: wl.u32_low_entropy(N, 2)),
                          ("f32_normal_signed_zeros", lambda: wl.f32_normal_signed_zeros(N, 3, 4))]:
            kv = gen()
            s2, _, ok2 = run(kv, max(2, K // 2), 1, check=True)
            variants[name] = {"gkeys_per_s": round(N / s2 / 1e9, 3), "ms_per_step": round(s2 * 1e3, 4),
                              "pct_hbm_68N": round(100 * 68.0 * N / s2 / 1e9 / HBM_PEAK_GBS, 2),
                              "bit_exact_check": ok2}
            log(f"variant {name}: {variants[name]}")
            del kv

    cpu = None
    if not a.no_cpu:
        from helpers import Oracle
        orc = Oracle()
        ck = keys.copy()
        cv = np.arange(N, dtype=np.uint32)
        t0 = time.perf_counter()
        orc.L.bl_sort_pairs_u32(ck.ctypes.data, cv.ctypes.data, N)
        dt = time.perf_counter() - t0
        cpu = {"value": round(N / dt / 1e9, 4), "unit": "Gkeys/s", "cores": int(orc.L.bl_threads()),
               "kind": "port",
               "sample": f"the full workload: one __gnu_parallel::stable_sort of 2^{a.n} (key, index) pairs "
                         f"(same keys), {dt:.2f} s"}
        log(f"cpu baseline: {cpu}")

    segmented = None if a.no_segmented else leg_segmented(trtc, native, wl, a, log)
    served = None if a.no_served else leg_served(trtc, native, wl, a, log)
    single_pass = None if a.no_single_pass else leg_single_pass(trtc, native, wl, a, log)
    small = None if a.no_small else leg_small_calls(trtc, native, wl, log, not a.no_cpu)
    consumers = None if a.no_consumers else leg_consumers(trtc, native, wl, a, log)
    copy_bw = leg_copy(trtc, N, 5, 2)
    log(f"copy bandwidth: {copy_bw}")

    line = {
        "metric": "Gkeys/s Sort_By_Key 2^28 u32 + %HBM roofline; served-mix p99 vs per-call",
        "value": round(gkeys, 4),
        "unit": "Gkeys/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": {"workload": f"TRTC_Sort_By_Key, 2^{a.n} uint32 keys uniform over [0,2^32) (mt19937 seed 2) "
                               f"+ uint32 original-index payload, ascending (Less), stable",
                   "n": N, "key_bits": 32, "value_bits": 32, "digit_bits": 8, "passes": 4},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "kernel": pass_name, "kernel_avg_ms": round(pass_ms, 4),
                     "bytes_per_launch": pass_bytes},
        "model_68N": {"gbs": round(model_gbs, 1), "pct_hbm": round(100 * model_gbs / HBM_PEAK_GBS, 2)},
        "kernels": prof,
        "variants": variants,
        "shard_2p25": shard,
        "cpu_baseline": cpu,
        "single_pass_2p29": single_pass,
        "batched_segments": segmented,
        "served_stream": served,
        "small_calls_2p20": small,
        "consumers_2p28": consumers,
        "device_copy": copy_bw,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
