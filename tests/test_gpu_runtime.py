"""GPU tests of the runtime around the kernels: the on-disk code-object cache
across processes (reference: internal/impl_context.inl:619-651, a kernel
cache keyed by source hash and target), kernel-debug mode (every launch is
followed by a synchronisation, impl_context.inl:783-797) and the occupancy
helpers behind n_kernel_calc_optimal_block_size / n_kernel_calc_number_blocks
(api.h:35-36, launch_calc.cpp:73-115)."""
import json
import os
import subprocess
import sys
import time

import numpy as np
import pytest

from helpers import ROOT

pytestmark = pytest.mark.gpu

CHILD = r'''
import ctypes, json, sys
sys.path.insert(0, sys.argv[1])
import numpy as np
import thrustrtc_amd as trtc
from thrustrtc_amd.native import native
x = np.arange(-50000, 50000, dtype=np.int32)
d = trtc.device_vector_from_numpy(x)
o = trtc.device_vector("int32_t", len(x))
pred = trtc.Functor({}, ["v"], "        return v % 7 == 3 || v % 11 == -2;\n")
c = trtc.Copy_If(d, o, pred)
st = (ctypes.c_ulonglong * 4)()
native.n_jit_cache_stats(st)
print(json.dumps({"count": int(c), "out": o.to_host()[:c].tolist(), "stats": list(st)}), flush=True)
'''


def say(*a):
    print(*a, flush=True)


def run_child(cache_dir):
    env = dict(os.environ, TRTC_CACHE_DIR=cache_dir)
    r = subprocess.run([sys.executable, "-c", CHILD, ROOT], env=env, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    return json.loads(r.stdout.strip().splitlines()[-1])


def cache_files(cache_dir):
    out = []
    for dp, _, fs in os.walk(cache_dir):
        out += [os.path.join(dp, f) for f in fs if f.endswith(".co")]
    return sorted(out)


def test_disk_cache_across_processes(tmp_path):
    """A fresh process compiles a user-functor Copy_If and writes
    <arch>/<key>.co; a second fresh process loads it from disk (no hipRTC
    compilation) with identical results; damaged entries (truncated to the
    name line, a flipped code byte) are rejected on the host, recompiled and
    rewritten, and the rewritten entry is used again."""
    cache = str(tmp_path / "cache")
    say("cache: first process (cold)")
    a = run_child(cache)
    x = np.arange(-50000, 50000, dtype=np.int32)
    exp = x[(np.fmod(x, 7) == 3) | (np.fmod(x, 11) == -2)]  # C++ % truncates
    assert a["count"] == len(exp) and a["out"] == exp.tolist()
    files = cache_files(cache)
    assert files and a["stats"][2] >= 1 and a["stats"][2] == len(files), (a["stats"], files)
    assert all(os.path.basename(os.path.dirname(f)).startswith("gfx950") for f in files)

    say("cache: second process (warm)")
    b = run_child(cache)
    assert b["out"] == a["out"]
    assert b["stats"][2] == 0 and b["stats"][1] == len(files) and b["stats"][3] == 0, b["stats"]

    say("cache: damaged entries")
    blob = open(files[0], "rb").read()
    open(files[0], "wb").write(blob[:blob.index(b"\n") + 1])  # header / name line only
    if len(files) > 1:
        blob1 = bytearray(open(files[1], "rb").read())
        blob1[-64] ^= 0x40
        open(files[1], "wb").write(bytes(blob1))
    damaged = min(2, len(files))
    c = run_child(cache)
    assert c["out"] == a["out"]
    assert c["stats"][3] == damaged and c["stats"][2] == damaged, c["stats"]
    d = run_child(cache)
    assert d["out"] == a["out"] and d["stats"][2] == 0 and d["stats"][3] == 0, d["stats"]


@pytest.fixture(scope="module")
def trtc():
    import thrustrtc_amd
    yield thrustrtc_amd
    thrustrtc_amd.Set_Kernel_Debug(False)


SPIN = r'''
    // bounded busy wait of ~ticks steady-clock ticks (also capped in iterations)
    unsigned long long t0 = wall_clock64();
    for (unsigned i = 0; i < 400000000u; i++) {
        if (wall_clock64() - t0 >= ticks) break;
    }
    out[0] = 1;
'''


def test_kernel_debug_synchronises_after_each_launch(trtc):
    """With kernel debug on, a launch returns only after the kernel finished;
    off, it returns at once.  Launch errors are reported by the launching call
    in both modes."""
    say("kernel debug mode")
    k = trtc.Kernel(["ticks", "out"], SPIN)
    out = trtc.device_vector("int32_t", 1)
    trtc.Set_Kernel_Debug(False)
    k.launch(1, 64, [trtc.DVUInt64(1000), out])  # compiles the kernel outside the timed calls
    trtc.Wait()
    ticks = 100_000_000 // 4  # 0.25 s at the 100 MHz steady clock of this GPU family
    trtc.Set_Kernel_Debug(False)
    t0 = time.perf_counter()
    k.launch(1, 64, [trtc.DVUInt64(ticks), out])
    t_async = time.perf_counter() - t0
    trtc.Wait()
    t_total = time.perf_counter() - t0
    trtc.Set_Kernel_Debug(True)
    t0 = time.perf_counter()
    k.launch(1, 64, [trtc.DVUInt64(ticks), out])
    t_debug = time.perf_counter() - t0
    trtc.Set_Kernel_Debug(False)
    say(f"  async launch {t_async * 1e3:.2f} ms, kernel {t_total * 1e3:.1f} ms, debug launch {t_debug * 1e3:.1f} ms")
    assert t_total > 0.1, "the spin kernel did not wait"
    assert t_async < 0.5 * t_total
    assert t_debug > 0.8 * t_total
    # a launch error (more threads per block than the device allows) fails the call, debug on or off
    for dbg in (False, True):
        trtc.Set_Kernel_Debug(dbg)
        with pytest.raises(SystemError):
            k.launch(1, 2048, [trtc.DVUInt64(1), out])
    trtc.Set_Kernel_Debug(False)
    k.launch(1, 64, [trtc.DVUInt64(1), out])  # the error is not sticky
    trtc.Wait()


def test_occupancy_helpers(trtc):
    """calc_optimal_block_size: a whole number of 64-lane waves, at most 1024;
    calc_number_blocks: resident blocks per CU times the CU count, shrinking
    as the per-block LDS request grows (160 KB of LDS per CU)."""
    say("occupancy helpers")
    k = trtc.Kernel(["v"], "    unsigned i = blockIdx.x * blockDim.x + threadIdx.x;\n    if (i < v.size()) v[i] += 1;\n")
    v = trtc.device_vector("int32_t", 1 << 16)
    bs = k.calc_optimal_block_size([v])
    assert 64 <= bs <= 1024 and bs % 64 == 0, bs
    cus = 256
    nb0 = k.calc_number_blocks([v], 256)
    assert nb0 >= cus and nb0 % cus == 0, nb0
    nb_big = k.calc_number_blocks([v], 256, 96 * 1024)
    assert nb_big == cus, nb_big  # one 96 KB block per CU
    nb_mid = k.calc_number_blocks([v], 256, 48 * 1024)
    assert nb_mid == 3 * cus, nb_mid
    assert nb_big <= nb_mid <= nb0
    # the reported grid runs: one pass over v with that many blocks
    v2 = trtc.device_vector_from_numpy(np.zeros(1 << 16, dtype=np.int32))
    k.launch((1 << 16) // 256, 256, [v2])
    assert (v2.to_host() == 1).all()
