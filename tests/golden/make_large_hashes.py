"""Writes tests/golden/large_hashes.json: 64-bit FNV-1a hashes of the
generated inputs and of the oracle's outputs at the full workload sizes
(2^20 sort and scan, 2^28 Sort_By_Key in three key distributions, 2^29 scan,
reduce and copy_if).  GPU tests compare the HIP results against these hashes,
so the full-size parity check needs neither the oracle's run time nor the
reference on the machine that runs them.  Run from the repository root:

    python tests/golden/make_large_hashes.py            # large_hashes.json
    python tests/golden/make_large_hashes.py --batched  # batched_segments_hashes.json

The first takes several minutes on 8 host cores (the 2^28 stable sorts
dominate).  The second writes, for the batched-segments workload (4096
segments, sizes log-uniform in [16, 2^20] from mt19937_64 seed 7, uint32 keys
uniform from mt19937 seed 2 with the uint32 index as payload, int32 values
uniform in [-1000, 1000] from mt19937 seed 5), the hash of every segment's
stable Sort_By_Key keys and payload and of its Inclusive_Scan, each segment
computed alone by the oracle.
"""
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
from helpers import Oracle, workloads  # noqa: E402

orc, wl = Oracle(), workloads()


def h(a):
    a = np.ascontiguousarray(a)
    return "%016x" % wl._lib.wl_fnv1a64(a.ctypes.data, a.nbytes)


def sort_case(keys, elem):
    perm = orc.sort_perm(keys, elem)
    out_k = keys[perm]
    out_v = perm.astype(np.uint32)
    return {"n": len(keys), "keys_in": h(keys), "keys_out": h(out_k), "values_out": h(out_v)}


def main():
    res = {}
    t0 = time.time()

    k = wl.u32_uniform(1 << 20, 0)
    res["sort_u32_2p20_uniform_seed0"] = {"n": 1 << 20, "keys_in": h(k), "keys_out": h(orc.sort(k, "uint32_t"))}
    x = wl.i32_range(1 << 20, 0, -100, 100)
    res["inclusive_scan_i32_2p20_seed0"] = {"n": 1 << 20, "in": h(x), "out": h(orc.inclusive_scan(x, "int32_t"))}
    print("2^20 cases", time.time() - t0, flush=True)

    n2 = 1 << 28
    res["sort_by_key_u32_2p28_uniform_seed2"] = sort_case(wl.u32_uniform(n2, 2), "uint32_t")
    print("uniform", time.time() - t0, flush=True)
    res["sort_by_key_u32_2p28_low_entropy_seed2"] = sort_case(wl.u32_low_entropy(n2, 2), "uint32_t")
    print("low entropy", time.time() - t0, flush=True)
    res["sort_by_key_f32_2p28_normal_seed3_zeros4"] = sort_case(wl.f32_normal_signed_zeros(n2, 3, 4), "float")
    print("f32", time.time() - t0, flush=True)

    n3 = 1 << 29
    x = wl.i32_range(n3, 5, -1000, 1000)
    res["inclusive_scan_i32_2p29_seed5"] = {"n": n3, "in": h(x), "out": h(orc.inclusive_scan(x, "int32_t"))}
    c = orc.copy_if_mod(x, 3, 0)
    res["copy_if_mod3_i32_2p29_seed5"] = {"n": n3, "in": h(x), "count": int(len(c)), "out": h(c)}
    del x, c
    f = wl.f32_unit(n3, 6)
    res["reduce_f32_2p29_seed6"] = {"n": n3, "in": h(f), "sum_fp64": orc.sum_f32_fp64(f)}
    print("2^29 cases", time.time() - t0, flush=True)

    with open(os.path.join(HERE, "large_hashes.json"), "w") as fo:
        json.dump(res, fo, indent=1, sort_keys=True)


def batched():
    t0 = time.time()
    sizes, total = wl.segment_sizes(4096, 7, 16, 1 << 20)
    offs = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint64)
    keys = wl.u32_uniform(total, 2)
    x = wl.i32_range(total, 5, -1000, 1000)
    segs = []
    for s in range(len(sizes)):
        lo, hi = int(offs[s]), int(offs[s + 1])
        perm = orc.sort_perm(keys[lo:hi], "uint32_t")
        segs.append(" ".join([h(keys[lo:hi][perm]), h((perm + lo).astype(np.uint32)),
                              h(orc.inclusive_scan(x[lo:hi], "int32_t"))]))
    res = {"segments": len(sizes), "total": total, "sizes": h(sizes), "keys_in": h(keys), "values_in": h(x),
           "per_segment": "keys_out values_out(global index payload) inclusive_scan_out",
           "hashes": segs}
    print("batched segments", time.time() - t0, flush=True)
    with open(os.path.join(HERE, "batched_segments_hashes.json"), "w") as fo:
        json.dump(res, fo, indent=0, separators=(",", ": "))


if __name__ == "__main__":
    if "--batched" in sys.argv[1:]:
        batched()
    else:
        main()
