"""Deterministic synthetic inputs of the benchmark workloads (ctypes over
libtrtc_workloads.so; std::mt19937-based, see csrc/workloads/workloads.cpp).

No package-relative imports: loadable by path without initialising a GPU.
"""
import ctypes
import os
import numpy as np

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libtrtc_workloads.so")


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is not built: run `python thrustrtc_amd/build.py`")
    lib = ctypes.CDLL(LIB_PATH)
    u32, u64, i32, f64, vp = ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int32, ctypes.c_double, ctypes.c_void_p
    lib.wl_gen_u32_uniform.argtypes = [u32, u64, vp]
    lib.wl_gen_u32_low_entropy.argtypes = [u32, u64, vp]
    lib.wl_gen_f32_normal_signed_zeros.argtypes = [u64, u32, ctypes.c_int, u64, vp]
    lib.wl_gen_i32_range.argtypes = [u32, i32, i32, u64, vp]
    lib.wl_gen_f32_unit.argtypes = [u32, u64, vp]
    lib.wl_gen_segment_sizes.argtypes = [u64, u64, u64, u64, vp]
    lib.wl_gen_segment_sizes.restype = u64
    lib.wl_gen_exponential.argtypes = [u64, f64, u64, vp]
    lib.wl_gen_unit_f64.argtypes = [u64, u64, vp]
    lib.wl_fnv1a64.argtypes = [vp, u64]
    lib.wl_fnv1a64.restype = u64
    return lib


_lib = _load()


def u32_uniform(n, seed):
    out = np.empty(n, dtype=np.uint32)
    _lib.wl_gen_u32_uniform(seed, n, out.ctypes.data)
    return out


def u32_low_entropy(n, seed):
    out = np.empty(n, dtype=np.uint32)
    _lib.wl_gen_u32_low_entropy(seed, n, out.ctypes.data)
    return out


def f32_normal_signed_zeros(n, seed_normal, seed_zero, plant_pos_zero=True):
    out = np.empty(n, dtype=np.float32)
    _lib.wl_gen_f32_normal_signed_zeros(seed_normal, seed_zero, 1 if plant_pos_zero else 0, n, out.ctypes.data)
    return out


def i32_range(n, seed, lo, hi):
    out = np.empty(n, dtype=np.int32)
    _lib.wl_gen_i32_range(seed, lo, hi, n, out.ctypes.data)
    return out


def f32_unit(n, seed):
    out = np.empty(n, dtype=np.float32)
    _lib.wl_gen_f32_unit(seed, n, out.ctypes.data)
    return out


def segment_sizes(count, seed, lo, hi):
    out = np.empty(count, dtype=np.uint64)
    total = _lib.wl_gen_segment_sizes(seed, lo, hi, count, out.ctypes.data)
    return out, int(total)


def exponential(n, seed, rate):
    out = np.empty(n, dtype=np.float64)
    _lib.wl_gen_exponential(seed, rate, n, out.ctypes.data)
    return out


def unit_f64(n, seed):
    out = np.empty(n, dtype=np.float64)
    _lib.wl_gen_unit_f64(seed, n, out.ctypes.data)
    return out


def fnv1a64(arr):
    arr = np.ascontiguousarray(arr)
    return int(_lib.wl_fnv1a64(arr.ctypes.data, arr.nbytes))
