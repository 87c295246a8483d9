"""Shared test helpers: load the native pieces by path (no GPU needed), build
them on demand, and give GPU tests the package."""
import ctypes
import importlib.util
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "thrustrtc_amd")
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_LIB = os.path.join(ORACLE_DIR, "liboracle.so")


def _load_path(name, path):
    spec = importlib.util.spec_from_file_location(name, path)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def ensure_built():
    lib = os.path.join(PKG, "libPyThrustRTC.so")
    wl = os.path.join(PKG, "libtrtc_workloads.so")
    if not (os.path.exists(lib) and os.path.exists(wl)):
        b = _load_path("trtc_build", os.path.join(PKG, "build.py"))
        b.build()
    if not os.path.exists(ORACLE_LIB):
        subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)


def native_module():
    ensure_built()
    return _load_path("trtc_native_standalone", os.path.join(PKG, "native.py"))


def oracle_algorithms():
    """oracle/algorithms.py (the CPU restatement of the other algorithms)."""
    return _load_path("trtc_oracle_algorithms", os.path.join(ROOT, "oracle", "algorithms.py"))


def workloads():
    ensure_built()
    return _load_path("trtc_workloads_standalone", os.path.join(PKG, "workloads.py"))


# ---------------------------------------------------------------- oracle
TC = {"int8_t": 0, "uint8_t": 1, "int16_t": 2, "uint16_t": 3, "int32_t": 4, "uint32_t": 5,
      "int64_t": 6, "uint64_t": 7, "float": 8, "double": 9}
NP = {"int8_t": np.int8, "uint8_t": np.uint8, "int16_t": np.int16, "uint16_t": np.uint16,
      "int32_t": np.int32, "uint32_t": np.uint32, "int64_t": np.int64, "uint64_t": np.uint64,
      "float": np.float32, "double": np.float64}
OPS = {"Plus": 0, "Maximum": 1, "Minimum": 2, "Multiplies": 3}


class Oracle:
    """ctypes front of oracle/liboracle.so (the CPU restatement)."""

    def __init__(self):
        ensure_built()
        L = ctypes.CDLL(ORACLE_LIB)
        vp, u64, i = ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int
        L.or_sort.argtypes = [i, vp, u64, i]
        L.or_sort_by_key.argtypes = [i, vp, vp, u64, u64, i]
        L.or_sort_perm.argtypes = [i, vp, u64, i, vp]
        L.or_inclusive_scan.argtypes = [i, vp, vp, u64, i]
        L.or_exclusive_scan.argtypes = [i, vp, vp, u64, vp, i]
        L.or_scan_by_key.argtypes = [i, i, vp, vp, vp, u64, i, vp, i]
        L.or_reduce.argtypes = [i, vp, u64, i, vp, i, vp]
        L.or_sum_f32_fp64.argtypes = [vp, u64]
        L.or_sum_f32_fp64.restype = ctypes.c_double
        L.or_reduce_ref_tree_f32.argtypes = [vp, u64, i, ctypes.c_float]
        L.or_reduce_ref_tree_f32.restype = ctypes.c_float
        L.or_copy_if_mod_i32.argtypes = [vp, vp, u64, ctypes.c_int32, ctypes.c_int32]
        L.or_copy_if_mod_i32.restype = u64
        L.or_copy_if_stencil_mod_i32.argtypes = [vp, vp, vp, u64, ctypes.c_int32, ctypes.c_int32]
        L.or_copy_if_stencil_mod_i32.restype = u64
        i64 = ctypes.c_int64
        L.or_unique.argtypes = [i, vp, vp, vp, vp, u64, u64, i64]
        L.or_unique.restype = u64
        L.or_reduce_by_key.argtypes = [i, i, vp, vp, vp, vp, u64, i, i64]
        L.or_reduce_by_key.restype = u64
        L.or_select.argtypes = [i, vp, vp, vp, u64, i, i64, i64, i]
        L.or_select.restype = u64
        L.or_partition.argtypes = [i, vp, vp, vp, vp, u64, i, i64, i64]
        L.or_partition.restype = u64
        L.or_transform_scan.argtypes = [i, vp, vp, u64, i, i, i, vp]
        L.or_count.argtypes = [i, vp, u64, i, i64, i64]
        L.or_count.restype = u64
        L.bl_threads.restype = i
        L.bl_sort_pairs_u32.argtypes = [vp, vp, u64]
        L.bl_sort_u32.argtypes = [vp, u64]
        L.bl_inclusive_scan_i32.argtypes = [vp, vp, u64]
        L.bl_reduce_f32.argtypes = [vp, u64]
        L.bl_reduce_f32.restype = ctypes.c_float
        L.bl_copy_if_mod3_i32.argtypes = [vp, vp, u64]
        L.bl_copy_if_mod3_i32.restype = u64
        L.bl_segmented_sort_pairs_u32.argtypes = [vp, vp, vp, u64]
        L.bl_segmented_inclusive_scan_i32.argtypes = [vp, vp, vp, u64]
        self.L = L

    def sort(self, keys, elem, descending=False):
        a = np.array(keys, dtype=NP[elem])
        assert self.L.or_sort(TC[elem], a.ctypes.data, len(a), int(descending)) == 0
        return a

    def sort_by_key(self, keys, values, elem, descending=False):
        k = np.array(keys, dtype=NP[elem])
        v = np.ascontiguousarray(np.array(values))
        assert self.L.or_sort_by_key(TC[elem], k.ctypes.data, v.ctypes.data, v.itemsize, len(k),
                                     int(descending)) == 0
        return k, v

    def sort_perm(self, keys, elem, descending=False):
        k = np.ascontiguousarray(np.array(keys, dtype=NP[elem]))
        p = np.empty(len(k), dtype=np.uint64)
        assert self.L.or_sort_perm(TC[elem], k.ctypes.data, len(k), int(descending), p.ctypes.data) == 0
        return p

    def inclusive_scan(self, x, elem, op="Plus"):
        a = np.ascontiguousarray(np.array(x, dtype=NP[elem]))
        out = np.empty_like(a)
        assert self.L.or_inclusive_scan(TC[elem], a.ctypes.data, out.ctypes.data, len(a), OPS[op]) == 0
        return out

    def exclusive_scan(self, x, elem, init=0, op="Plus"):
        a = np.ascontiguousarray(np.array(x, dtype=NP[elem]))
        out = np.empty_like(a)
        iv = np.array([init], dtype=NP[elem])
        assert self.L.or_exclusive_scan(TC[elem], a.ctypes.data, out.ctypes.data, len(a), iv.ctypes.data,
                                        OPS[op]) == 0
        return out

    def scan_by_key(self, keys, vals, kelem, velem, exclusive=False, init=0, op="Plus"):
        k = np.ascontiguousarray(np.array(keys, dtype=NP[kelem]))
        v = np.ascontiguousarray(np.array(vals, dtype=NP[velem]))
        out = np.empty_like(v)
        iv = np.array([init], dtype=NP[velem])
        assert self.L.or_scan_by_key(TC[kelem], TC[velem], k.ctypes.data, v.ctypes.data, out.ctypes.data,
                                     len(k), int(exclusive), iv.ctypes.data, OPS[op]) == 0
        return out

    def reduce(self, x, elem, init=None, op="Plus"):
        a = np.ascontiguousarray(np.array(x, dtype=NP[elem]))
        iv = np.array([0 if init is None else init], dtype=NP[elem])
        r = np.zeros(1, dtype=NP[elem])
        assert self.L.or_reduce(TC[elem], a.ctypes.data, len(a), int(init is not None), iv.ctypes.data,
                                OPS[op], r.ctypes.data) == 0
        return r[0]

    def sum_f32_fp64(self, x):
        a = np.ascontiguousarray(x, dtype=np.float32)
        return self.L.or_sum_f32_fp64(a.ctypes.data, len(a))

    def reduce_ref_tree_f32(self, x, init=None):
        a = np.ascontiguousarray(x, dtype=np.float32)
        return self.L.or_reduce_ref_tree_f32(a.ctypes.data, len(a), int(init is not None),
                                             0.0 if init is None else init)

    def copy_if_mod(self, x, m, r):
        a = np.ascontiguousarray(np.array(x, dtype=np.int32))
        out = np.empty_like(a)
        k = self.L.or_copy_if_mod_i32(a.ctypes.data, out.ctypes.data, len(a), m, r)
        return out[:k]

    def copy_if_stencil_mod(self, x, st, m, r):
        a = np.ascontiguousarray(np.array(x, dtype=np.int32))
        s = np.ascontiguousarray(np.array(st, dtype=np.int32))
        out = np.empty_like(a)
        k = self.L.or_copy_if_stencil_mod_i32(a.ctypes.data, s.ctypes.data, out.ctypes.data, len(a), m, r)
        return out[:k]

    # ---- scan / compaction consumers.  Predicates (pmode): ("mod", a, b) is
    # v % a == b, ("eq", a) is v == a, ("nonzero",) is v != 0.
    @staticmethod
    def _pred(p):
        if p[0] == "mod":
            return 0, p[1], p[2]
        if p[0] == "eq":
            return 1, p[1], 0
        return 2, 0, 0

    def unique(self, keys, kelem, vals=None, div=1):
        k = np.ascontiguousarray(np.array(keys, dtype=NP[kelem]))
        ko = np.empty_like(k)
        v = None if vals is None else np.ascontiguousarray(np.array(vals))
        vo = None if v is None else np.empty_like(v)
        c = self.L.or_unique(TC[kelem], k.ctypes.data, ko.ctypes.data, None if v is None else v.ctypes.data,
                             None if vo is None else vo.ctypes.data, 0 if v is None else v.itemsize, len(k), div)
        return (ko[:c],) if v is None else (ko[:c], vo[:c])

    def reduce_by_key(self, keys, vals, kelem, velem, op="Plus", div=1):
        k = np.ascontiguousarray(np.array(keys, dtype=NP[kelem]))
        v = np.ascontiguousarray(np.array(vals, dtype=NP[velem]))
        ko, vo = np.empty_like(k), np.empty_like(v)
        c = self.L.or_reduce_by_key(TC[kelem], TC[velem], k.ctypes.data, v.ctypes.data, ko.ctypes.data,
                                    vo.ctypes.data, len(k), OPS[op], div)
        assert c != 2 ** 64 - 1
        return ko[:c], vo[:c]

    def select(self, x, elem, pred, stencil=None, negate=False):
        a = np.ascontiguousarray(np.array(x, dtype=NP[elem]))
        st = None if stencil is None else np.ascontiguousarray(np.array(stencil, dtype=np.int32))
        out = np.empty_like(a)
        m, pa, pb = self._pred(pred)
        c = self.L.or_select(TC[elem], a.ctypes.data, None if st is None else st.ctypes.data, out.ctypes.data,
                             len(a), m, pa, pb, int(negate))
        return out[:c]

    def partition(self, x, elem, pred, stencil=None):
        a = np.ascontiguousarray(np.array(x, dtype=NP[elem]))
        st = None if stencil is None else np.ascontiguousarray(np.array(stencil, dtype=np.int32))
        t, f = np.empty_like(a), np.empty_like(a)
        m, pa, pb = self._pred(pred)
        c = self.L.or_partition(TC[elem], a.ctypes.data, None if st is None else st.ctypes.data, t.ctypes.data,
                                f.ctypes.data, len(a), m, pa, pb)
        return t[:c], f[:len(a) - c]

    def transform_scan(self, x, elem, unary, op="Plus", exclusive=False, init=0):
        a = np.ascontiguousarray(np.array(x, dtype=NP[elem]))
        out = np.empty_like(a)
        iv = np.array([init], dtype=NP[elem])
        u = {"Identity": 0, "Negate": 1, "Square": 2}[unary]
        assert self.L.or_transform_scan(TC[elem], a.ctypes.data, out.ctypes.data, len(a), u, OPS[op],
                                        int(exclusive), iv.ctypes.data) == 0
        return out

    def count(self, x, elem, pred):
        a = np.ascontiguousarray(np.array(x, dtype=NP[elem]))
        m, pa, pb = self._pred(pred)
        return self.L.or_count(TC[elem], a.ctypes.data, len(a), m, pa, pb)


def gpu_available():
    """True when a HIP device is visible (probed in a child process so that
    this process does not initialise the GPU)."""
    code = "import ctypes,sys;L=ctypes.CDLL('libamdhip64.so');c=ctypes.c_int(0);sys.exit(0 if L.hipGetDeviceCount(ctypes.byref(c))==0 and c.value>0 else 1)"
    try:
        return subprocess.run(["python3", "-c", code], timeout=120).returncode == 0
    except Exception:
        return False
