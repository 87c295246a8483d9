"""Hot-path algorithms with the reference's Python signatures
(python/ThrustRTC/Sorting.py, PrefixSums.py, Reductions.py, Reordering.py,
Copying.py, Transformations.py)."""
import ctypes
import numpy as np
from .native import native, check_i, check_u, check_cptr
from .DeviceViewable import ValueFromDVPtr


def _p(obj):
    return None if obj is None else obj.m_cptr


# ---- Sorting.py
def Sort(vec, comp=None):
    check_i(native.n_sort(vec.m_cptr, _p(comp)))


def Sort_By_Key(keys, values, comp=None):
    check_i(native.n_sort_by_key(keys.m_cptr, values.m_cptr, _p(comp)))


def Is_Sorted(vec, comp=None):
    res = check_i(native.n_is_sorted(vec.m_cptr, _p(comp)))
    if res < 0:
        return None
    return res != 0


# ---- PrefixSums.py
def Inclusive_Scan(vec_in, vec_out, binary_op=None):
    check_i(native.n_inclusive_scan(vec_in.m_cptr, vec_out.m_cptr, _p(binary_op)))


def Exclusive_Scan(vec_in, vec_out, value_init=None, binary_op=None):
    check_i(native.n_exclusive_scan(vec_in.m_cptr, vec_out.m_cptr, _p(value_init), _p(binary_op)))


def Inclusive_Scan_By_Key(vec_key, vec_value, vec_out, binary_pred=None, binary_op=None):
    check_i(native.n_inclusive_scan_by_key(vec_key.m_cptr, vec_value.m_cptr, vec_out.m_cptr, _p(binary_pred),
                                           _p(binary_op)))


def Exclusive_Scan_By_Key(vec_key, vec_value, vec_out, value_init=None, binary_pred=None, binary_op=None):
    check_i(native.n_exclusive_scan_by_key(vec_key.m_cptr, vec_value.m_cptr, vec_out.m_cptr, _p(value_init),
                                           _p(binary_pred), _p(binary_op)))


# ---- Reductions.py
def Reduce(vec, value_init=None, binary_op=None):
    dvptr = check_cptr(native.n_reduce(vec.m_cptr, _p(value_init), _p(binary_op)))
    ret = ValueFromDVPtr(dvptr)
    native.n_dv_destroy(dvptr)
    return ret


# ---- Reordering.py
def Copy_If(vec_in, vec_out, pred):
    return check_u(native.n_copy_if(vec_in.m_cptr, vec_out.m_cptr, pred.m_cptr))


def Copy_If_Stencil(vec_in, vec_stencil, vec_out, pred):
    return check_u(native.n_copy_if_stencil(vec_in.m_cptr, vec_stencil.m_cptr, vec_out.m_cptr, pred.m_cptr))


# ---- Copying.py / Transformations.py
def Copy(vec_in, vec_out):
    check_i(native.n_copy(vec_in.m_cptr, vec_out.m_cptr))


def Fill(vec, value):
    check_i(native.n_fill(vec.m_cptr, value.m_cptr))


# ---- extensions: batched segmented primitives
def _offsets(offsets):
    arr = np.ascontiguousarray(np.asarray(offsets, dtype=np.uint64))
    return arr, arr.ctypes.data_as(ctypes.POINTER(ctypes.c_ulonglong)), len(arr) - 1


def Segmented_Sort_By_Key(keys, values, offsets, comp=None):
    """Sort each segment [offsets[s], offsets[s+1]) independently (stable)."""
    arr, ptr, nseg = _offsets(offsets)
    check_i(native.n_segmented_sort_by_key(keys.m_cptr, values.m_cptr, ptr, nseg, _p(comp)))


def Segmented_Inclusive_Scan(vec_in, vec_out, offsets):
    """Inclusive Plus scan of each segment [offsets[s], offsets[s+1]) independently."""
    arr, ptr, nseg = _offsets(offsets)
    check_i(native.n_segmented_inclusive_scan(vec_in.m_cptr, vec_out.m_cptr, ptr, nseg))
