"""Hot-path algorithms with the reference's Python signatures
(python/ThrustRTC/Sorting.py, PrefixSums.py, Reductions.py, Reordering.py,
Copying.py, Transformations.py)."""
import ctypes
import numpy as np
from .native import native, check_i, check_u, check_ull, check_cptr
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


def Count(vec, value):
    return check_ull(native.n_count(vec.m_cptr, value.m_cptr))


def Count_If(vec, pred):
    return check_ull(native.n_count_if(vec.m_cptr, pred.m_cptr))


def Transform_Reduce(vec, unary_op, value_init, binary_op):
    dvptr = check_cptr(native.n_transform_reduce(vec.m_cptr, unary_op.m_cptr, value_init.m_cptr, binary_op.m_cptr))
    ret = ValueFromDVPtr(dvptr)
    native.n_dv_destroy(dvptr)
    return ret


def Reduce_By_Key(key_in, value_in, key_out, value_out, binary_pred=None, binary_op=None):
    return check_u(native.n_reduce_by_key(key_in.m_cptr, value_in.m_cptr, key_out.m_cptr, value_out.m_cptr,
                                          _p(binary_pred), _p(binary_op)))


# ---- PrefixSums.py (transform scans)
def Transform_Inclusive_Scan(vec_in, vec_out, unary_op, binary_op):
    check_i(native.n_transform_inclusive_scan(vec_in.m_cptr, vec_out.m_cptr, unary_op.m_cptr, binary_op.m_cptr))


def Transform_Exclusive_Scan(vec_in, vec_out, unary_op, value_init, binary_op):
    check_i(native.n_transform_exclusive_scan(vec_in.m_cptr, vec_out.m_cptr, unary_op.m_cptr, value_init.m_cptr,
                                              binary_op.m_cptr))


# ---- Reordering.py
def Remove(vec, value):
    return check_u(native.n_remove(vec.m_cptr, value.m_cptr))


def Remove_Copy(vec_in, vec_out, value):
    return check_u(native.n_remove_copy(vec_in.m_cptr, vec_out.m_cptr, value.m_cptr))


def Remove_If(vec, pred):
    return check_u(native.n_remove_if(vec.m_cptr, pred.m_cptr))


def Remove_Copy_If(vec_in, vec_out, pred):
    return check_u(native.n_remove_copy_if(vec_in.m_cptr, vec_out.m_cptr, pred.m_cptr))


def Remove_If_Stencil(vec, stencil, pred):
    return check_u(native.n_remove_if_stencil(vec.m_cptr, stencil.m_cptr, pred.m_cptr))


def Remove_Copy_If_Stencil(vec_in, stencil, vec_out, pred):
    return check_u(native.n_remove_copy_if_stencil(vec_in.m_cptr, stencil.m_cptr, vec_out.m_cptr, pred.m_cptr))


def Unique(vec, binary_pred=None):
    return check_u(native.n_unique(vec.m_cptr, _p(binary_pred)))


def Unique_Copy(vec_in, vec_out, binary_pred=None):
    return check_u(native.n_unique_copy(vec_in.m_cptr, vec_out.m_cptr, _p(binary_pred)))


def Unique_By_Key(keys, values, binary_pred=None):
    return check_u(native.n_unique_by_key(keys.m_cptr, values.m_cptr, _p(binary_pred)))


def Unique_By_Key_Copy(keys_in, values_in, keys_out, values_out, binary_pred=None):
    return check_u(native.n_unique_by_key_copy(keys_in.m_cptr, values_in.m_cptr, keys_out.m_cptr, values_out.m_cptr,
                                               _p(binary_pred)))


def Partition(vec, pred):
    return check_u(native.n_partition(vec.m_cptr, pred.m_cptr))


def Partition_Stencil(vec, stencil, pred):
    return check_u(native.n_partition_stencil(vec.m_cptr, stencil.m_cptr, pred.m_cptr))


def Partition_Copy(vec_in, vec_true, vec_false, pred):
    return check_u(native.n_partition_copy(vec_in.m_cptr, vec_true.m_cptr, vec_false.m_cptr, pred.m_cptr))


def Partition_Copy_Stencil(vec_in, stencil, vec_true, vec_false, pred):
    return check_u(native.n_partition_copy_stencil(vec_in.m_cptr, stencil.m_cptr, vec_true.m_cptr, vec_false.m_cptr,
                                                   pred.m_cptr))


def Copy_If(vec_in, vec_out, pred):
    return check_u(native.n_copy_if(vec_in.m_cptr, vec_out.m_cptr, pred.m_cptr))


def Copy_If_Stencil(vec_in, vec_stencil, vec_out, pred):
    return check_u(native.n_copy_if_stencil(vec_in.m_cptr, vec_stencil.m_cptr, vec_out.m_cptr, pred.m_cptr))


# ---- Copying.py / Transformations.py
def Copy(vec_in, vec_out):
    check_i(native.n_copy(vec_in.m_cptr, vec_out.m_cptr))


def Fill(vec, value):
    check_i(native.n_fill(vec.m_cptr, value.m_cptr))



# ---- Transformations.py / Copying.py (the other element-wise algorithms)
def Replace(vec, old_value, new_value):
    check_i(native.n_replace(vec.m_cptr, old_value.m_cptr, new_value.m_cptr))


def Replace_If(vec, pred, new_value):
    check_i(native.n_replace_if(vec.m_cptr, pred.m_cptr, new_value.m_cptr))


def Replace_Copy(vec_in, vec_out, old_value, new_value):
    check_i(native.n_replace_copy(vec_in.m_cptr, vec_out.m_cptr, old_value.m_cptr, new_value.m_cptr))


def Replace_Copy_If(vec_in, vec_out, pred, new_value):
    check_i(native.n_replace_copy_if(vec_in.m_cptr, vec_out.m_cptr, pred.m_cptr, new_value.m_cptr))


def For_Each(vec, f):
    check_i(native.n_for_each(vec.m_cptr, f.m_cptr))


def Adjacent_Difference(vec_in, vec_out, binary_op=None):
    check_i(native.n_adjacent_difference(vec_in.m_cptr, vec_out.m_cptr, _p(binary_op)))


def Sequence(vec, value_init=None, value_step=None):
    check_i(native.n_sequence(vec.m_cptr, _p(value_init), _p(value_step)))


def Tabulate(vec, op):
    check_i(native.n_tabulate(vec.m_cptr, op.m_cptr))


def Transform(vec_in, vec_out, op):
    check_i(native.n_transform(vec_in.m_cptr, vec_out.m_cptr, op.m_cptr))


def Transform_Binary(vec_in1, vec_in2, vec_out, op):
    check_i(native.n_transform_binary(vec_in1.m_cptr, vec_in2.m_cptr, vec_out.m_cptr, op.m_cptr))


def Transform_If(vec_in, vec_out, op, pred):
    check_i(native.n_transform_if(vec_in.m_cptr, vec_out.m_cptr, op.m_cptr, pred.m_cptr))


def Transform_If_Stencil(vec_in, vec_stencil, vec_out, op, pred):
    check_i(native.n_transform_if_stencil(vec_in.m_cptr, vec_stencil.m_cptr, vec_out.m_cptr, op.m_cptr, pred.m_cptr))


def Transform_Binary_If_Stencil(vec_in1, vec_in2, vec_stencil, vec_out, op, pred):
    check_i(native.n_transform_binary_if_stencil(vec_in1.m_cptr, vec_in2.m_cptr, vec_stencil.m_cptr, vec_out.m_cptr,
                                                 op.m_cptr, pred.m_cptr))


def Gather(vec_map, vec_in, vec_out):
    check_i(native.n_gather(vec_map.m_cptr, vec_in.m_cptr, vec_out.m_cptr))


def Gather_If(vec_map, vec_stencil, vec_in, vec_out, pred=None):
    check_i(native.n_gather_if(vec_map.m_cptr, vec_stencil.m_cptr, vec_in.m_cptr, vec_out.m_cptr, _p(pred)))


def Scatter(vec_in, vec_map, vec_out):
    check_i(native.n_scatter(vec_in.m_cptr, vec_map.m_cptr, vec_out.m_cptr))


def Scatter_If(vec_in, vec_map, vec_stencil, vec_out, pred=None):
    check_i(native.n_scatter_if(vec_in.m_cptr, vec_map.m_cptr, vec_stencil.m_cptr, vec_out.m_cptr, _p(pred)))


def Swap(vec1, vec2):
    check_i(native.n_swap(vec1.m_cptr, vec2.m_cptr))


# ---- Reductions.py (the other reductions)
def _bool(res):
    return None if res < 0 else res != 0


def Equal(vec1, vec2, binary_pred=None):
    return _bool(check_i(native.n_equal(vec1.m_cptr, vec2.m_cptr, _p(binary_pred))))


def Min_Element(vec, comp=None):
    return check_ull(native.n_min_element(vec.m_cptr, _p(comp)))


def Max_Element(vec, comp=None):
    return check_ull(native.n_max_element(vec.m_cptr, _p(comp)))


def MinMax_Element(vec, comp=None):
    arr = (ctypes.c_ulonglong * 2)()
    check_i(native.n_minmax_element(vec.m_cptr, _p(comp), arr))
    return (arr[0], arr[1])


def Inner_Product(vec1, vec2, value_init, binary_op1=None, binary_op2=None):
    dvptr = check_cptr(native.n_inner_product(vec1.m_cptr, vec2.m_cptr, value_init.m_cptr, _p(binary_op1),
                                              _p(binary_op2)))
    ret = ValueFromDVPtr(dvptr)
    native.n_dv_destroy(dvptr)
    return ret


def All_Of(vec, pred):
    return _bool(check_i(native.n_all_of(vec.m_cptr, pred.m_cptr)))


def Any_Of(vec, pred):
    return _bool(check_i(native.n_any_of(vec.m_cptr, pred.m_cptr)))


def None_Of(vec, pred):
    return _bool(check_i(native.n_none_of(vec.m_cptr, pred.m_cptr)))


def Is_Partitioned(vec, pred):
    return _bool(check_i(native.n_is_partitioned(vec.m_cptr, pred.m_cptr)))


# ---- Searching.py
def _index(res, vec):
    return None if res >= vec.size() else res


def Find(vec, value):
    return _index(check_i(native.n_find(vec.m_cptr, value.m_cptr)), vec)


def Find_If(vec, pred):
    return _index(check_i(native.n_find_if(vec.m_cptr, pred.m_cptr)), vec)


def Find_If_Not(vec, pred):
    return _index(check_i(native.n_find_if_not(vec.m_cptr, pred.m_cptr)), vec)


def Mismatch(vec1, vec2, pred=None):
    return _index(check_i(native.n_mismatch(vec1.m_cptr, vec2.m_cptr, _p(pred))), vec1)


def Lower_Bound(vec, value, comp=None):
    return check_ull(native.n_lower_bound(vec.m_cptr, value.m_cptr, _p(comp)))


def Upper_Bound(vec, value, comp=None):
    return check_ull(native.n_upper_bound(vec.m_cptr, value.m_cptr, _p(comp)))


def Binary_Search(vec, value, comp=None):
    return _bool(check_i(native.n_binary_search(vec.m_cptr, value.m_cptr, _p(comp))))


def Partition_Point(vec, pred):
    return check_ull(native.n_partition_point(vec.m_cptr, pred.m_cptr))


def Is_Sorted_Until(vec, comp=None):
    return check_ull(native.n_is_sorted_until(vec.m_cptr, _p(comp)))


def Lower_Bound_V(vec, values, result, comp=None):
    check_i(native.n_lower_bound_v(vec.m_cptr, values.m_cptr, result.m_cptr, _p(comp)))


def Upper_Bound_V(vec, values, result, comp=None):
    check_i(native.n_upper_bound_v(vec.m_cptr, values.m_cptr, result.m_cptr, _p(comp)))


def Binary_Search_V(vec, values, result, comp=None):
    check_i(native.n_binary_search_v(vec.m_cptr, values.m_cptr, result.m_cptr, _p(comp)))


# ---- Merging.py
def Merge(vec1, vec2, vec_out, comp=None):
    check_i(native.n_merge(vec1.m_cptr, vec2.m_cptr, vec_out.m_cptr, _p(comp)))


def Merge_By_Key(keys1, keys2, value1, value2, keys_out, value_out, comp=None):
    check_i(native.n_merge_by_key(keys1.m_cptr, keys2.m_cptr, value1.m_cptr, value2.m_cptr, keys_out.m_cptr,
                                  value_out.m_cptr, _p(comp)))

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


def Segmented_Sort(keys, offsets, comp=None):
    """Keys-only stable sort of each segment independently."""
    arr, ptr, nseg = _offsets(offsets)
    check_i(native.n_segmented_sort(keys.m_cptr, ptr, nseg, _p(comp)))


def Segmented_Exclusive_Scan(vec_in, vec_out, offsets):
    """Exclusive Plus scan (init 0) of each segment independently."""
    arr, ptr, nseg = _offsets(offsets)
    check_i(native.n_segmented_exclusive_scan(vec_in.m_cptr, vec_out.m_cptr, ptr, nseg))


def Segmented_Reduce(vec_in, offsets, results, binary_op=None):
    """results[s] = reduction of segment s (no init; every segment non-empty)."""
    arr, ptr, nseg = _offsets(offsets)
    check_i(native.n_segmented_reduce(vec_in.m_cptr, ptr, nseg, results.m_cptr, _p(binary_op)))


def Set_Batch_Autotune(on):
    native.n_set_batch_autotune(1 if on else 0)


def Set_Batch_Schedule(schedule=None):
    """Fixed tile schedule for every batched sort, e.g. "<=2048:256x8 rest:512x16";
    None returns to automatic (tuned) schedules."""
    check_i(native.n_set_batch_schedule(None if schedule is None else schedule.encode()))


def _json_out(fn, *args):
    import json
    n = fn(*args, None, 0)
    buf = ctypes.create_string_buffer(int(n) + 1)
    fn(*args, buf, len(buf))
    return json.loads(buf.value.decode())


def Batch_Tuning_Report():
    """The batched-sort autotuner's runs: signature, candidate times, choice."""
    return _json_out(native.n_batch_tuning_report)


class Batch:
    """Segmented operations recorded once, captured into one HIP graph by
    instantiate() and replayed by launch() (one hipGraphLaunch).  The vectors
    must stay alive while the batch exists (the batch keeps references)."""

    def __init__(self):
        self.m_cptr = check_cptr(native.n_batch_create())
        self._keep = []

    def __del__(self):
        if getattr(self, "m_cptr", None):
            native.n_batch_destroy(self.m_cptr)
            self.m_cptr = None

    def add_sort(self, keys, offsets, comp=None):
        arr, ptr, nseg = _offsets(offsets)
        self._keep += [keys, comp]
        check_i(native.n_batch_add_sort(self.m_cptr, keys.m_cptr, ptr, nseg, _p(comp)))

    def add_sort_by_key(self, keys, values, offsets, comp=None):
        arr, ptr, nseg = _offsets(offsets)
        self._keep += [keys, values, comp]
        check_i(native.n_batch_add_sort_by_key(self.m_cptr, keys.m_cptr, values.m_cptr, ptr, nseg, _p(comp)))

    def add_inclusive_scan(self, vec_in, vec_out, offsets, binary_op=None):
        arr, ptr, nseg = _offsets(offsets)
        self._keep += [vec_in, vec_out, binary_op]
        check_i(native.n_batch_add_inclusive_scan(self.m_cptr, vec_in.m_cptr, vec_out.m_cptr, ptr, nseg,
                                                  _p(binary_op)))

    def add_reduce(self, vec_in, offsets, results, binary_op=None):
        arr, ptr, nseg = _offsets(offsets)
        self._keep += [vec_in, results, binary_op]
        check_i(native.n_batch_add_reduce(self.m_cptr, vec_in.m_cptr, ptr, nseg, results.m_cptr, _p(binary_op)))

    def instantiate(self):
        check_i(native.n_batch_instantiate(self.m_cptr))

    def launch(self):
        check_i(native.n_batch_launch(self.m_cptr))

    def describe(self):
        return _json_out(native.n_batch_describe, self.m_cptr)


# ---- extension: request server (include/server.h)
REQ_SORT_BY_KEY, REQ_SORT, REQ_INCLUSIVE_SCAN, REQ_REDUCE = 0, 1, 2, 3


class Server:
    """Request server over device-resident inputs: HBM-budget admission and
    continuous batching on `slots` streams (batching=False: one request at a
    time through the single-call path).  Outputs belong to the server until
    release()."""

    def __init__(self, hbm_budget=64 << 30, slots=3, batching=True):
        self.m_cptr = check_cptr(native.n_server_create(hbm_budget, slots, 1 if batching else 0))

    def __del__(self):
        if getattr(self, "m_cptr", None):
            native.n_server_destroy(self.m_cptr)
            self.m_cptr = None

    def set_batch_limits(self, max_requests=4096, max_elements=1 << 28):
        """A batch closes at max_requests requests or before its elements pass
        max_elements (it always takes at least one request)."""
        native.n_server_set_batch_limits(self.m_cptr, max_requests, max_elements)

    def submit(self, op, key_elem, in_a, n, val_elem=None, in_b=None):
        """in_a / in_b: device addresses (ints) or DVVector-likes."""
        def addr(x):
            if x is None or isinstance(x, int):
                return x
            return native.n_dvvectorlike_data(x.m_cptr)
        t = native.n_server_submit(self.m_cptr, op, key_elem.encode(), None if val_elem is None else val_elem.encode(),
                                   addr(in_a), addr(in_b), n)
        if t < 0:
            raise ValueError("invalid request")
        return t

    def wait(self, ticket):
        check_i(native.n_server_wait(self.m_cptr, ticket))

    def done(self, ticket):
        return native.n_server_done(self.m_cptr, ticket) == 1

    def output(self, ticket, which=0):
        return native.n_server_output(self.m_cptr, ticket, which)

    def latency_us(self, ticket):
        return native.n_server_latency_us(self.m_cptr, ticket)

    def release(self, ticket):
        native.n_server_release(self.m_cptr, ticket)

    def stats(self):
        return _json_out(native.n_server_stats, self.m_cptr)

    def replay(self, arrival_us, ops, key_elems, val_elems, in_a, in_b, ns):
        """Submit request i at arrival_us[i] (microseconds from the start);
        returns (makespan seconds, latency_us array)."""
        cnt = len(arrival_us)
        arr = np.ascontiguousarray(arrival_us, dtype=np.float64)
        opa = np.ascontiguousarray(ops, dtype=np.int32)
        na = np.ascontiguousarray(ns, dtype=np.uint64)
        lat = np.zeros(cnt, dtype=np.float64)
        ke = (ctypes.c_char_p * cnt)(*[k.encode() for k in key_elems])
        ve = (ctypes.c_char_p * cnt)(*[None if v is None else v.encode() for v in val_elems])
        pa = (ctypes.c_void_p * cnt)(*in_a)
        pb = (ctypes.c_void_p * cnt)(*[None if b is None else b for b in in_b])
        dp = ctypes.POINTER(ctypes.c_double)
        span = native.n_server_replay(self.m_cptr, cnt, arr.ctypes.data_as(dp), opa.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                      ke, ve, pa, pb, na.ctypes.data_as(ctypes.POINTER(ctypes.c_ulonglong)),
                                      lat.ctypes.data_as(dp))
        return span, lat
