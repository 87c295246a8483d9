"""ctypes binding of libPyThrustRTC.so (the C ABI declared in include/trtc_capi.h).

Mirrors the reference's Native.py (python/ThrustRTC/Native.py:1-36): the
library handle `native` and the return-code checkers.  This module has no
package-relative imports so tests can load it without initialising a GPU.
"""
import ctypes
import os

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libPyThrustRTC.so")

_vp, _i, _u, _ull, _ll = ctypes.c_void_p, ctypes.c_int, ctypes.c_uint, ctypes.c_ulonglong, ctypes.c_longlong
_f, _d, _s = ctypes.c_float, ctypes.c_double, ctypes.c_char_p

# name: (restype, argtypes)
SIGNATURES = {
    "n_string_array_create": (_vp, [_ull, ctypes.POINTER(_s)]),
    "n_string_array_destroy": (None, [_vp]),
    "n_pointer_array_create": (_vp, [_ull, ctypes.POINTER(_vp)]),
    "n_pointer_array_destroy": (None, [_vp]),
    "n_dim3_create": (_vp, [_u, _u, _u]),
    "n_dim3_destroy": (None, [_vp]),
    "n_set_libnvrtc_path": (None, [_s]),
    "n_trtc_try_init": (_i, []),
    "n_get_ptx_arch": (_i, []),
    "n_set_verbose": (None, [_u]),
    "n_set_kernel_debug": (None, [_u]),
    "n_add_include_dir": (None, [_s]),
    "n_add_built_in_header": (None, [_s, _s]),
    "n_add_inlcude_filename": (None, [_s]),
    "n_add_code_block": (None, [_s]),
    "n_add_constant_object": (None, [_s, _vp]),
    "n_wait": (None, []),
    "n_kernel_create": (_vp, [_vp, _s]),
    "n_kernel_destroy": (None, [_vp]),
    "n_kernel_num_params": (_i, [_vp]),
    "n_kernel_calc_optimal_block_size": (_i, [_vp, _vp, _u]),
    "n_kernel_calc_number_blocks": (_i, [_vp, _vp, _i, _u]),
    "n_kernel_launch": (_i, [_vp, _vp, _vp, _vp, _i]),
    "n_for_create": (_vp, [_vp, _s, _s]),
    "n_for_destroy": (None, [_vp]),
    "n_for_num_params": (_i, [_vp]),
    "n_for_launch": (_i, [_vp, _i, _i, _vp]),
    "n_for_launch_n": (_i, [_vp, _i, _vp]),
    "n_dv_name_view_cls": (_s, [_vp]),
    "n_dv_destroy": (None, [_vp]),
    "n_dvint8_create": (_vp, [_i]), "n_dvint8_value": (_i, [_vp]),
    "n_dvuint8_create": (_vp, [_u]), "n_dvuint8_value": (_u, [_vp]),
    "n_dvint16_create": (_vp, [_i]), "n_dvint16_value": (_i, [_vp]),
    "n_dvuint16_create": (_vp, [_u]), "n_dvuint16_value": (_u, [_vp]),
    "n_dvint32_create": (_vp, [_i]), "n_dvint32_value": (_i, [_vp]),
    "n_dvuint32_create": (_vp, [_u]), "n_dvuint32_value": (_u, [_vp]),
    "n_dvint64_create": (_vp, [_ll]), "n_dvint64_value": (_ll, [_vp]),
    "n_dvuint64_create": (_vp, [_ull]), "n_dvuint64_value": (_ull, [_vp]),
    "n_dvfloat_create": (_vp, [_f]), "n_dvfloat_value": (_f, [_vp]),
    "n_dvdouble_create": (_vp, [_d]), "n_dvdouble_value": (_d, [_vp]),
    "n_dvbool_create": (_vp, [_i]), "n_dvbool_value": (_i, [_vp]),
    "n_dvcomplex64_create": (_vp, [_f, _f]), "n_dvcomplex64_real": (_f, [_vp]), "n_dvcomplex64_imag": (_f, [_vp]),
    "n_dvcomplex128_create": (_vp, [_d, _d]), "n_dvcomplex128_real": (_d, [_vp]), "n_dvcomplex128_imag": (_d, [_vp]),
    "n_dvvectorlike_name_elem_cls": (_s, [_vp]),
    "n_dvvectorlike_size": (_ull, [_vp]),
    "n_dvvector_create": (_vp, [_s, _ull, _vp]),
    "n_dvvector_to_host": (None, [_vp, _vp, _ull, _ull]),
    "n_dvvector_from_dvs": (_vp, [_vp]),
    "n_dvvectoradaptor_create": (_vp, [_s, _ull, _vp]),
    "n_dvrange_create": (_vp, [_vp, _ull, _ull]),
    "n_functor_create": (_vp, [_vp, _vp, _vp, _s]),
    "n_built_in_functor_create": (_vp, [_s]),
    "n_fill": (_i, [_vp, _vp]),
    "n_copy": (_i, [_vp, _vp]),
    "n_reduce": (_vp, [_vp, _vp, _vp]),
    "n_is_sorted": (_i, [_vp, _vp]),
    "n_inclusive_scan": (_i, [_vp, _vp, _vp]),
    "n_exclusive_scan": (_i, [_vp, _vp, _vp, _vp]),
    "n_inclusive_scan_by_key": (_i, [_vp, _vp, _vp, _vp, _vp]),
    "n_exclusive_scan_by_key": (_i, [_vp, _vp, _vp, _vp, _vp, _vp]),
    "n_count": (_ull, [_vp, _vp]),
    "n_count_if": (_ull, [_vp, _vp]),
    "n_reduce_by_key": (_u, [_vp, _vp, _vp, _vp, _vp, _vp]),
    "n_transform_reduce": (_vp, [_vp, _vp, _vp, _vp]),
    "n_transform_inclusive_scan": (_i, [_vp, _vp, _vp, _vp]),
    "n_transform_exclusive_scan": (_i, [_vp, _vp, _vp, _vp, _vp]),
    "n_remove": (_u, [_vp, _vp]),
    "n_remove_copy": (_u, [_vp, _vp, _vp]),
    "n_remove_if": (_u, [_vp, _vp]),
    "n_remove_copy_if": (_u, [_vp, _vp, _vp]),
    "n_remove_if_stencil": (_u, [_vp, _vp, _vp]),
    "n_remove_copy_if_stencil": (_u, [_vp, _vp, _vp, _vp]),
    "n_unique": (_u, [_vp, _vp]),
    "n_unique_copy": (_u, [_vp, _vp, _vp]),
    "n_unique_by_key": (_u, [_vp, _vp, _vp]),
    "n_unique_by_key_copy": (_u, [_vp, _vp, _vp, _vp, _vp]),
    "n_partition": (_u, [_vp, _vp]),
    "n_partition_stencil": (_u, [_vp, _vp, _vp]),
    "n_partition_copy": (_u, [_vp, _vp, _vp, _vp]),
    "n_partition_copy_stencil": (_u, [_vp, _vp, _vp, _vp, _vp]),
    "n_dvconstant_create": (_vp, [_vp, _ull]),
    "n_dvcounter_create": (_vp, [_vp, _ull]),
    "n_dvdiscard_create": (_vp, [_s, _ull]),
    "n_dvpermutation_create": (_vp, [_vp, _vp]),
    "n_dvreverse_create": (_vp, [_vp]),
    "n_dvtransform_create": (_vp, [_vp, _s, _vp]),
    "n_dvzipped_create": (_vp, [_vp, _vp]),
    "n_dvcustomvector_create": (_vp, [_vp, _vp, _s, _s, _s, _ull, _u]),
    "n_dvtuple_create": (_vp, [_vp, _vp]),
    "n_replace": (_i, [_vp, _vp, _vp]),
    "n_replace_if": (_i, [_vp, _vp, _vp]),
    "n_replace_copy": (_i, [_vp, _vp, _vp, _vp]),
    "n_replace_copy_if": (_i, [_vp, _vp, _vp, _vp]),
    "n_for_each": (_i, [_vp, _vp]),
    "n_adjacent_difference": (_i, [_vp, _vp, _vp]),
    "n_sequence": (_i, [_vp, _vp, _vp]),
    "n_tabulate": (_i, [_vp, _vp]),
    "n_transform": (_i, [_vp, _vp, _vp]),
    "n_transform_binary": (_i, [_vp, _vp, _vp, _vp]),
    "n_transform_if": (_i, [_vp, _vp, _vp, _vp]),
    "n_transform_if_stencil": (_i, [_vp, _vp, _vp, _vp, _vp]),
    "n_transform_binary_if_stencil": (_i, [_vp, _vp, _vp, _vp, _vp, _vp]),
    "n_gather": (_i, [_vp, _vp, _vp]),
    "n_gather_if": (_i, [_vp, _vp, _vp, _vp, _vp]),
    "n_scatter": (_i, [_vp, _vp, _vp]),
    "n_scatter_if": (_i, [_vp, _vp, _vp, _vp, _vp]),
    "n_swap": (_i, [_vp, _vp]),
    "n_equal": (_i, [_vp, _vp, _vp]),
    "n_min_element": (_ull, [_vp, _vp]),
    "n_max_element": (_ull, [_vp, _vp]),
    "n_minmax_element": (_i, [_vp, _vp, ctypes.POINTER(_ull)]),
    "n_inner_product": (_vp, [_vp, _vp, _vp, _vp, _vp]),
    "n_all_of": (_i, [_vp, _vp]),
    "n_any_of": (_i, [_vp, _vp]),
    "n_none_of": (_i, [_vp, _vp]),
    "n_is_partitioned": (_i, [_vp, _vp]),
    "n_find": (_ll, [_vp, _vp]),
    "n_find_if": (_ll, [_vp, _vp]),
    "n_find_if_not": (_ll, [_vp, _vp]),
    "n_mismatch": (_ll, [_vp, _vp, _vp]),
    "n_lower_bound": (_ull, [_vp, _vp, _vp]),
    "n_upper_bound": (_ull, [_vp, _vp, _vp]),
    "n_binary_search": (_i, [_vp, _vp, _vp]),
    "n_lower_bound_v": (_i, [_vp, _vp, _vp, _vp]),
    "n_upper_bound_v": (_i, [_vp, _vp, _vp, _vp]),
    "n_binary_search_v": (_i, [_vp, _vp, _vp, _vp]),
    "n_partition_point": (_ull, [_vp, _vp]),
    "n_is_sorted_until": (_ull, [_vp, _vp]),
    "n_merge": (_i, [_vp, _vp, _vp, _vp]),
    "n_merge_by_key": (_i, [_vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "n_copy_if": (_u, [_vp, _vp, _vp]),
    "n_copy_if_stencil": (_u, [_vp, _vp, _vp, _vp]),
    "n_sort": (_i, [_vp, _vp]),
    "n_sort_by_key": (_i, [_vp, _vp, _vp]),
    "n_segmented_sort_by_key": (_i, [_vp, _vp, ctypes.POINTER(_ull), _ull, _vp]),
    "n_segmented_inclusive_scan": (_i, [_vp, _vp, ctypes.POINTER(_ull), _ull]),
    "n_segmented_sort": (_i, [_vp, ctypes.POINTER(_ull), _ull, _vp]),
    "n_segmented_exclusive_scan": (_i, [_vp, _vp, ctypes.POINTER(_ull), _ull]),
    "n_segmented_reduce": (_i, [_vp, ctypes.POINTER(_ull), _ull, _vp, _vp]),
    "n_set_batch_autotune": (None, [_i]),
    "n_set_batch_schedule": (_i, [_s]),
    "n_batch_tuning_report": (_ull, [ctypes.c_char_p, _ull]),
    "n_batch_create": (_vp, []),
    "n_batch_destroy": (None, [_vp]),
    "n_batch_add_sort": (_i, [_vp, _vp, ctypes.POINTER(_ull), _ull, _vp]),
    "n_batch_add_sort_by_key": (_i, [_vp, _vp, _vp, ctypes.POINTER(_ull), _ull, _vp]),
    "n_batch_add_inclusive_scan": (_i, [_vp, _vp, _vp, ctypes.POINTER(_ull), _ull, _vp]),
    "n_batch_add_reduce": (_i, [_vp, _vp, ctypes.POINTER(_ull), _ull, _vp, _vp]),
    "n_batch_instantiate": (_i, [_vp]),
    "n_batch_launch": (_i, [_vp]),
    "n_batch_describe": (_ull, [_vp, ctypes.c_char_p, _ull]),
    "n_server_create": (_vp, [_ull, _i, _i]),
    "n_server_destroy": (None, [_vp]),
    "n_server_set_batch_limits": (None, [_vp, _ull, _ull]),
    "n_server_submit": (ctypes.c_longlong, [_vp, _i, _s, _s, _vp, _vp, _ull]),
    "n_server_wait": (_i, [_vp, ctypes.c_longlong]),
    "n_server_done": (_i, [_vp, ctypes.c_longlong]),
    "n_server_output": (_vp, [_vp, ctypes.c_longlong, _i]),
    "n_server_latency_us": (_d, [_vp, ctypes.c_longlong]),
    "n_server_release": (None, [_vp, ctypes.c_longlong]),
    "n_server_stats": (_ull, [_vp, ctypes.c_char_p, _ull]),
    "n_server_replay": (_d, [_vp, _ull, ctypes.POINTER(_d), ctypes.POINTER(_i), ctypes.POINTER(_s), ctypes.POINTER(_s),
                             ctypes.POINTER(_vp), ctypes.POINTER(_vp), ctypes.POINTER(_ull), ctypes.POINTER(_d)]),
    "n_set_stream": (None, [_vp]),
    "n_get_stream": (_vp, []),
    "n_dvvectorlike_data": (_vp, [_vp]),
    "n_jit_compile_check": (_i, [_s, _s]),
    "n_inclusive_scan_noalloc": (_i, [_vp, _vp, _vp]),
    "n_jit_cache_stats": (None, [ctypes.POINTER(_ull)]),
    "n_jit_cache_entry_valid": (_i, [ctypes.c_char_p, _ull]),
    "n_profile_begin": (_i, []),
    "n_profile_end": (_ull, [ctypes.c_char_p, _ull]),
}


def load(path=LIB_PATH):
    """Load the library and declare every C ABI signature.  Raises ImportError
    (never falls back to anything else) when the native library is missing."""
    if not os.path.exists(path):
        raise ImportError(f"{path} is not built: run `python thrustrtc_amd/build.py`")
    lib = ctypes.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


native = load()


def check_i(v):
    if v == -1:
        raise SystemError("An internal error happend")
    elif v == -2:
        raise ValueError("Wrong number of arguments.")
    elif v == -100:
        raise ImportError("cannot import ThrustRTC")
    return v


def check_u(v):
    if v == 0xFFFFFFFF:
        raise SystemError("An internal error happend")
    return v


def check_ull(v):
    if v == 0xFFFFFFFFFFFFFFFF:
        raise SystemError("An internal error happend")
    return v


def check_cptr(v):
    if not v:
        raise SystemError("An internal error happend")
    return v
