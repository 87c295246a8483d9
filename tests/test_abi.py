"""CPU checks of the drop-in boundary: the C-ABI library loads, exports every
symbol include/trtc_capi.h declares, keeps the reference's error conventions,
and its host logic (functor code generation, struct layout, hipRTC
instantiation of the kernel templates for gfx950) works without a device.
No compute call is made here."""
import ctypes
import json
import os
import re
import struct

import pytest

from helpers import native_module, gpu_available, ROOT

HEADER = os.path.join(ROOT, "include", "trtc_capi.h")


@pytest.fixture(scope="module")
def nm():
    return native_module()


def header_symbols():
    text = open(HEADER).read()
    return re.findall(r"TRTC_CAPI\s+[^;]*?\b(n_\w+)\s*\(", text)


def test_every_declared_symbol_is_exported(nm):
    syms = header_symbols()
    assert len(syms) > 80
    lib = ctypes.CDLL(nm.LIB_PATH)
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing
    # the Python binding declares a signature for each of them
    assert set(syms) == set(nm.SIGNATURES), set(syms) ^ set(nm.SIGNATURES)


def test_every_reference_function_declared_and_exported(nm):
    # every function of the reference's python/api.h, same name
    ref = json.load(open(os.path.join(ROOT, "tests", "golden", "api_h_functions.json")))["functions"]
    assert len(ref) == 149
    syms = set(header_symbols())
    assert not [f for f in ref if f not in syms]
    assert not [f for f in ref if not hasattr(nm.native, f)]


def test_reference_hot_path_subset_present():
    # the reference's python/api.h names of the primitive core
    need = ["n_sort", "n_sort_by_key", "n_inclusive_scan", "n_exclusive_scan", "n_inclusive_scan_by_key",
            "n_exclusive_scan_by_key", "n_reduce", "n_copy_if", "n_copy_if_stencil", "n_trtc_try_init",
            "n_set_verbose", "n_set_kernel_debug", "n_wait", "n_get_ptx_arch", "n_dvvector_create",
            "n_dvvector_to_host", "n_dvvectorlike_size", "n_dvvectorlike_name_elem_cls",
            "n_dvvectoradaptor_create", "n_dvrange_create", "n_dvvector_from_dvs", "n_functor_create",
            "n_built_in_functor_create", "n_string_array_create", "n_pointer_array_create", "n_dim3_create",
            "n_dv_destroy", "n_dv_name_view_cls", "n_kernel_create", "n_for_create"]
    syms = set(header_symbols())
    assert not [n for n in need if n not in syms]


def test_try_init_without_gpu(nm):
    if gpu_available():
        pytest.skip("a GPU is visible")
    assert nm.native.n_trtc_try_init() == -100
    with pytest.raises(ImportError):
        nm.check_i(-100)
    with pytest.raises(SystemError):
        nm.check_i(-1)
    with pytest.raises(ValueError):
        nm.check_i(-2)


def _strs(nm, items):
    arr = (ctypes.c_char_p * max(1, len(items)))(*[s.encode() for s in items])
    return nm.native.n_string_array_create(len(items), arr)


def _ptrs(nm, items):
    arr = (ctypes.c_void_p * max(1, len(items)))(*items)
    return nm.native.n_pointer_array_create(len(items), arr)


def test_scalars_roundtrip(nm):
    n = nm.native
    for create, value, v in [(n.n_dvint8_create, n.n_dvint8_value, -5), (n.n_dvuint16_create, n.n_dvuint16_value, 65535),
                             (n.n_dvint64_create, n.n_dvint64_value, -2**40), (n.n_dvdouble_create, n.n_dvdouble_value, 1.25)]:
        p = create(v)
        assert value(p) == v
        n.n_dv_destroy(p)
    p = n.n_dvint32_create(7)
    assert n.n_dv_name_view_cls(p) == b"int32_t"
    n.n_dv_destroy(p)


def test_functor_codegen_and_jit_compile(nm):
    n = nm.native
    f = n.n_functor_create(_ptrs(nm, []), _strs(nm, []), _strs(nm, ["x"]), b"        return x % 3 == 0;\n")
    name = n.n_dv_name_view_cls(f).decode()
    assert re.fullmatch(r"_S_[0-9a-f]{16}", name)
    expr = f"trtc_copy_if_k<VectorView<int32_t>, VectorView<int32_t>, VectorView<int32_t>, {name}>"
    assert n.n_jit_compile_check(b"gfx950", expr.encode()) == 0
    # scan with a user op, int8 in / int32 out (the JIT path's type mixing)
    op = n.n_functor_create(_ptrs(nm, []), _strs(nm, []), _strs(nm, ["a", "b"]), b"        return a + b;\n")
    oname = n.n_dv_name_view_cls(op).decode()
    expr = f"trtc_scan_k<VectorView<int8_t>, VectorView<int32_t>, {oname}, int32_t>"
    assert n.n_jit_compile_check(b"gfx950", expr.encode()) == 0


def test_bad_functor_fails_to_compile_cleanly(nm):
    n = nm.native
    f = n.n_functor_create(_ptrs(nm, []), _strs(nm, []), _strs(nm, ["x"]), b"        return x %% ;\n")
    name = n.n_dv_name_view_cls(f).decode()
    expr = f"trtc_copy_if_k<VectorView<int32_t>, VectorView<int32_t>, VectorView<int32_t>, {name}>"
    assert n.n_jit_compile_check(b"gfx950", expr.encode()) == -1


def test_functor_capture_layout(nm):
    # members laid out with C rules: int8 at 0, VectorView (align 8) at 8,
    # double at 24, size 32
    n = nm.native
    i8 = n.n_dvint8_create(-3)
    d = n.n_dvdouble_create(2.5)
    # a VectorView without a device: an adaptor over a fake pointer
    vv = n.n_dvvectoradaptor_create(b"int32_t", 10, ctypes.c_void_p(0x1000))
    f = n.n_functor_create(_ptrs(nm, [i8, vv, d]), _strs(nm, ["a", "v", "d"]), _strs(nm, ["x"]),
                           b"        return x + a + v[0] + d;\n")
    # read the view through a scan name expression compile (layout must match
    # the device struct: a static_assert-free check is the compile itself)
    name = n.n_dv_name_view_cls(f).decode()
    expr = f"trtc_reduce_k<VectorView<double>, {name}, double>"
    # a reduce op must be binary; use the functor as a copy_if predicate instead
    expr = f"trtc_copy_if_k<VectorView<double>, VectorView<double>, VectorView<double>, {name}>"
    assert n.n_jit_compile_check(b"gfx950", expr.encode()) == 0


def test_embedded_headers_match_sources():
    inc = open(os.path.join(ROOT, "thrustrtc_amd", "csrc", "host", "embedded_headers.inc")).read()
    for h in ["trtc_builtin.hpp", "trtc_prims.hpp", "trtc_radix.hpp", "trtc_kernels.hpp"]:
        src = open(os.path.join(ROOT, "thrustrtc_amd", "csrc", "device", h)).read()
        assert src in inc, h


def test_merge_sort_kernels_compile_for_user_comparator(nm):
    # the comparison-sort kernels instantiate on a user comparator functor
    n = nm.native
    f = n.n_functor_create(_ptrs(nm, []), _strs(nm, []), _strs(nm, ["x", "y"]), b"        return (x & 255) < (y & 255);\n")
    name = n.n_dv_name_view_cls(f).decode()
    for k in ("trtc_msort_tile_k", "trtc_msort_merge_k", "trtc_msort_split_k"):
        assert n.n_jit_compile_check(b"gfx950", f"{k}<uint32_t, {name}>".encode()) == 0
    assert n.n_jit_compile_check(b"gfx950", b"trtc_msort_merge_k<double, Greater>") == 0


def _entry(lowered, code):
    """A disk-cache entry as the library writes it (the format is part of the
    cache's contract: header with code length and a 64-bit FNV-1a checksum
    whose offset basis is the library's 1469598103934665603)."""
    h = 1469598103934665603
    for b in code:
        h = ((h ^ b) * 0x100000001b3) & 0xFFFFFFFFFFFFFFFF
    return b"TRTCCO2 %d %016x %s\n" % (len(code), h, lowered) + code


def test_cache_entry_validator(nm):
    """The host check every disk-cache entry passes before the loader sees it:
    a good entry passes; truncated, extended, bit-flipped, empty, name-only,
    old-format and foreign files are rejected."""
    n = nm.native
    code = bytes(range(256)) * 7 + b"\x7fELF tail"
    good = _entry(b"_Z3fooPi", code)
    ok = lambda blob: n.n_jit_cache_entry_valid(blob, len(blob))
    assert ok(good) == 1
    assert ok(good[:-1]) == 0                         # truncated code
    assert ok(good + b"x") == 0                       # trailing bytes
    flipped = bytearray(good)
    flipped[-100] ^= 0x10
    assert ok(bytes(flipped)) == 0                    # damaged code
    assert ok(b"") == 0
    assert ok(b"TRTCCO2 0 0000000000000000 name\n") == 0   # name only
    assert ok(b"_Z3fooPi\n" + code) == 0               # earlier format (name line + code)
    assert ok(b"\x7fELF" + code) == 0                  # foreign file
    assert ok(_entry(b"", code)) == 0                  # no kernel name
    assert ok(good.replace(b"TRTCCO2 %d" % len(code), b"TRTCCO2 %d" % (len(code) + 1))) == 0
