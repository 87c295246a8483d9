"""Device vectors (reference: python/ThrustRTC/DVVector.py)."""
import ctypes
import numpy as np
from .native import native, check_cptr
from .DeviceViewable import DeviceViewable
from .utils import ObjArray

_NP_TO_C = [
    (np.int8, "int8_t"), (np.uint8, "uint8_t"), (np.int16, "int16_t"), (np.uint16, "uint16_t"),
    (np.int32, "int32_t"), (np.uint32, "uint32_t"), (np.int64, "int64_t"), (np.uint64, "uint64_t"),
    (np.float32, "float"), (np.float64, "double"), (np.bool_, "bool"),
    (np.complex64, "cuFloatComplex"), (np.complex128, "cuDoubleComplex"),
]
_C_TO_NP = {c: n for n, c in _NP_TO_C}
_C_TO_NP.update({"int": np.int32, "unsigned int": np.uint32, "char": np.int8, "long long": np.int64,
                 "unsigned long long": np.uint64, "size_t": np.uint64, "short": np.int16,
                 "unsigned short": np.uint16, "unsigned char": np.uint8, "long": np.int64,
                 "unsigned long": np.uint64, "signed char": np.int8})


def elem_cls_of_dtype(dtype):
    dtype = np.dtype(dtype)
    for n, c in _NP_TO_C:
        if dtype == np.dtype(n):
            return c
    raise TypeError(f"unsupported dtype {dtype}")


def dtype_of_elem_cls(elem_cls):
    return _C_TO_NP[elem_cls]


class DVVectorLike(DeviceViewable):
    def name_elem_cls(self):
        return native.n_dvvectorlike_name_elem_cls(self.m_cptr).decode("utf-8")

    def size(self):
        return native.n_dvvectorlike_size(self.m_cptr)

    def range(self, begin=0, end=-1):
        return DVRange(self, begin, end)


class DVRange(DVVectorLike):
    def __init__(self, src, begin=0, end=-1):
        self.m_src = src
        self.m_cptr = native.n_dvrange_create(src.m_cptr, ctypes.c_ulonglong(begin).value,
                                              ctypes.c_ulonglong(end).value)


class DVVector(DVVectorLike):
    def __init__(self, cptr):
        self.m_cptr = cptr

    def to_host(self, begin=0, end=-1):
        nptype = dtype_of_elem_cls(self.name_elem_cls())
        if end == -1:
            end = self.size()
        ret = np.empty(end - begin, dtype=nptype)
        native.n_dvvector_to_host(self.m_cptr, ret.ctypes.data, begin, end)
        return ret


def device_vector(elem_cls, size, ptr_host_data=None):
    if size < 1:
        raise ValueError("0-size vector not allowed")
    return DVVector(check_cptr(native.n_dvvector_create(elem_cls.encode("utf-8"), size, ptr_host_data)))


def device_vector_from_numpy(nparr):
    nparr = np.ascontiguousarray(nparr)
    return device_vector(elem_cls_of_dtype(nparr.dtype), len(nparr), nparr.ctypes.data)


def device_vector_from_list(lst, elem_cls):
    nparr = np.array(lst, dtype=dtype_of_elem_cls(elem_cls))
    return device_vector(elem_cls, len(lst), nparr.ctypes.data)


def device_vector_from_dvs(lst_dv):
    if len(lst_dv) < 1:
        raise ValueError("0-size vector not allowed")
    dvarr = ObjArray(lst_dv)
    return DVVector(check_cptr(native.n_dvvector_from_dvs(dvarr.m_cptr)))


class DVVectorAdaptor(DVVectorLike):
    """Borrows device memory owned by someone else (the counterpart of the
    reference's DVNumbaVector / DVCupyVector)."""

    def __init__(self, elem_cls, size, dev_ptr, owner=None):
        self.owner = owner
        self.m_cptr = check_cptr(native.n_dvvectoradaptor_create(elem_cls.encode("utf-8"), size, dev_ptr))

    def data_ptr(self):
        return native.n_dvvectorlike_data(self.m_cptr)


_TORCH_TO_C = {"torch.int8": "int8_t", "torch.uint8": "uint8_t", "torch.int16": "int16_t",
               "torch.int32": "int32_t", "torch.int64": "int64_t", "torch.float32": "float",
               "torch.float64": "double", "torch.bool": "bool", "torch.uint16": "uint16_t",
               "torch.uint32": "uint32_t", "torch.uint64": "uint64_t"}


class DVTorchVector(DVVectorAdaptor):
    """Wraps a contiguous 1-D torch tensor on the GPU (kept alive by this object)."""

    def __init__(self, tensor):
        if not tensor.is_contiguous():
            raise ValueError("tensor must be contiguous")
        elem = _TORCH_TO_C[str(tensor.dtype)]
        super().__init__(elem, tensor.numel(), tensor.data_ptr(), owner=tensor)
