"""Fake vectors (reference: python/ThrustRTC/FakeVectors.py): vector-likes
whose elements are computed from an index on the device.  Each keeps the
objects it views alive for as long as it lives."""
import ctypes
from .native import native, check_cptr
from .DVVector import DVVectorLike
from .DeviceViewable import DeviceViewable
from .utils import ObjArray, StrArray


class DVConstant(DVVectorLike):
    """Every element is `dvobj`'s value (size defaults to unbounded)."""

    def __init__(self, dvobj, size=-1):
        self.m_dvobj = dvobj
        self.m_cptr = check_cptr(native.n_dvconstant_create(dvobj.m_cptr, ctypes.c_ulonglong(size).value))


class DVCounter(DVVectorLike):
    """Element i is init + i."""

    def __init__(self, dvobj_init, size=-1):
        self.m_dvobj_init = dvobj_init
        self.m_cptr = check_cptr(native.n_dvcounter_create(dvobj_init.m_cptr, ctypes.c_ulonglong(size).value))


class DVDiscard(DVVectorLike):
    """Write-only; what is written is dropped."""

    def __init__(self, elem_cls, size=-1):
        self.m_cptr = check_cptr(native.n_dvdiscard_create(elem_cls.encode("utf-8"), ctypes.c_ulonglong(size).value))


class DVPermutation(DVVectorLike):
    """Element i is vec_value[vec_index[i]]."""

    def __init__(self, vec_value, vec_index):
        self.m_vec_value = vec_value
        self.m_vec_index = vec_index
        self.m_cptr = check_cptr(native.n_dvpermutation_create(vec_value.m_cptr, vec_index.m_cptr))


class DVReverse(DVVectorLike):
    """Element i is vec_value[size - 1 - i]."""

    def __init__(self, vec_value):
        self.m_vec_value = vec_value
        self.m_cptr = check_cptr(native.n_dvreverse_create(vec_value.m_cptr))


class DVTransform(DVVectorLike):
    """Element i is op(vec_in[i]) converted to elem_cls."""

    def __init__(self, vec_in, elem_cls, op):
        self.m_vec_in = vec_in
        self.m_op = op
        self.m_cptr = check_cptr(native.n_dvtransform_create(vec_in.m_cptr, elem_cls.encode("utf-8"), op.m_cptr))


class DVZipped(DVVectorLike):
    """Vectors of one length seen as one vector of structs with members named
    by elem_names."""

    def __init__(self, vecs, elem_names):
        self.m_vecs = vecs
        o_vecs = ObjArray(vecs)
        o_names = StrArray(elem_names)
        self.m_cptr = check_cptr(native.n_dvzipped_create(o_vecs.m_cptr, o_names.m_cptr))


class DVCustomVector(DVVectorLike):
    """Element `name_idx` is what code_body returns, with the objects of
    arg_map (name -> device-viewable) in scope."""

    def __init__(self, arg_map, name_idx, code_body, elem_cls, size=-1, read_only=True):
        self.m_arg_map = arg_map
        o_names = StrArray(list(arg_map.keys()))
        o_elems = ObjArray(list(arg_map.values()))
        self.m_cptr = check_cptr(native.n_dvcustomvector_create(
            o_elems.m_cptr, o_names.m_cptr, name_idx.encode("utf-8"), code_body.encode("utf-8"),
            elem_cls.encode("utf-8"), ctypes.c_ulonglong(size).value, 1 if read_only else 0))


class DVTuple(DeviceViewable):
    """A struct of named device-viewable values (reference:
    python/ThrustRTC/DVTuple.py); the same members as a DVZipped element make
    the same struct, so it can be copied into zipped vectors."""

    def __init__(self, elem_map):
        self.m_elem_map = elem_map
        o_names = StrArray(list(elem_map.keys()))
        o_elems = ObjArray(list(elem_map.values()))
        self.m_cptr = check_cptr(native.n_dvtuple_create(o_elems.m_cptr, o_names.m_cptr))
