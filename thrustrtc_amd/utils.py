"""Argument-list handles (reference: python/ThrustRTC/utils.py)."""
import ctypes
from .native import native


class StrArray:
    def __init__(self, arr):
        enc = [s.encode("utf-8") for s in arr]
        c = (ctypes.c_char_p * max(1, len(enc)))(*enc)
        self.m_cptr = native.n_string_array_create(len(enc), c)

    def __del__(self):
        native.n_string_array_destroy(self.m_cptr)


class ObjArray:
    def __init__(self, arr):
        ptrs = [obj.m_cptr for obj in arr]
        c = (ctypes.c_void_p * max(1, len(ptrs)))(*ptrs)
        self.m_cptr = native.n_pointer_array_create(len(ptrs), c)

    def __del__(self):
        native.n_pointer_array_destroy(self.m_cptr)


class Dim3:
    def __init__(self, t):
        tp = [1, 1, 1]
        if type(t) is tuple:
            tp[0:len(t)] = t[:]
        else:
            tp[0] = t
        self.m_cptr = native.n_dim3_create(tp[0], tp[1], tp[2])

    def __del__(self):
        native.n_dim3_destroy(self.m_cptr)
