"""Functors (reference: python/ThrustRTC/Functor.py)."""
from .native import native, check_cptr
from .DeviceViewable import DeviceViewable
from .utils import StrArray, ObjArray


class Functor(DeviceViewable):
    def __init__(self, arg_map, functor_params, code_body):
        self.m_arg_map = arg_map
        o_param_names = StrArray(list(arg_map.keys()))
        o_elems = ObjArray(list(arg_map.values()))
        o_functor_params = StrArray(functor_params)
        self.m_cptr = check_cptr(native.n_functor_create(o_elems.m_cptr, o_param_names.m_cptr,
                                                         o_functor_params.m_cptr, code_body.encode("utf-8")))


def _builtin(name):
    def __init__(self):
        self.m_cptr = check_cptr(native.n_built_in_functor_create(name.encode("utf-8")))
    return type(name, (DeviceViewable,), {"__init__": __init__})


Identity = _builtin("Identity")
Maximum = _builtin("Maximum")
Minimum = _builtin("Minimum")
EqualTo = _builtin("EqualTo")
NotEqualTo = _builtin("NotEqualTo")
Greater = _builtin("Greater")
Less = _builtin("Less")
GreaterEqual = _builtin("GreaterEqual")
LessEqual = _builtin("LessEqual")
Plus = _builtin("Plus")
Minus = _builtin("Minus")
Multiplies = _builtin("Multiplies")
Divides = _builtin("Divides")
Modulus = _builtin("Modulus")
Negate = _builtin("Negate")
