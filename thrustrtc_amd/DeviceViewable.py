"""Device-viewable scalars (reference: python/ThrustRTC/DeviceViewable.py)."""
from .native import native, check_cptr


def ValueFromDVPtr(cptr):
    cls = native.n_dv_name_view_cls(cptr).decode("utf-8")
    getters = {
        "int8_t": native.n_dvint8_value, "uint8_t": native.n_dvuint8_value,
        "int16_t": native.n_dvint16_value, "uint16_t": native.n_dvuint16_value,
        "int32_t": native.n_dvint32_value, "uint32_t": native.n_dvuint32_value,
        "int64_t": native.n_dvint64_value, "uint64_t": native.n_dvuint64_value,
        "float": native.n_dvfloat_value, "double": native.n_dvdouble_value,
    }
    if cls in getters:
        return getters[cls](cptr)
    if cls == "bool":
        return native.n_dvbool_value(cptr) != 0
    if cls == "cuFloatComplex":
        return complex(native.n_dvcomplex64_real(cptr), native.n_dvcomplex64_imag(cptr))
    if cls == "cuDoubleComplex":
        return complex(native.n_dvcomplex128_real(cptr), native.n_dvcomplex128_imag(cptr))
    return "[Device-viewable object, type: %s]" % cls


class DeviceViewable:
    def name_view_cls(self):
        return native.n_dv_name_view_cls(self.m_cptr).decode("utf-8")

    def __del__(self):
        if getattr(self, "m_cptr", None):
            native.n_dv_destroy(self.m_cptr)

    def value(self):
        return "[Device-viewable object, type: %s]" % self.name_view_cls()


def _scalar(name, create, getter, conv=lambda v: v):
    def __init__(self, value):
        self.m_cptr = check_cptr(create(value))

    def value(self):
        return conv(getter(self.m_cptr))

    return type(name, (DeviceViewable,), {"__init__": __init__, "value": value})


DVInt8 = _scalar("DVInt8", native.n_dvint8_create, native.n_dvint8_value)
DVUInt8 = _scalar("DVUInt8", native.n_dvuint8_create, native.n_dvuint8_value)
DVInt16 = _scalar("DVInt16", native.n_dvint16_create, native.n_dvint16_value)
DVUInt16 = _scalar("DVUInt16", native.n_dvuint16_create, native.n_dvuint16_value)
DVInt32 = _scalar("DVInt32", native.n_dvint32_create, native.n_dvint32_value)
DVUInt32 = _scalar("DVUInt32", native.n_dvuint32_create, native.n_dvuint32_value)
DVInt64 = _scalar("DVInt64", native.n_dvint64_create, native.n_dvint64_value)
DVUInt64 = _scalar("DVUInt64", native.n_dvuint64_create, native.n_dvuint64_value)
DVFloat = _scalar("DVFloat", native.n_dvfloat_create, native.n_dvfloat_value)
DVDouble = _scalar("DVDouble", native.n_dvdouble_create, native.n_dvdouble_value)
DVBool = _scalar("DVBool", lambda v: native.n_dvbool_create(1 if v else 0), native.n_dvbool_value,
                 lambda v: v != 0)


class DVComplex64(DeviceViewable):
    def __init__(self, value):
        self.m_cptr = check_cptr(native.n_dvcomplex64_create(value.real, value.imag))

    def value(self):
        return complex(native.n_dvcomplex64_real(self.m_cptr), native.n_dvcomplex64_imag(self.m_cptr))


class DVComplex128(DeviceViewable):
    def __init__(self, value):
        self.m_cptr = check_cptr(native.n_dvcomplex128_create(value.real, value.imag))

    def value(self):
        return complex(native.n_dvcomplex128_real(self.m_cptr), native.n_dvcomplex128_imag(self.m_cptr))
