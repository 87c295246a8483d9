// Device-viewable objects: anything that can be passed to a kernel as a byte
// blob ("view") together with the name of its device-side type.
// Interface kept identical to the reference (DeviceViewable.h:1-175) so that
// C++, JNI and C++/CLI callers compile unchanged.
#ifndef _DeviceViewable_h
#define _DeviceViewable_h

#include <string>
#include <vector>
#include <cstdint>
#include <cstring>

typedef std::vector<char> ViewBuf;

// root class of all device-viewable objects
class DeviceViewable
{
public:
    DeviceViewable() {}
    virtual ~DeviceViewable() {}
    virtual ViewBuf view() const = 0;
    const std::string& name_view_cls() const { return m_name_view_cls; }

protected:
    std::string m_name_view_cls;
};

struct CapturedDeviceViewable
{
    const char* obj_name;
    const DeviceViewable* obj;
};

class SomeDeviceViewable : public DeviceViewable
{
public:
    SomeDeviceViewable(const char* name_view_cls, const void* data_view = "", size_t size_view = 1)
    {
        m_name_view_cls = name_view_cls;
        m_view_buf.resize(size_view);
        memcpy(m_view_buf.data(), data_view, size_view);
    }
    virtual ViewBuf view() const { return m_view_buf; }

private:
    ViewBuf m_view_buf;
};

#define TRTC_DECLARE_DV_BASIC(clsname, type)                                   \
    class clsname : public SomeDeviceViewable                                 \
    {                                                                         \
    public:                                                                   \
        clsname(type in) : SomeDeviceViewable(#type, &in, sizeof(type)) {}    \
    };

TRTC_DECLARE_DV_BASIC(DVChar, char)
TRTC_DECLARE_DV_BASIC(DVSChar, signed char)
TRTC_DECLARE_DV_BASIC(DVUChar, unsigned char)
TRTC_DECLARE_DV_BASIC(DVShort, short)
TRTC_DECLARE_DV_BASIC(DVUShort, unsigned short)
TRTC_DECLARE_DV_BASIC(DVInt, int)
TRTC_DECLARE_DV_BASIC(DVUInt, unsigned int)
TRTC_DECLARE_DV_BASIC(DVLong, long)
TRTC_DECLARE_DV_BASIC(DVULong, unsigned long)
TRTC_DECLARE_DV_BASIC(DVLongLong, long long)
TRTC_DECLARE_DV_BASIC(DVULongLong, unsigned long long)
TRTC_DECLARE_DV_BASIC(DVFloat, float)
TRTC_DECLARE_DV_BASIC(DVDouble, double)
TRTC_DECLARE_DV_BASIC(DVBool, bool)

TRTC_DECLARE_DV_BASIC(DVInt8, int8_t)
TRTC_DECLARE_DV_BASIC(DVUInt8, uint8_t)
TRTC_DECLARE_DV_BASIC(DVInt16, int16_t)
TRTC_DECLARE_DV_BASIC(DVUInt16, uint16_t)
TRTC_DECLARE_DV_BASIC(DVInt32, int32_t)
TRTC_DECLARE_DV_BASIC(DVUInt32, uint32_t)
TRTC_DECLARE_DV_BASIC(DVInt64, int64_t)
TRTC_DECLARE_DV_BASIC(DVUInt64, uint64_t)

TRTC_DECLARE_DV_BASIC(DVSizeT, size_t)

// Complex scalars keep the reference's type names so that code strings and
// Python callers that name them keep working; on the device they are plain
// {x, y} structs.
class DVComplex64 : public DeviceViewable
{
public:
    DVComplex64(float real, float imag) : m_real(real), m_imag(imag) { m_name_view_cls = "cuFloatComplex"; }
    virtual ViewBuf view() const
    {
        ViewBuf buf(sizeof(float) * 2);
        memcpy(&buf[0], &m_real, sizeof(float));
        memcpy(&buf[sizeof(float)], &m_imag, sizeof(float));
        return buf;
    }

private:
    float m_real, m_imag;
};

class DVComplex128 : public DeviceViewable
{
public:
    DVComplex128(double real, double imag) : m_real(real), m_imag(imag) { m_name_view_cls = "cuDoubleComplex"; }
    virtual ViewBuf view() const
    {
        ViewBuf buf(sizeof(double) * 2);
        memcpy(&buf[0], &m_real, sizeof(double));
        memcpy(&buf[sizeof(double)], &m_imag, sizeof(double));
        return buf;
    }

private:
    double m_real, m_imag;
};

// Wraps a view buffer of a named scalar type into a new device-viewable.
DeviceViewable* dv_from_viewbuf(const ViewBuf& buf, const char* type);

#endif
