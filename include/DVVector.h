// Device vectors (reference: DVVector.h:1-60).  DVVector owns device memory
// (zero-filled when created without host data, as in the reference);
// DVVectorAdaptor borrows a device pointer (e.g. a torch tensor's storage).
#ifndef _DVVector_h
#define _DVVector_h

#include "TRTC_api.h"
#include "DeviceViewable.h"
#include "TRTCContext.h"

class THRUST_RTC_API DVVectorLike : public DeviceViewable
{
public:
    const std::string& name_elem_cls() const { return m_elem_cls; }
    const std::string& name_ref_type() const { return m_ref_type; }
    size_t elem_size() const { return m_elem_size; }
    size_t size() const { return m_size; }

    DVVectorLike(const char* elem_cls, const char* ref_type, size_t size);
    virtual ~DVVectorLike() {}
    virtual bool is_readable() const { return true; }
    virtual bool is_writable() const { return false; }

protected:
    std::string m_elem_cls;
    std::string m_ref_type;
    size_t m_elem_size;
    size_t m_size;
};

class THRUST_RTC_API DVVector : public DVVectorLike
{
public:
    void* data() const { return m_data; }

    DVVector(const char* elem_cls, size_t size, void* hdata = nullptr);
    ~DVVector();
    virtual bool is_writable() const { return true; }

    void to_host(void* hdata, size_t begin = 0, size_t end = (size_t)(-1)) const;
    virtual ViewBuf view() const;

private:
    void* m_data;
};

class THRUST_RTC_API DVVectorAdaptor : public DVVectorLike
{
public:
    void* data() const { return m_data; }

    DVVectorAdaptor(const char* elem_cls, size_t size, void* ddata);
    DVVectorAdaptor(const DVVector& vec, size_t begin = 0, size_t end = (size_t)(-1));
    DVVectorAdaptor(const DVVectorAdaptor& vec, size_t begin = 0, size_t end = (size_t)(-1));

    virtual bool is_writable() const { return true; }
    virtual ViewBuf view() const;

private:
    void* m_data;
};

// A sub-range [begin, end) of any vector-like (reference: fake_vectors/DVRange.h).
class THRUST_RTC_API DVRange : public DVVectorLike
{
public:
    DVRange(const DVVectorLike& vec_value, size_t begin = 0, size_t end = (size_t)(-1));
    virtual bool is_readable() const { return m_readable; }
    virtual bool is_writable() const { return m_writable; }
    virtual ViewBuf view() const;
    // device address of element 0 when the underlying vector is contiguous
    void* contiguous() const { return m_contig; }

private:
    bool m_readable, m_writable;
    ViewBuf m_view_vec_value;
    size_t m_begin, m_end;
    void* m_contig;
};

#endif
