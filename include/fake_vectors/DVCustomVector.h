// A vector defined by user code: element idx is what code_body returns for
// index name_idx, with the captured objects in scope (reference:
// fake_vectors/DVCustomVector.h:1-25).  read_only = false makes the body
// return a reference that kernels may assign through.
#ifndef _TRTC_DVCustomVector_h
#define _TRTC_DVCustomVector_h

#include "../DVVector.h"

class THRUST_RTC_API DVCustomVector : public DVVectorLike
{
public:
    DVCustomVector(const std::vector<CapturedDeviceViewable>& arg_map, const char* name_idx, const char* code_body,
                   const char* elem_cls, size_t size = (size_t)(-1), bool read_only = true);
    virtual bool is_readable() const { return true; }
    virtual bool is_writable() const { return !m_read_only; }
    virtual ViewBuf view() const;

private:
    bool m_read_only;
    std::vector<ViewBuf> m_view_args;
    std::vector<size_t> m_arg_offsets;
    size_t m_offsets[3];
};

#endif
