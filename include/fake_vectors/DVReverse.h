// Element i is vec_value[size - 1 - i] (reference: fake_vectors/DVReverse.h:1-22).
// Device view: ReverseView<V>.
#ifndef _TRTC_DVReverse_h
#define _TRTC_DVReverse_h

#include "../DVVector.h"

class THRUST_RTC_API DVReverse : public DVVectorLike
{
public:
    DVReverse(const DVVectorLike& vec_value);
    virtual bool is_readable() const { return m_readable; }
    virtual bool is_writable() const { return m_writable; }
    virtual ViewBuf view() const;

private:
    bool m_readable, m_writable;
    ViewBuf m_view_value;
};

#endif
