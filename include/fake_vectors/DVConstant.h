// A read-only vector whose every element is one value (reference:
// fake_vectors/DVConstant.h:1-20).  Device view: ConstantView<T>.
#ifndef _TRTC_DVConstant_h
#define _TRTC_DVConstant_h

#include "../DVVector.h"

class THRUST_RTC_API DVConstant : public DVVectorLike
{
public:
    DVConstant(const DeviceViewable& dvobj, size_t size = (size_t)(-1));
    virtual ViewBuf view() const;

private:
    ViewBuf m_value;
    size_t m_offsets[3];
};

#endif
