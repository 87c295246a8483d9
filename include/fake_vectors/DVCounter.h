// A read-only vector whose element i is init + i (reference:
// fake_vectors/DVCounter.h:1-20).  Device view: CounterView<T>.
#ifndef _TRTC_DVCounter_h
#define _TRTC_DVCounter_h

#include "../DVVector.h"

class THRUST_RTC_API DVCounter : public DVVectorLike
{
public:
    DVCounter(const DeviceViewable& dvobj_init, size_t size = (size_t)(-1));
    virtual ViewBuf view() const;

private:
    ViewBuf m_value_init;
    size_t m_offsets[3];
};

#endif
