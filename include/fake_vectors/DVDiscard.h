// A write-only vector that drops what is written to it (reference:
// fake_vectors/DVDiscard.h:1-18).  Device view: DiscardView<T>.
#ifndef _TRTC_DVDiscard_h
#define _TRTC_DVDiscard_h

#include "../DVVector.h"

class THRUST_RTC_API DVDiscard : public DVVectorLike
{
public:
    DVDiscard(const char* elem_cls, size_t size = (size_t)(-1));
    virtual bool is_readable() const { return false; }
    virtual bool is_writable() const { return true; }
    virtual ViewBuf view() const;
};

#endif
