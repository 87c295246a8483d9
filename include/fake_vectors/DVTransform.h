// A read-only vector whose element i is op(vec_in[i]) converted to elem_cls
// (reference: fake_vectors/DVTransform.h:1-20).  Device view:
// TransformView<T, VIn, Op>.
#ifndef _TRTC_DVTransform_h
#define _TRTC_DVTransform_h

#include "../DVVector.h"
#include "../functor.h"

class THRUST_RTC_API DVTransform : public DVVectorLike
{
public:
    DVTransform(const DVVectorLike& vec_in, const char* elem_cls, const Functor& op);
    virtual ViewBuf view() const;

private:
    ViewBuf m_view_in, m_view_op;
    size_t m_offsets[3];
};

#endif
