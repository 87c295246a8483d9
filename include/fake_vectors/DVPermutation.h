// Element i is vec_value[vec_index[i]] (reference:
// fake_vectors/DVPermutation.h:1-25).  Device view: PermutationView<V, I>.
#ifndef _TRTC_DVPermutation_h
#define _TRTC_DVPermutation_h

#include "../DVVector.h"

class THRUST_RTC_API DVPermutation : public DVVectorLike
{
public:
    DVPermutation(const DVVectorLike& vec_value, const DVVectorLike& vec_index);
    virtual bool is_readable() const { return m_readable; }
    virtual bool is_writable() const { return m_writable; }
    virtual ViewBuf view() const;

private:
    bool m_readable, m_writable;
    ViewBuf m_view_value, m_view_index;
    size_t m_offsets[3];
};

#endif
