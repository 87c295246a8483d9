// Several vectors of one length seen as one vector of structs whose members
// are named by elem_names (reference: fake_vectors/DVZipped.h:1-25).  The
// element type and the reference type (members that refer to the zipped
// vectors' elements) are generated structs; readable / writable when every
// zipped vector is.
#ifndef _TRTC_DVZipped_h
#define _TRTC_DVZipped_h

#include "../DVVector.h"

class THRUST_RTC_API DVZipped : public DVVectorLike
{
public:
    DVZipped(const std::vector<DVVectorLike*>& vecs, const std::vector<const char*>& elem_names);
    virtual bool is_readable() const { return m_readable; }
    virtual bool is_writable() const { return m_writable; }
    virtual ViewBuf view() const;

private:
    bool m_readable, m_writable;
    std::vector<ViewBuf> m_view_elems;
    std::vector<size_t> m_offsets;
};

#endif
