// A struct of named device-viewable values (reference: DVTuple.h:8-18).  Its
// device type is the struct the same members in the same order generate for a
// DVZipped element, so a DVTuple is assignable to a zipped vector's elements.
#ifndef _TRTC_DVTuple_h
#define _TRTC_DVTuple_h

#include "TRTC_api.h"
#include "DeviceViewable.h"
#include "TRTCContext.h"

class THRUST_RTC_API DVTuple : public DeviceViewable
{
public:
    DVTuple(const std::vector<CapturedDeviceViewable>& elem_map);
    virtual ViewBuf view() const;

private:
    std::vector<ViewBuf> m_view_elems;
    std::vector<size_t> m_offsets;
};

#endif
