// vec[i] = value (reference: fill.h:9).
#ifndef _TRTC_fill_h
#define _TRTC_fill_h

#include "TRTC_api.h"
#include "TRTCContext.h"
#include "DeviceViewable.h"
#include "DVVector.h"
#include "functor.h"

bool THRUST_RTC_API TRTC_Fill(DVVectorLike& vec, const DeviceViewable& value);

#endif
