// Element-wise exchange of two vectors (reference: swap.h:9).
#ifndef _TRTC_swap_h
#define _TRTC_swap_h

#include "TRTC_api.h"
#include "TRTCContext.h"
#include "DeviceViewable.h"
#include "DVVector.h"
#include "functor.h"

bool THRUST_RTC_API TRTC_Swap(DVVectorLike& vec1, DVVectorLike& vec2);

#endif
