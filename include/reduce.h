// Reductions (reference: reduce.h:1-20).
#ifndef _TRTC_reduce_h
#define _TRTC_reduce_h

#include "TRTC_api.h"
#include "TRTCContext.h"
#include "DeviceViewable.h"
#include "DVVector.h"
#include "functor.h"

bool THRUST_RTC_API TRTC_Reduce(const DVVectorLike& vec, ViewBuf& ret);
bool THRUST_RTC_API TRTC_Reduce(const DVVectorLike& vec, const DeviceViewable& init, ViewBuf& ret);
bool THRUST_RTC_API TRTC_Reduce(const DVVectorLike& vec, const DeviceViewable& init, const Functor& binary_op, ViewBuf& ret);

#endif
