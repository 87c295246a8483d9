// Sorting (reference: sort.h:1-24).  All sorts are stable.  Less / Greater on
// 8..64-bit integer and float/double keys run the onesweep radix sort.
#ifndef _TRTC_sort_h
#define _TRTC_sort_h

#include "TRTC_api.h"
#include "TRTCContext.h"
#include "DeviceViewable.h"
#include "DVVector.h"
#include "functor.h"

bool THRUST_RTC_API TRTC_Sort(DVVectorLike& vec);
bool THRUST_RTC_API TRTC_Sort(DVVectorLike& vec, const Functor& comp);
bool THRUST_RTC_API TRTC_Sort_By_Key(DVVectorLike& keys, DVVectorLike& values);
bool THRUST_RTC_API TRTC_Sort_By_Key(DVVectorLike& keys, DVVectorLike& values, const Functor& comp);

bool THRUST_RTC_API TRTC_Is_Sorted(const DVVectorLike& vec, bool& result);
bool THRUST_RTC_API TRTC_Is_Sorted(const DVVectorLike& vec, const Functor& comp, bool& result);

// ---- extension (no counterpart in the reference) ----
// Segmented stable sort by key: segment s is [offsets[s], offsets[s+1]) of
// keys/values (offsets: host array of num_segments + 1 entries).  Each
// segment ends up exactly as a separate TRTC_Sort_By_Key call would leave it.
bool THRUST_RTC_API TRTC_Segmented_Sort_By_Key(DVVectorLike& keys, DVVectorLike& values,
                                               const size_t* offsets, size_t num_segments,
                                               const Functor& comp);


// Whether the vector is sorted / the length of its sorted prefix (sort.h:16-19).
bool THRUST_RTC_API TRTC_Is_Sorted_Until(const DVVectorLike& vec, size_t& result);
bool THRUST_RTC_API TRTC_Is_Sorted_Until(const DVVectorLike& vec, const Functor& comp, size_t& result);

#endif
