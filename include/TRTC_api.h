// Export macro of the C++ API (reference: TRTC_api.h:1-20).
#ifndef _TRTC_API_h
#define _TRTC_API_h
#ifndef THRUST_RTC_API
#define THRUST_RTC_API __attribute__((visibility("default")))
#endif
#endif
