// Utility kernels (fill / iota / byte gather).
#define TRTC_DEFINE_UTIL_KERNELS
#include "aot_common.h"

namespace {
struct RegUtil
{
    RegUtil()
    {
        TRTC_REG("trtc_fill_u32", &trtc_fill_u32);
        TRTC_REG("trtc_radix_copyback_k", &trtc_radix_copyback_k);
        TRTC_REG("trtc_clear_u32_agent", &trtc_clear_u32_agent);
        TRTC_REG("trtc_iota_u32", &trtc_iota_u32);
        TRTC_REG("trtc_copy_bytes16", &trtc_copy_bytes16);
        TRTC_REG("trtc_gather_bytes", &trtc_gather_bytes);
    }
} reg_util;
}  // namespace
