// Digit passes of the batched segments for the tile shapes the batch
// autotuner and the request server choose among, instantiated per key kind
// (TRTC_SEG_KIND: 0 unsigned, 1 signed, 2 float) by aot_seg_pass_k*.hip.
#pragma once
#include "aot_common.h"

namespace {
template <int KB, int VB, int BLOCK, int ITEMS, int KIND> void reg_seg_pass()
{
    std::string p = std::to_string(KB) + ", " + std::to_string(VB) + ", " + std::to_string(BLOCK) + ", " +
                    std::to_string(ITEMS) + ", " + std::to_string(KIND) + ">";
    TRTC_REG(("trtc_seg_radix_pass_k<" + p).c_str(), (&trtc_seg_radix_pass_k<KB, VB, BLOCK, ITEMS, KIND>));
}
template <int VB, int KIND> void reg_seg_pass_long()
{
    std::string p = std::to_string(VB) + ", " + std::to_string(KIND) + ">";
    TRTC_REG(("trtc_seg_radix_pass_long_k<" + p).c_str(), (&trtc_seg_radix_pass_long_k<VB, KIND>));
}
template <int KIND> struct RegSegPass
{
    RegSegPass()
    {
        // 4-byte keys with 4-byte or no values: the long-tile pass (1024 x 32)
        reg_seg_pass_long<4, KIND>();
        reg_seg_pass_long<0, KIND>();
        // 4-byte keys: keys only and 4-byte values
        reg_seg_pass<4, 4, 1024, 16, KIND>();
        reg_seg_pass<4, 4, 512, 16, KIND>();
        reg_seg_pass<4, 4, 512, 8, KIND>();
        reg_seg_pass<4, 4, 256, 16, KIND>();
        reg_seg_pass<4, 4, 256, 8, KIND>();
        reg_seg_pass<4, 0, 1024, 16, KIND>();
        reg_seg_pass<4, 0, 512, 16, KIND>();
        reg_seg_pass<4, 0, 256, 16, KIND>();
        reg_seg_pass<4, 0, 256, 8, KIND>();
        // 8-byte keys with 8-byte or no values
        reg_seg_pass<8, 8, 1024, 8, KIND>();
        reg_seg_pass<8, 8, 512, 8, KIND>();
        reg_seg_pass<8, 8, 256, 8, KIND>();
        reg_seg_pass<8, 0, 1024, 8, KIND>();
        reg_seg_pass<8, 0, 512, 8, KIND>();
        reg_seg_pass<8, 0, 256, 8, KIND>();
    }
};
}  // namespace
