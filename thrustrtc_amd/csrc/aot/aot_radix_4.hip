// Onesweep radix sort kernels for 4-byte keys (all value widths, both
// status widths).
#include "aot_common.h"

namespace {
template <int VB, class S> void reg_pass(const char* sn)
{
    // 512 threads: key kind read at run time; 1024 threads: one per key kind
    std::string name = std::string("trtc_radix_pass_k<4, ") + std::to_string(VB) + ", " + sn + ", ";
    TRTC_REG((name + "512, -1>").c_str(), (&trtc_radix_pass_k<4, VB, S, 512, -1>));
    TRTC_REG((name + "1024, 0>").c_str(), (&trtc_radix_pass_k<4, VB, S, 1024, 0>));
    TRTC_REG((name + "1024, 1>").c_str(), (&trtc_radix_pass_k<4, VB, S, 1024, 1>));
    TRTC_REG((name + "1024, 2>").c_str(), (&trtc_radix_pass_k<4, VB, S, 1024, 2>));
}
template <int VB, class S> void reg_long(const char* sn)
{
    std::string name = std::string("trtc_radix_pass_long_k<") + std::to_string(VB) + ", " + sn + ", ";
    TRTC_REG((name + "0>").c_str(), (&trtc_radix_pass_long_k<VB, S, 0>));
    TRTC_REG((name + "1>").c_str(), (&trtc_radix_pass_long_k<VB, S, 1>));
    TRTC_REG((name + "2>").c_str(), (&trtc_radix_pass_long_k<VB, S, 2>));
}
struct RegRadix
{
    RegRadix()
    {
        TRTC_REG("trtc_radix_hist_k<4>", (&trtc_radix_hist_k<4>));
        reg_pass<0, unsigned int>("unsigned int");
        reg_pass<1, unsigned int>("unsigned int");
        reg_pass<2, unsigned int>("unsigned int");
        reg_pass<4, unsigned int>("unsigned int");
        reg_pass<8, unsigned int>("unsigned int");
        reg_pass<0, unsigned long long>("unsigned long long");
        reg_pass<1, unsigned long long>("unsigned long long");
        reg_pass<2, unsigned long long>("unsigned long long");
        reg_pass<4, unsigned long long>("unsigned long long");
        reg_pass<8, unsigned long long>("unsigned long long");
        reg_long<0, unsigned int>("unsigned int");
        reg_long<4, unsigned int>("unsigned int");
        reg_long<0, unsigned long long>("unsigned long long");
        reg_long<4, unsigned long long>("unsigned long long");
    }
} reg_radix;
}  // namespace
