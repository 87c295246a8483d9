// Batched-segment kernels: digit histograms, segmented scans and reductions
// for the built-in element types (the digit passes are in aot_seg_pass_k*.hip,
// one file per key kind so that they compile in parallel).
#include "aot_common.h"

namespace {
template <int KB> void reg_seg_hist()
{
    TRTC_REG(("trtc_seg_radix_hist_k<" + std::to_string(KB) + ">").c_str(), (&trtc_seg_radix_hist_k<KB>));
}
template <class T, int ITEMS> void reg_seg_scan(const char* tn, const char* op)
{
    std::string name = std::string("trtc_seg_scan_k<") + op + ", " + tn + ", " + std::to_string(ITEMS) + ">";
    if (std::string(op) == "Plus") TRTC_REG(name.c_str(), (&trtc_seg_scan_k<Plus, T, ITEMS>));
    else if (std::string(op) == "Maximum") TRTC_REG(name.c_str(), (&trtc_seg_scan_k<Maximum, T, ITEMS>));
    else TRTC_REG(name.c_str(), (&trtc_seg_scan_k<Minimum, T, ITEMS>));
}
template <class T> void reg_seg_reduce(const char* tn)
{
    TRTC_REG((std::string("trtc_seg_reduce_k<Plus, ") + tn + ">").c_str(), (&trtc_seg_reduce_k<Plus, T>));
    TRTC_REG((std::string("trtc_seg_reduce_k<Maximum, ") + tn + ">").c_str(), (&trtc_seg_reduce_k<Maximum, T>));
    TRTC_REG((std::string("trtc_seg_reduce_k<Minimum, ") + tn + ">").c_str(), (&trtc_seg_reduce_k<Minimum, T>));
}
struct RegSeg
{
    RegSeg()
    {
        reg_seg_hist<4>();
        reg_seg_hist<8>();
        for (const char* op : {"Plus", "Maximum", "Minimum"}) {
            reg_seg_scan<int32_t, 32>("int32_t", op);
            reg_seg_scan<uint32_t, 32>("uint32_t", op);
            reg_seg_scan<float, 32>("float", op);
            reg_seg_scan<int64_t, 16>("int64_t", op);
            reg_seg_scan<double, 16>("double", op);
        }
        reg_seg_reduce<int32_t>("int32_t");
        reg_seg_reduce<uint32_t>("uint32_t");
        reg_seg_reduce<float>("float");
        reg_seg_reduce<double>("double");
        reg_seg_reduce<int64_t>("int64_t");
    }
} reg_seg;
}  // namespace
