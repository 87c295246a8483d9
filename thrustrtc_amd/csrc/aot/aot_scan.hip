// Scan and reduce for the built-in arithmetic types with Plus / Maximum / Minimum.
#include "aot_common.h"

namespace {
template <class T, class Op> void reg_scan(const char* tn, const char* on)
{
    std::string name = "trtc_scan_k<" + vv(tn) + ", " + vv(tn) + ", " + on + ", " + tn + ">";
    TRTC_REG(name.c_str(), (&trtc_scan_k<VectorView<T>, VectorView<T>, Op, T>));
    std::string rname = "trtc_reduce_k<" + vv(tn) + ", " + on + ", " + tn + ">";
    TRTC_REG(rname.c_str(), (&trtc_reduce_k<VectorView<T>, Op, T>));
}
template <class T> void reg_type(const char* tn)
{
    reg_scan<T, Plus>(tn, "Plus");
    reg_scan<T, Maximum>(tn, "Maximum");
    reg_scan<T, Minimum>(tn, "Minimum");
}
struct RegScan
{
    RegScan()
    {
        reg_type<int32_t>("int32_t");
        reg_type<uint32_t>("uint32_t");
        reg_type<int64_t>("int64_t");
        reg_type<uint64_t>("uint64_t");
        reg_type<float>("float");
        reg_type<double>("double");
    }
} reg_scan_all;
}  // namespace
