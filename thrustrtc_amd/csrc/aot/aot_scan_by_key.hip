// Scan-by-key with EqualTo / Plus for common key/value type pairs.
#include "aot_common.h"

namespace {
template <class K, class V> void reg_sbk(const char* kn, const char* vn)
{
    std::string name = "trtc_scan_by_key_k<" + vv(kn) + ", " + vv(vn) + ", " + vv(vn) + ", EqualTo, Plus, " + vn + ">";
    TRTC_REG(name.c_str(), (&trtc_scan_by_key_k<VectorView<K>, VectorView<V>, VectorView<V>, EqualTo, Plus, V>));
}
struct RegSbk
{
    RegSbk()
    {
        reg_sbk<int32_t, int32_t>("int32_t", "int32_t");
        reg_sbk<int32_t, uint32_t>("int32_t", "uint32_t");
        reg_sbk<int32_t, float>("int32_t", "float");
        reg_sbk<int32_t, double>("int32_t", "double");
        reg_sbk<uint32_t, int32_t>("uint32_t", "int32_t");
        reg_sbk<uint32_t, uint32_t>("uint32_t", "uint32_t");
        reg_sbk<uint32_t, float>("uint32_t", "float");
        reg_sbk<uint32_t, double>("uint32_t", "double");
    }
} reg_sbk_all;
}  // namespace
