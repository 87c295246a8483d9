// n_inclusive_scan_noalloc: TRTC_Inclusive_Scan for callers that cannot
// allocate device memory (code under HIP graph capture, tight serving loops,
// harnesses that forbid hipMalloc).  The scan's only device memory beyond the
// input and output, its decoupled-lookback state (one tile counter and two
// status words per tile), lives in the file-scope array `scratch` below.
//
// Limits and sharing:
//  * scratch is kScratchBytes = 1 MiB per GPU, held until the process ends;
//    with 16384-element tiles that covers n <= 65535 * 16384 for elements of
//    up to 4 bytes and n <= 32767 * 16384 for 8-byte elements.  Larger inputs
//    are rejected with an error (nothing is launched).
//  * every call on a GPU uses that GPU's one array, and each call clears the
//    part it uses with a fill kernel first: calls that overlap on one GPU
//    (other threads or streams) race and must be serialised by the caller.
//  * nothing is compiled either: the scan kernel must be an ahead-of-time one
//    (Plus / Maximum / Minimum on plain vectors of the built-in arithmetic
//    types) or one that an earlier n_inclusive_scan call with the same views
//    and functor compiled in this process.
#include <hip/hip_runtime.h>

#include <mutex>

#include "trtc_capi.h"
#include "../host/trtc_internal.h"

namespace {

constexpr size_t kScratchBytes = (size_t)1 << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mtx;
void* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device: looked up once per
// device with that device current (each GPU holds its own copy of the array
// at its own address).  The array's name is only a host-side handle; device
// work is given this pointer.
void* scratch_on_current_device()
{
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
        TRTC_ERR("n_inclusive_scan_noalloc: no current device");
        return nullptr;
    }
    std::lock_guard<std::mutex> lk(g_mtx);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            TRTC_ERR("n_inclusive_scan_noalloc: hipGetSymbolAddress failed on device %d: %s", dev,
                     hipGetErrorString(e));
            return nullptr;
        }
        g_scratch_ptr[dev] = p;
    }
    return g_scratch_ptr[dev];
}

}  // namespace

extern "C" TRTC_CAPI int n_inclusive_scan_noalloc(void* ptr_vec_in, void* ptr_vec_out, void* ptr_binary_op)
{
    if (!ptr_vec_in || !ptr_vec_out) return -1;
    if (!trtc_host::ensure_init()) return -1;
    void* p = scratch_on_current_device();
    if (!p) return -1;
    const DVVectorLike& in = *(const DVVectorLike*)ptr_vec_in;
    DVVectorLike& out = *(DVVectorLike*)ptr_vec_out;
    bool ok = ptr_binary_op
                  ? trtc_host::inclusive_scan_fixed_scratch(in, out, *(const Functor*)ptr_binary_op, p, kScratchBytes)
                  : trtc_host::inclusive_scan_fixed_scratch(in, out, Functor("Plus"), p, kScratchBytes);
    return ok ? 0 : -1;
}
