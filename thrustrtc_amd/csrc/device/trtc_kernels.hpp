// __global__ entry points of every primitive-core kernel.
//
// The same templates are instantiated two ways: ahead of time by hipcc for the
// built-in element types and functors (thrustrtc_amd/csrc/aot/*.hip), and at
// run time by hipRTC for user functors, fancy views and other type mixes, where
// the host asks for the instantiation by its name expression, e.g.
//   trtc_scan_k<VectorView<int32_t>, VectorView<int32_t>, Plus, int32_t>
// Both paths look kernels up by that same string.
#pragma once
#include "trtc_scan.hpp"
#include "trtc_reduce.hpp"
#include "trtc_copy_if.hpp"
#include "trtc_scan_by_key.hpp"
#include "trtc_radix.hpp"
#include "trtc_msort.hpp"
#include "trtc_select.hpp"
#include "trtc_merge.hpp"

#define TRTC_BLOCK 256

// Non-template utility kernels: defined in exactly one translation unit.
#if defined(TRTC_DEFINE_UTIL_KERNELS)
extern "C" __global__ void __launch_bounds__(TRTC_BLOCK) trtc_fill_u32(unsigned* p, size_t n, unsigned v)
{
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = v;
}

// Zeroes n 32-bit words with relaxed agent-scope atomic stores: lookback
// state cleared this way is seen by the next kernel's agent-scope atomics
// however the memory is cached (used for the library's static scratch array,
// which lives in a code object's global segment rather than in hipMalloc'd
// memory).
extern "C" __global__ void __launch_bounds__(TRTC_BLOCK) trtc_clear_u32_agent(unsigned* p, size_t n)
{
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        __hip_atomic_store(p + i, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Copy-back of a single radix sort's plan (nothing when the plan needs none).
extern "C" __global__ void __launch_bounds__(TRTC_BLOCK) trtc_radix_copyback_k(trtc::RadixCopyArgs a)
{
    trtc::radix_copyback(a);
}

// Byte-wise gather for value types whose size is not 1/2/4/8:
//   dst[i] = src[idx[i]], elem_size bytes each.
extern "C" __global__ void __launch_bounds__(TRTC_BLOCK)
trtc_gather_bytes(const unsigned char* src, unsigned char* dst, const unsigned* idx, size_t n, size_t elem_size)
{
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const unsigned char* s = src + (size_t)idx[i] * elem_size;
        unsigned char* d = dst + i * elem_size;
        for (size_t b = 0; b < elem_size; b++) d[b] = s[b];
    }
}

// Contiguous copy of nbytes between 16-byte aligned buffers: 16-byte vectors
// grid-stride (one per thread when the grid covers them all, as the host
// launches it), then the last nbytes % 16 bytes by the first threads of
// workgroup 0.
extern "C" __global__ void __launch_bounds__(TRTC_BLOCK)
trtc_copy_bytes16(const unsigned char* src, unsigned char* dst, size_t nbytes)
{
    typedef unsigned v4u __attribute__((ext_vector_type(4)));
    const size_t nvec = nbytes / 16;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < nvec; i += stride)
        ((v4u*)dst)[i] = __builtin_nontemporal_load((const v4u*)src + i);
    if (blockIdx.x == 0 && threadIdx.x < (nbytes & 15)) dst[nvec * 16 + threadIdx.x] = src[nvec * 16 + threadIdx.x];
}

// dst[i] = i (the payload of a keys-with-index sort)
extern "C" __global__ void __launch_bounds__(TRTC_BLOCK) trtc_iota_u32(unsigned* dst, size_t n)
{
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = (unsigned)i;
}
#endif

// Single scans and compactions use 512-thread tiles (8192 elements of 4
// bytes): half as many lookbacks as 256-thread tiles.
#define TRTC_SCAN_BLOCK 512

// The whole-input scan takes 1024-thread tiles for 8-byte elements (2^28
// int64: 1.05 -> 0.90 ms against 512 x 16); 4-byte ones keep 512 x 32 (1024
// x 32: 0.88 -> 0.89-0.93 ms), wider ones 512 threads for LDS.
template <class T> struct ScanBlock
{
    static constexpr int value = sizeof(T) == 8 ? 1024 : TRTC_SCAN_BLOCK;
};

template <class Src, class Out, class Op, class T>
__global__ void __launch_bounds__(ScanBlock<T>::value) trtc_scan_k(Src src, Out out, Op op, T init, trtc::ScanArgs a)
{
    trtc::scan_tile<ScanBlock<T>::value, trtc::ScanItems<T>::value, T>(src, out, op, init, a);
}

template <class Src, class Op, class T>
__global__ void __launch_bounds__(TRTC_BLOCK) trtc_reduce_k(Src src, Op op, T init, trtc::ReduceArgs a)
{
    trtc::ReduceDispatch<TRTC_BLOCK, Src, Op, T>::template run<trtc::ReduceItems<T>::value>(src, op, init, a);
}

// Position of the first smallest (MAX = 0) / first largest (MAX = 1) element
// of a contiguous arithmetic vector; the (index, value) pair lands in a.result.
template <class Vec, class Comp, int MAX>
__global__ void __launch_bounds__(TRTC_BLOCK) trtc_extreme_k(Vec vec, Comp comp, trtc::ReduceArgs a)
{
    typedef typename Vec::value_t T;
    static_assert(trtc::IsArith<T>::value, "trtc_extreme_k takes arithmetic vectors");
    trtc::extreme_streaming<TRTC_BLOCK, T, Comp, MAX != 0>(vec._data, comp, a);
}

// Copy_If (and Remove) in 1024-thread tiles: 2^29 int32 x % 3 == 0 0.805 ->
// 0.745 ms against 512 threads x 32 items; items per thread follow the wider
// of the input and the stencil (the LDS transpose buffer holds either).
#define TRTC_COPY_IF_BLOCK 1024
template <class In, class Sten, class Out, class Pred>
__global__ void __launch_bounds__(TRTC_COPY_IF_BLOCK)
trtc_copy_if_k(In in, Sten sten, Out out, Pred pred, trtc::CopyIfArgs a)
{
    typedef typename In::value_t T;
    typedef typename Sten::value_t ST;
    constexpr int ITEMS = sizeof(T) >= sizeof(ST) ? trtc::CopyIfItems<T>::value : trtc::CopyIfItems<ST>::value;
    trtc::copy_if_tile<TRTC_COPY_IF_BLOCK, ITEMS, T>(in, sten, out, pred, a);
}

// Selections on the compaction machinery (Unique, Partition, Reduce_By_Key)
// and the by-key scan.  Items per thread follow the wider of the key and value
// (or input and stencil) types.  Unique takes 1024-thread tiles (2^28 int32
// 0.360 -> 0.339 ms); the others keep 512: at 1024 Reduce_By_Key and
// Partition measured the same, Scan_By_Key slower (0.70 -> 0.86 ms) and
// Unique_By_Key spills at 128 VGPRs.
#define TRTC_SEL_BLOCK 1024
template <class KIn, class KOut, class Pred>
__global__ void __launch_bounds__(TRTC_SEL_BLOCK) trtc_unique_k(KIn kin, KOut kout, Pred pred, trtc::CopyIfArgs a)
{
    typedef typename KIn::value_t K;
    trtc::unique_tile<TRTC_SEL_BLOCK, trtc::CopyIfItems<K>::value, false, K, K>(kin, kin, kout, kout, pred, a);
}

// Block of the by-key scan: 1024 threads when the wider of the key and value
// types is 8 bytes (2^27 int64 keys and values 0.81 -> 0.76 ms), as the
// whole-input scan.  Unique_By_Key and Partition measured slower that way
// (0.62 -> 0.64, 0.68 -> 0.71 ms) and Reduce_By_Key the same; they keep 512.
template <class A, class B> struct WideBlock
{
    static constexpr int value = (sizeof(A) >= sizeof(B) ? sizeof(A) : sizeof(B)) == 8 ? 1024 : TRTC_SCAN_BLOCK;
};

template <class KIn, class VIn, class KOut, class VOut, class Pred>
__global__ void __launch_bounds__(TRTC_SCAN_BLOCK)
trtc_unique_by_key_k(KIn kin, VIn vin, KOut kout, VOut vout, Pred pred, trtc::CopyIfArgs a)
{
    typedef typename KIn::value_t K;
    typedef typename VIn::value_t V;
    constexpr int ITEMS = sizeof(K) >= sizeof(V) ? trtc::CopyIfItems<K>::value : trtc::CopyIfItems<V>::value;
    trtc::unique_tile<TRTC_SCAN_BLOCK, ITEMS, true, K, V>(kin, vin, kout, vout, pred, a);
}

template <class In, class Sten, class TOut, class FOut, class Pred>
__global__ void __launch_bounds__(TRTC_SCAN_BLOCK)
trtc_partition_k(In in, Sten sten, TOut tout, FOut fout, Pred pred, trtc::CopyIfArgs a)
{
    typedef typename In::value_t T;
    typedef typename Sten::value_t ST;
    constexpr int ITEMS = sizeof(T) >= sizeof(ST) ? trtc::CopyIfItems<T>::value : trtc::CopyIfItems<ST>::value;
    trtc::partition_tile<TRTC_SCAN_BLOCK, ITEMS, T>(in, sten, tout, fout, pred, a);
}

template <class Keys, class Vals, class KOut, class VOut, class Pred, class Op, class V>
__global__ void __launch_bounds__(TRTC_SCAN_BLOCK)
trtc_reduce_by_key_k(Keys keys, Vals vals, KOut kout, VOut vout, Pred pred, Op op, trtc::CopyIfArgs a)
{
    typedef typename Keys::value_t K;
    constexpr int ITEMS = sizeof(K) >= sizeof(V) ? trtc::CopyIfItems<K>::value : trtc::CopyIfItems<V>::value;
    trtc::reduce_by_key_tile<TRTC_SCAN_BLOCK, ITEMS, V, K>(keys, vals, kout, vout, pred, op, a);
}

template <class Keys, class Vals, class Out, class Pred, class Op, class V>
__global__ void __launch_bounds__((WideBlock<typename Keys::value_t, V>::value))
trtc_scan_by_key_k(Keys keys, Vals vals, Out out, Pred pred, Op op, V init, trtc::ScanByKeyArgs a)
{
    typedef typename Keys::value_t K;
    constexpr int ITEMS = sizeof(K) >= sizeof(V) ? trtc::ScanByKeyItems<K>::value : trtc::ScanByKeyItems<V>::value;
    trtc::scan_by_key_tile<WideBlock<K, V>::value, ITEMS, V, K>(keys, vals, out, pred, op, init, a);
}

// Histogram pre-pass: 1024 threads, one workgroup per CU (up to 128 KB of
// LDS counters).
#define TRTC_HIST_BLOCK 1024
template <int KB>
__global__ void __launch_bounds__(TRTC_HIST_BLOCK) trtc_radix_hist_k(trtc::RadixHistArgs a)
{
    trtc::radix_hist<TRTC_HIST_BLOCK, KB>(a);
}

// Radix digit pass: BLOCK = 512 or 1024 threads; 16 keys per thread for keys
// and values of up to 4 bytes, 8 per thread when either is 8 bytes.  Large
// sorts use 1024 threads (16384-key tiles for 4-byte pairs, one workgroup per
// CU at 150 KB LDS): twice as long digit runs per tile, 2^28 u32/u32 pass
// 1.341 -> 1.291 ms against 512 threads; smaller sorts keep 512 threads for
// more tiles per CU.
template <int KB, int VB> struct RadixItems
{
    static constexpr int value = (KB > 4 || VB > 4) ? 8 : 16;
};

// KIND: the key kind (trtc::KeyKind) fixed at compile time, or -1 to read it
// from the arguments; the 1024-thread variants are compiled per kind (2^28
// u32/u32 pass 1.160 -> 1.144 ms), the 512-thread ones read it.
template <int KB, int VB, class S, int BLOCK, int KIND>
__global__ void __launch_bounds__(BLOCK) trtc_radix_pass_k(trtc::RadixPassArgs a)
{
    trtc::radix_pass<BLOCK, RadixItems<KB, VB>::value, KB, VB, S, KIND>(a);
}

// Long-tile digit pass (1024 x 32 keys, ranked in 32-lane units) for large
// sorts of 4-byte keys with 4-byte or no values (trtc::radix_pass_long).  Its
// workgroups take tiles until none is left; the grid must be a multiple of 8
// (tile runs go to virtual XCD blockIdx.x mod 8), normally one per CU.
#define TRTC_LONG_BLOCK 1024
#define TRTC_LONG_ITEMS 32
template <int VB, class S, int KIND>
__global__ void __launch_bounds__(TRTC_LONG_BLOCK) trtc_radix_pass_long_k(trtc::RadixPassArgs a)
{
    trtc::radix_pass_long<TRTC_LONG_BLOCK, TRTC_LONG_ITEMS, VB, S, KIND>(a);
}

// ---- batched segments (one launch over many independent sequences) ----
// Tile shapes are template arguments so the batch autotuner can pick among
// instantiations; every one keeps one thread per digit (BLOCK >= 256).  KIND
// fixes the key kind at compile time (trtc::KeyKind; -1 reads it).
template <int KB, int VB, int BLOCK, int ITEMS, int KIND>
__global__ void __launch_bounds__(BLOCK) trtc_seg_radix_pass_k(trtc::SegRadixArgs a)
{
    trtc::seg_radix_pass<BLOCK, ITEMS, KB, VB, KIND>(a);
}

// The long-tile pass over the large segments of a batch (1024 x 32 keys; grid
// a multiple of 8, as trtc_radix_pass_long_k).
template <int VB, int KIND>
__global__ void __launch_bounds__(TRTC_LONG_BLOCK) trtc_seg_radix_pass_long_k(trtc::SegRadixArgs a)
{
    trtc::seg_radix_pass_long<TRTC_LONG_BLOCK, TRTC_LONG_ITEMS, VB, KIND>(a);
}

template <int KB>
__global__ void __launch_bounds__(TRTC_HIST_BLOCK) trtc_seg_radix_hist_k(trtc::SegRadixArgs a)
{
    trtc::seg_radix_hist<TRTC_HIST_BLOCK, KB>(a);
}

template <class Op, class T, int ITEMS>
__global__ void __launch_bounds__(ScanBlock<T>::value) trtc_seg_scan_k(Op op, T init, trtc::SegScanArgs a)
{
    trtc::seg_scan_tile<ScanBlock<T>::value, ITEMS, T>(op, init, a);
}

// 16384 elements per reduce tile (4 waves x 4096, chunks of 64*ITEMS)
#define TRTC_SEG_REDUCE_TILE 16384
template <class Op, class T>
__global__ void __launch_bounds__(TRTC_BLOCK) trtc_seg_reduce_k(Op op, trtc::SegReduceArgs a)
{
    trtc::seg_reduce_tile<TRTC_BLOCK, trtc::ReduceItems<T>::value, TRTC_SEG_REDUCE_TILE, T>(op, a);
}

// ---- stable comparison sort (user comparators) ----
#define TRTC_MSORT_BLOCK 256
#define TRTC_MSORT_ITEMS 8
template <class T, class Comp>
__global__ void __launch_bounds__(TRTC_MSORT_BLOCK) trtc_msort_tile_k(T* keys, unsigned* idx, Comp comp,
                                                                     trtc::MergeSortArgs a)
{
    trtc::msort_tile<TRTC_MSORT_BLOCK, TRTC_MSORT_ITEMS, T>(keys, idx, comp, a);
}

template <class T, class Comp>
__global__ void __launch_bounds__(TRTC_MSORT_BLOCK)
trtc_msort_merge_k(const T* kin, const unsigned* iin, T* kout, unsigned* iout, Comp comp, trtc::MergeSortArgs a)
{
    trtc::msort_merge<TRTC_MSORT_BLOCK, TRTC_MSORT_ITEMS, T>(kin, iin, kout, iout, comp, a);
}

template <class T, class Comp>
__global__ void __launch_bounds__(256) trtc_msort_split_k(const T* kin, Comp comp, trtc::MergeSortArgs a)
{
    trtc::msort_splits<TRTC_MSORT_BLOCK * TRTC_MSORT_ITEMS, T>(kin, comp, a);
}

// Merge / Merge_By_Key of two sorted inputs (views of any kind), one tile of
// BLOCK * ITEMS outputs per workgroup.
template <class AK, class BK, class OK, class Comp, int BLOCK, int ITEMS>
__global__ void __launch_bounds__(BLOCK) trtc_merge_k(AK ak, BK bk, OK ok, Comp comp, trtc::MergeArgs a)
{
    typedef typename AK::value_t K;
    trtc::merge_tile<BLOCK, ITEMS, false, K>(ak, bk, ak, bk, ok, ok, comp, a);
}

template <class AK, class BK, class Comp, int TILE>
__global__ void __launch_bounds__(256) trtc_merge_split_k(AK ak, BK bk, Comp comp, trtc::MergeArgs a,
                                                          unsigned long long* splits)
{
    trtc::merge_splits<TILE>(ak, bk, comp, a, splits);
}

template <class AK, class BK, class AV, class BV, class OK, class OV, class Comp, int BLOCK, int ITEMS>
__global__ void __launch_bounds__(BLOCK)
trtc_merge_by_key_k(AK ak, BK bk, AV av, BV bv, OK ok, OV ov, Comp comp, trtc::MergeArgs a)
{
    typedef typename AK::value_t K;
    trtc::merge_tile<BLOCK, ITEMS, true, K>(ak, bk, av, bv, ok, ov, comp, a);
}
