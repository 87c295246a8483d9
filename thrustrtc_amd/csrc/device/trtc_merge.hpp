// Merge / Merge_By_Key of two sorted inputs (merge.cpp): a stable merge in
// which equal keys take the first input first (std::merge).
//
// One workgroup produces one tile of BLOCK*ITEMS outputs: the merge-path
// splits of all tile starts come from a split launch (one thread per tile
// boundary: the binary searches run in parallel instead of two dependent
// chains of global reads per workgroup; ties go to the first input); the
// tile stages its two input ranges (keys, and values for Merge_By_Key) in
// LDS, every thread merges its ITEMS outputs from its own split, and the
// merged keys (and their source positions) go back through LDS so that keys
// and values are stored by consecutive threads at consecutive outputs.
#pragma once
#include "trtc_msort.hpp"

namespace trtc {

// merge_split over views (any indexable with operator[]).
template <class AV, class BV, class Comp>
TRTC_DEV size_t merge_split_views(AV& A, size_t lenA, BV& B, size_t lenB, size_t d, Comp& comp)
{
    size_t lo = d > lenB ? d - lenB : 0;
    size_t hi = d < lenA ? d : lenA;
    while (lo < hi) {
        size_t mid = (lo + hi) >> 1;
        if (!comp(B[d - mid - 1], A[mid])) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// Splits of the tile starts b*TILE, b = 0 .. tiles (the last one is n).
template <int TILE, class AK, class BK, class Comp>
TRTC_DEV void merge_splits(AK& ak, BK& bk, Comp& comp, const MergeArgs& a, unsigned long long* splits)
{
    const size_t b = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t n = a.na + a.nb;
    const size_t tiles = (n + TILE - 1) / TILE;
    if (b > tiles) return;
    const size_t d = b * TILE < n ? b * TILE : n;
    splits[b] = merge_split_views(ak, a.na, bk, a.nb, d, comp);
}

template <int BLOCK, int ITEMS, bool VALS, class K, class AK, class BK, class AV, class BV, class OK, class OV,
          class Comp>
TRTC_DEV void merge_tile(AK& ak, BK& bk, AV& av, BV& bv, OK& ok, OV& ov, Comp& comp, const MergeArgs& a)
{
    constexpr int TILE = BLOCK * ITEMS;
    constexpr int PADDED = LdsPadded<TILE>::n;
    typedef typename OV::value_t VO;
    // values of up to 8 bytes are staged in LDS beside the keys; wider ones
    // are gathered from the inputs by source position
    constexpr bool STAGE_V = VALS && sizeof(VO) <= 8;
    __shared__ K s_k[PADDED];
    __shared__ unsigned s_src[VALS ? PADDED : 1];
    __shared__ VO s_v[STAGE_V ? PADDED : 1];
    __shared__ size_t s_split[2];
    const size_t la = a.na, lb = a.nb, n = la + lb;
    const size_t out0 = (size_t)blockIdx.x * TILE;
    const size_t d1 = out0 + TILE < n ? out0 + TILE : n;
    if (threadIdx.x < 2) {
        s_split[threadIdx.x] = a.splits ? (size_t)a.splits[blockIdx.x + threadIdx.x]
                                        : merge_split_views(ak, la, bk, lb, threadIdx.x ? d1 : out0, comp);
    }
    __syncthreads();
    const size_t a0 = s_split[0], a1 = s_split[1];
    const size_t b0 = out0 - a0, b1 = d1 - a1;
    const unsigned na = (unsigned)(a1 - a0), nb = (unsigned)(b1 - b0);
    const unsigned total = na + nb;
    for (unsigned p = threadIdx.x; p < total; p += BLOCK) {
        s_k[lds_pad(p)] = p < na ? (K)ak[a0 + p] : (K)bk[b0 + p - na];
        if constexpr (STAGE_V) s_v[lds_pad(p)] = p < na ? (VO)av[a0 + p] : (VO)bv[b0 + p - na];
    }
    __syncthreads();
    const unsigned q0 = threadIdx.x * ITEMS;
    K k[ITEMS];
    unsigned src[ITEMS];  // position in the staged ranges: < na from the first input
    unsigned cnt = 0;
    if (q0 < total) {
        unsigned i = lds_split(s_k, 0, na, na, nb, q0, comp);  // ties to the first input
        unsigned j = q0 - i;
        cnt = total - q0 < (unsigned)ITEMS ? total - q0 : (unsigned)ITEMS;
#pragma unroll
        for (int e = 0; e < ITEMS; e++) {
            if ((unsigned)e >= cnt) break;
            bool takeA;
            if (i >= na) takeA = false;
            else if (j >= nb) takeA = true;
            else takeA = !comp(s_k[lds_pad(na + j)], s_k[lds_pad(i)]);
            src[e] = takeA ? i : na + j;
            k[e] = s_k[lds_pad(src[e])];
            if (takeA) i++;
            else j++;
        }
    }
    __syncthreads();
#pragma unroll
    for (int e = 0; e < ITEMS; e++)
        if ((unsigned)e < cnt) {
            s_k[lds_pad(q0 + e)] = k[e];
            if constexpr (VALS) s_src[lds_pad(q0 + e)] = src[e];
        }
    __syncthreads();
    typedef typename OK::value_t KO;
    for (unsigned p = threadIdx.x; p < total; p += BLOCK) {
        ok[out0 + p] = (KO)s_k[lds_pad(p)];
        if constexpr (VALS) {
            const unsigned q = s_src[lds_pad(p)];
            if constexpr (STAGE_V) ov[out0 + p] = s_v[lds_pad(q)];
            else ov[out0 + p] = q < na ? (VO)av[a0 + q] : (VO)bv[b0 + q - na];
        }
    }
}

}  // namespace trtc
