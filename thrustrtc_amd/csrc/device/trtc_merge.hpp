// Merge / Merge_By_Key of two sorted inputs (merge.cpp): a stable merge in
// which equal keys take the first input first (std::merge).
//
// One workgroup produces one tile of BLOCK*ITEMS outputs: it finds where the
// tile's first and last diagonals cut the two inputs (merge-path binary
// search over the views, ties to the first input), stages the two input
// ranges in LDS, and every thread merges its ITEMS outputs from its own
// split, writing keys (and the values of their source elements).
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

template <int BLOCK, int ITEMS, bool VALS, class K, class AK, class BK, class AV, class BV, class OK, class OV,
          class Comp>
TRTC_DEV void merge_tile(AK& ak, BK& bk, AV& av, BV& bv, OK& ok, OV& ov, Comp& comp, const MergeArgs& a)
{
    constexpr int TILE = BLOCK * ITEMS;
    __shared__ K s_k[TILE];
    __shared__ size_t s_split[2];
    const size_t la = a.na, lb = a.nb, n = la + lb;
    const size_t out0 = (size_t)blockIdx.x * TILE;
    const size_t d1 = out0 + TILE < n ? out0 + TILE : n;
    if (threadIdx.x < 2) s_split[threadIdx.x] = merge_split_views(ak, la, bk, lb, threadIdx.x ? d1 : out0, comp);
    __syncthreads();
    const size_t a0 = s_split[0], a1 = s_split[1];
    const size_t b0 = out0 - a0, b1 = d1 - a1;
    const unsigned na = (unsigned)(a1 - a0), nb = (unsigned)(b1 - b0);
    for (unsigned p = threadIdx.x; p < na + nb; p += BLOCK) s_k[p] = p < na ? (K)ak[a0 + p] : (K)bk[b0 + p - na];
    __syncthreads();
    const unsigned total = na + nb;
    const unsigned q0 = threadIdx.x * ITEMS;
    if (q0 >= total) return;
    unsigned i = (unsigned)merge_split(s_k, na, s_k + na, nb, q0, comp);
    unsigned j = q0 - i;
    const unsigned cnt = total - q0 < (unsigned)ITEMS ? total - q0 : (unsigned)ITEMS;
    typedef typename OK::value_t KO;
#pragma unroll
    for (int e = 0; e < ITEMS; e++) {
        if ((unsigned)e >= cnt) break;
        bool takeA;
        if (i >= na) takeA = false;
        else if (j >= nb) takeA = true;
        else takeA = !comp(s_k[na + j], s_k[i]);
        const size_t p = out0 + q0 + e;
        if (takeA) {
            ok[p] = (KO)s_k[i];
            if constexpr (VALS) ov[p] = (typename OV::value_t)av[a0 + i];
            i++;
        } else {
            ok[p] = (KO)s_k[na + j];
            if constexpr (VALS) ov[p] = (typename OV::value_t)bv[b0 + j];
            j++;
        }
    }
}

}  // namespace trtc
