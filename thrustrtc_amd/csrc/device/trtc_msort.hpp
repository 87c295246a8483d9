// Stable comparison sort for user comparators (the reference's merge sort,
// merge_sort.cpp: ties keep input order, i.e. std::stable_sort with comp).
//
// 1. Tile sort: a tile of BLOCK*ITEMS keys (with their input indices) is
//    sorted in LDS: every thread sorts its ITEMS consecutive keys (odd-even
//    transposition, stable), then runs are merged pairwise up to the tile
//    size by merge path (each thread: one diagonal binary search, then ITEMS
//    sequential merge steps; equal keys take the left run first).
// 2. Merge passes: runs of width w are merged pairwise into runs of 2w.  Each
//    workgroup produces one output tile: it finds where the tile's first and
//    last output diagonals cut the two input runs (merge-path binary search,
//    ties to the left run), stages both input ranges in LDS, and every
//    thread merges its ITEMS outputs from its own split.
#pragma once
#include "trtc_prims.hpp"

namespace trtc {

// Merge-path split of diagonal d of runs A (lenA) and B (lenB): the number of
// the first d outputs taken from A (equal keys: A first).
template <class T, class Comp>
TRTC_DEV size_t merge_split(const T* A, size_t lenA, const T* B, size_t lenB, size_t d, Comp& comp)
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

// LDS position of staged item p: one pad word per 32, so that the ITEMS
// consecutive items of a thread (stride ITEMS between lanes) fall in distinct
// banks when every thread reads or writes its e-th item at once.
TRTC_DEV unsigned lds_pad(unsigned p) { return p + (p >> 5); }
template <int TILE> struct LdsPadded
{
    static constexpr int n = TILE + TILE / 32;
};

// merge_split over padded LDS: A = s[pad(a + i)], i < la; B = s[pad(b + j)], j < lb.
template <class T, class Comp>
TRTC_DEV unsigned lds_split(const T* s, unsigned a, unsigned la, unsigned b, unsigned lb, unsigned d, Comp& comp)
{
    unsigned lo = d > lb ? d - lb : 0;
    unsigned hi = d < la ? d : la;
    while (lo < hi) {
        const unsigned mid = (lo + hi) >> 1;
        if (!comp(s[lds_pad(b + d - mid - 1)], s[lds_pad(a + mid)])) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

template <int BLOCK, int ITEMS, class T, class Comp>
TRTC_DEV void msort_tile(T* keys, unsigned* idx, Comp& comp, const MergeSortArgs& a)
{
    constexpr int TILE = BLOCK * ITEMS;
    __shared__ T s_k[LdsPadded<TILE>::n];
    __shared__ unsigned s_i[LdsPadded<TILE>::n];
    const size_t base = (size_t)blockIdx.x * TILE;
    const unsigned valid = (unsigned)((a.n - base) < (size_t)TILE ? (a.n - base) : (size_t)TILE);

    for (unsigned p = threadIdx.x; p < valid; p += BLOCK) {
        s_k[lds_pad(p)] = keys[base + p];
        s_i[lds_pad(p)] = a.with_idx ? idx[base + p] : 0u;
    }
    __syncthreads();
    // 1. each thread: stable sort of its ITEMS consecutive keys (odd-even
    //    transposition: adjacent swaps of strict inversions only, constant
    //    register indices)
    T k[ITEMS];
    unsigned ix[ITEMS];
    const unsigned p0 = threadIdx.x * ITEMS;
    const unsigned mine = p0 >= valid ? 0 : (valid - p0 < (unsigned)ITEMS ? valid - p0 : (unsigned)ITEMS);
#pragma unroll
    for (int e = 0; e < ITEMS; e++)
        if ((unsigned)e < mine) {
            k[e] = s_k[lds_pad(p0 + e)];
            ix[e] = s_i[lds_pad(p0 + e)];
        }
#pragma unroll
    for (int r = 0; r < ITEMS; r++) {
#pragma unroll
        for (int e = r & 1; e + 1 < ITEMS; e += 2) {
            if ((unsigned)(e + 1) < mine && comp(k[e + 1], k[e])) {
                T t = k[e];
                k[e] = k[e + 1];
                k[e + 1] = t;
                unsigned ti = ix[e];
                ix[e] = ix[e + 1];
                ix[e + 1] = ti;
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int e = 0; e < ITEMS; e++)
        if ((unsigned)e < mine) {
            s_k[lds_pad(p0 + e)] = k[e];
            s_i[lds_pad(p0 + e)] = ix[e];
        }
    __syncthreads();
    // 2. pairwise run merges inside the tile, by merge path: the thread that
    //    owns outputs [p0, p0 + ITEMS) of a run pair finds where its first
    //    diagonal cuts the two runs (one binary search, ties to the left run)
    //    and merges its ITEMS outputs sequentially
    for (unsigned run = ITEMS; run < (unsigned)TILE; run <<= 1) {
        unsigned cnt = 0;
        if (p0 < valid) {
            const unsigned g = p0 / (2 * run) * (2 * run);  // start of the run pair
            const unsigned la = (g + run <= valid ? run : valid - g);
            const unsigned lb = g + run < valid ? ((g + 2 * run <= valid ? 2 * run : valid - g) - run) : 0;
            const unsigned d = p0 - g;
            cnt = la + lb - d < (unsigned)ITEMS ? la + lb - d : (unsigned)ITEMS;
            unsigned i = lds_split(s_k, g, la, g + la, lb, d, comp);
            unsigned j = d - i;
#pragma unroll
            for (int e = 0; e < ITEMS; e++) {
                if ((unsigned)e >= cnt) break;
                bool takeA;
                if (i >= la) takeA = false;
                else if (j >= lb) takeA = true;
                else takeA = !comp(s_k[lds_pad(g + la + j)], s_k[lds_pad(g + i)]);
                const unsigned q = lds_pad(takeA ? g + i : g + la + j);
                k[e] = s_k[q];
                ix[e] = s_i[q];
                if (takeA) i++;
                else j++;
            }
        }
        __syncthreads();
#pragma unroll
        for (int e = 0; e < ITEMS; e++)
            if ((unsigned)e < cnt) {
                s_k[lds_pad(p0 + e)] = k[e];
                s_i[lds_pad(p0 + e)] = ix[e];
            }
        __syncthreads();
    }
    for (unsigned p = threadIdx.x; p < valid; p += BLOCK) {
        keys[base + p] = s_k[lds_pad(p)];
        if (a.with_idx) idx[base + p] = s_i[lds_pad(p)];
    }
}

// The run pair of output position out0 in a merge pass over runs of width w:
// pair start, lengths of its left (A) and right (B) runs.
TRTC_DEV void msort_pair(size_t out0, size_t n, size_t w, size_t& pair, size_t& lenA, size_t& lenB)
{
    pair = out0 / (2 * w) * (2 * w);
    lenA = pair + w < n ? w : n - pair;
    lenB = pair + w < n ? ((pair + 2 * w <= n ? 2 * w : n - pair) - w) : 0;
}

// Splits of every merge tile's first diagonal, one thread per tile: all the
// binary searches of a pass run at once instead of heading each workgroup.
template <int TILE, class T, class Comp>
TRTC_DEV void msort_splits(const T* kin, Comp& comp, const MergeSortArgs& a)
{
    const size_t b = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t tiles = (a.n + TILE - 1) / TILE;
    if (b >= tiles) return;
    size_t pair, lenA, lenB;
    msort_pair(b * TILE, a.n, a.width, pair, lenA, lenB);
    a.splits[b] = merge_split(kin + pair, lenA, kin + pair + lenA, lenB, b * TILE - pair, comp);
}

template <int BLOCK, int ITEMS, class T, class Comp>
TRTC_DEV void msort_merge(const T* kin, const unsigned* iin, T* kout, unsigned* iout, Comp& comp,
                          const MergeSortArgs& a)
{
    constexpr int TILE = BLOCK * ITEMS;
    __shared__ T s_k[LdsPadded<TILE>::n];
    __shared__ unsigned s_i[LdsPadded<TILE>::n];
    __shared__ size_t s_split[2];
    const size_t out0 = (size_t)blockIdx.x * TILE;
    size_t pair, lenA, lenB;
    msort_pair(out0, a.n, a.width, pair, lenA, lenB);
    const T* A = kin + pair;
    const T* B = kin + pair + lenA;
    const size_t d0 = out0 - pair;
    const size_t d1 = (d0 + TILE < lenA + lenB) ? d0 + TILE : lenA + lenB;
    if (threadIdx.x < 2) {
        size_t sp;
        if (!a.splits)
            sp = merge_split(A, lenA, B, lenB, threadIdx.x ? d1 : d0, comp);
        else if (threadIdx.x == 0)
            sp = a.splits[blockIdx.x];
        else  // the next tile starts at d1 in the same pair, or the pair ends at d1
            sp = d1 == lenA + lenB ? lenA : a.splits[blockIdx.x + 1];
        s_split[threadIdx.x] = sp;
    }
    __syncthreads();
    const size_t a0 = s_split[0], a1 = s_split[1];
    const size_t b0 = d0 - a0, b1 = d1 - a1;
    const unsigned na = (unsigned)(a1 - a0), nb = (unsigned)(b1 - b0);
    for (unsigned p = threadIdx.x; p < na + nb; p += BLOCK) {
        if (p < na) {
            s_k[lds_pad(p)] = A[a0 + p];
            s_i[lds_pad(p)] = a.with_idx ? iin[pair + a0 + p] : 0u;
        } else {
            s_k[lds_pad(p)] = B[b0 + p - na];
            s_i[lds_pad(p)] = a.with_idx ? iin[pair + lenA + b0 + p - na] : 0u;
        }
    }
    __syncthreads();
    // this thread's outputs: diagonals [t*ITEMS, t*ITEMS + ITEMS) of the staged ranges
    const unsigned total = na + nb;
    const unsigned q0 = threadIdx.x * ITEMS;
    T k[ITEMS];
    unsigned ix[ITEMS];
    unsigned cnt = 0;
    if (q0 < total) {
        unsigned i = lds_split(s_k, 0, na, na, nb, q0, comp);
        unsigned j = q0 - i;
        cnt = total - q0 < (unsigned)ITEMS ? total - q0 : (unsigned)ITEMS;
#pragma unroll
        for (int e = 0; e < ITEMS; e++) {
            if ((unsigned)e >= cnt) break;
            bool takeA;
            if (i >= na) takeA = false;
            else if (j >= nb) takeA = true;
            else takeA = !comp(s_k[lds_pad(na + j)], s_k[lds_pad(i)]);
            const unsigned q = lds_pad(takeA ? i : na + j);
            k[e] = s_k[q];
            ix[e] = s_i[q];
            if (takeA) i++;
            else j++;
        }
    }
    __syncthreads();
#pragma unroll
    for (int e = 0; e < ITEMS; e++)
        if ((unsigned)e < cnt) {
            s_k[lds_pad(q0 + e)] = k[e];
            s_i[lds_pad(q0 + e)] = ix[e];
        }
    __syncthreads();
    for (unsigned p = threadIdx.x; p < total; p += BLOCK) {
        kout[out0 + p] = s_k[lds_pad(p)];
        if (a.with_idx) iout[out0 + p] = s_i[lds_pad(p)];
    }
}

}  // namespace trtc
