// Single-pass stable compaction (Copy_If / Copy_If_Stencil) for gfx950.
//
// Semantics follow copy.cpp:16-28 and general_copy_if.cpp:26-42: out receives,
// in input order, every in[i] whose predicate holds (pred(in[i]), or
// pred(stencil[i]) for the stencil form); elements of out past the count are
// left untouched; the count is returned to the host.
//
// One workgroup = one tile: blocked loads through a wave-local LDS transpose,
// predicate flags, a thread/wave/workgroup count, a decoupled lookback over
// 64-bit tile counts, then the kept elements are packed in LDS and written as
// one contiguous, coalesced run at the tile's output offset.
#pragma once
#include "trtc_prims.hpp"

namespace trtc {

template <int BLOCK, int ITEMS, class T, class In, class Sten, class Out, class Pred>
TRTC_DEV void copy_if_tile(In& in, Sten& sten, Out& out, Pred& pred, const CopyIfArgs& a)
{
    constexpr int WAVES = BLOCK / 64;
    constexpr int WELEMS = 64 * ITEMS;
    constexpr size_t TILE = (size_t)BLOCK * ITEMS;
    typedef typename Sten::value_t ST;
    constexpr bool HAS_ST = !_IsNoStencil<Sten>::value;
    typedef unsigned long long C;
    constexpr int ESZ = sizeof(T) > sizeof(ST) ? sizeof(T) : sizeof(ST);
    // one transpose buffer per wave, used for the input, then the stencil
    __shared__ __attribute__((aligned(16))) unsigned char s_raw[WAVES][WaveTile<ITEMS>::padded * ESZ];
    __shared__ C s_wtot[WAVES];
    __shared__ C s_prefix;
    __shared__ unsigned s_tile;

    if (threadIdx.x == 0) s_tile = atomicAdd(a.lb.tile_counter, 1u);
    __syncthreads();
    const size_t tile = s_tile;
    const unsigned wave = threadIdx.x / 64;
    const unsigned lane = lane_id();
    const size_t n = a.n;
    const size_t tile_base = tile * TILE;
    const size_t wave_base = tile_base + (size_t)wave * WELEMS;
    const bool full = tile_base + TILE <= n;
    T* lds = (T*)s_raw[wave];
    ST* ldss = (ST*)s_raw[wave];

    // contiguous views of 4/8-byte elements: 16-byte pieces
    T x[ITEMS];
    int nvalid = ITEMS;
    const bool direct_in = load_blocked_vec<ITEMS>(in, wave_base, full, lds, x);
    if (!direct_in) {
#pragma unroll
        for (int k = 0; k < ITEMS; k++) {
            int i = k * 64 + (int)lane;
            const size_t g = wave_base + i;  // past the end: a clamped index, so the loads issue together
            const T v = (T)in[full || g < n ? g : n - 1];
            if (full || g < n) lds[WaveTile<ITEMS>::pad(i)] = v;
        }
        wave_sync();
        if (!full) {
            size_t first = wave_base + (size_t)lane * ITEMS;
            nvalid = first >= n ? 0 : (int)((n - first) < (size_t)ITEMS ? (n - first) : ITEMS);
        }
#pragma unroll
        for (int j = 0; j < ITEMS; j++)
            if (j < nvalid) x[j] = lds[WaveTile<ITEMS>::pad((int)lane * ITEMS + j)];
    }
    unsigned flags = 0;
    unsigned cnt = 0;
    if constexpr (HAS_ST) {
        ST sv[ITEMS];
        const bool direct_st = load_blocked_vec<ITEMS>(sten, wave_base, full, ldss, sv);
        if (!direct_st) {
            wave_sync();
#pragma unroll
            for (int k = 0; k < ITEMS; k++) {
                int i = k * 64 + (int)lane;
                const size_t g = wave_base + i;
                const ST v = sten[full || g < n ? g : n - 1];
                if (full || g < n) ldss[WaveTile<ITEMS>::pad(i)] = v;
            }
            wave_sync();
#pragma unroll
            for (int j = 0; j < ITEMS; j++)
                if (j < nvalid) sv[j] = ldss[WaveTile<ITEMS>::pad((int)lane * ITEMS + j)];
        }
#pragma unroll
        for (int j = 0; j < ITEMS; j++) {
            if (j < nvalid) {
                bool f = (bool)pred(sv[j]);
                flags |= (unsigned)f << j;
                cnt += f;
            }
        }
    } else {
#pragma unroll
        for (int j = 0; j < ITEMS; j++) {
            if (j < nvalid) {
                bool f = (bool)pred(x[j]);
                flags |= (unsigned)f << j;
                cnt += f;
            }
        }
    }
    Plus plus;
    C winc = wave_inclusive_scan((C)cnt, plus);
    if (lane == 63) s_wtot[wave] = winc;
    C lexcl = winc - cnt;
    __syncthreads();
    C wpre = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < WAVES; w++) {
        C t = s_wtot[w];
        if ((unsigned)w < wave) wpre += t;
        tot += t;
    }
    C P = 0;
    if (tile == 0) {
        if (threadIdx.x == 0) publish<C>(a.lb.incl, 0, tot);
    } else {
        if (threadIdx.x == 0) publish<C>(a.lb.agg, tile, tot);
        if (wave == 0) {
            C pre = lookback<C>(a.lb, tile, plus);
            if (lane == 0) {
                publish<C>(a.lb.incl, tile, pre + tot);
                s_prefix = pre;
            }
        }
        __syncthreads();
        P = s_prefix;
    }
    if (tile == a.num_tiles - 1 && threadIdx.x == 0) *a.count = P + tot;

    // pack the kept elements of the tile into LDS (tile-local order), then
    // write them out as one contiguous run
    __syncthreads();  // every wave is done reading its transpose buffer
    T* pack = (T*)&s_raw[0][0];
    unsigned local = (unsigned)(wpre + lexcl);
#pragma unroll
    for (int j = 0; j < ITEMS; j++) {
        if ((flags >> j) & 1u) pack[local++] = x[j];
    }
    __syncthreads();
    typedef typename Out::value_t OutT;
    const size_t cap = out.size();  // never write past the output vector
    for (unsigned i = threadIdx.x; i < (unsigned)tot; i += BLOCK)
        if (P + i < cap) out[P + i] = (OutT)pack[i];
}

// Items per thread of the compaction tiles (copy_if, unique, partition,
// reduce_by_key; 512 threads): 128 bytes per lane for 4- and 8-byte elements.
// The host's tile size follows the same table.
template <class T> struct CopyIfItems
{
    static constexpr int value = sizeof(T) <= 4 ? 32 : (sizeof(T) <= 8 ? 16 : 4);
};

}  // namespace trtc
