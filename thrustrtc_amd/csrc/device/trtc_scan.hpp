// Single-pass inclusive / exclusive scan (decoupled lookback) for gfx950.
//
// Semantics follow scan.cpp:4-39 and general_scan.cpp:8-88 of the reference:
//   inclusive  out[i] = in[0] op in[1] op ... op in[i]
//   exclusive  out[0] = init, out[i] = init op in[0] op ... op in[i-1]
// The scan runs in the output element type T (each input is converted to T
// first), and the left operand of op is always the earlier part of the
// sequence, so any associative op works, commutative or not.
//
// Work decomposition: one workgroup = one tile of BLOCK*ITEMS elements, tile
// ids handed out in order by an atomic counter.  Each wave owns 64*ITEMS
// consecutive elements: coalesced striped loads into LDS, a wave-local
// transpose to a blocked layout (lane l holds ITEMS consecutive elements), a
// sequential thread scan, a wave scan, a workgroup combine, one lookback by
// wave 0, then the reverse transpose and coalesced striped stores.  A tile
// reads all of its inputs before it writes any output and writes no other
// tile's elements, so in-place scans (in == out) are safe.
#pragma once
#include "trtc_prims.hpp"

namespace trtc {

template <class T> struct MaybeT
{
    T v;
    bool has;
};

// One tile of a sequence: lookback index `tile` (status slot), `first` = the
// lookback index of the sequence's first tile, elements [base, min(base+TILE,
// n)) of src/out.  The loads may read up to `readable` (>= n): a tile that
// ends early inside its sequence still loads its whole extent through
// 16-byte pieces and drops the elements past n.
template <int BLOCK, int ITEMS, class T, class Src, class Out, class Op>
TRTC_DEV void scan_tile_at(Src& src, Out& out, Op& op, const T& init, int exclusive, const LookbackArgs& lb,
                           size_t tile, size_t first, size_t tile_base, size_t n, size_t readable)
{
    constexpr int WAVES = BLOCK / 64;
    constexpr int WELEMS = 64 * ITEMS;
    constexpr size_t TILE = (size_t)BLOCK * ITEMS;
    __shared__ T s_x[WAVES][WaveTile<ITEMS>::padded];
    __shared__ T s_wtot[WAVES];
    __shared__ int s_wvalid[WAVES];
    __shared__ T s_prefix;

    const unsigned wave = threadIdx.x / 64;
    const unsigned lane = lane_id();
    const size_t wave_base = tile_base + (size_t)wave * WELEMS;
    const bool full = tile_base + TILE <= n;
    T* lds = s_x[wave];

    // contiguous views of 4/8-byte elements: 16-byte pieces
    T x[ITEMS];
    int nvalid = ITEMS;
    if (!full) {
        size_t first = wave_base + (size_t)lane * ITEMS;
        nvalid = first >= n ? 0 : (int)((n - first) < (size_t)ITEMS ? (n - first) : ITEMS);
    }
    const bool direct = load_blocked_vec<ITEMS>(src, wave_base, tile_base + TILE <= readable, lds, x);
    if (!direct) {
    // striped load -> LDS
    if (full) {
#pragma unroll
        for (int k = 0; k < ITEMS; k++) {
            int i = k * 64 + (int)lane;
            lds[WaveTile<ITEMS>::pad(i)] = (T)src[wave_base + i];
        }
    } else {
#pragma unroll
        for (int k = 0; k < ITEMS; k++) {
            int i = k * 64 + (int)lane;
            // past the end a clamped index, so that the loads issue together
            // (a load inside the bounds test waits before the next one)
            const size_t g = wave_base + i;
            const T v = (T)src[g < n ? g : n - 1];
            if (g < n) lds[WaveTile<ITEMS>::pad(i)] = v;
        }
    }
    wave_sync();

    // blocked read: lane holds elements lane*ITEMS .. lane*ITEMS+ITEMS-1
#pragma unroll
    for (int j = 0; j < ITEMS; j++)
        if (j < nvalid) x[j] = lds[WaveTile<ITEMS>::pad((int)lane * ITEMS + j)];
    }

    // thread aggregate
    T tagg = x[0];
#pragma unroll
    for (int j = 1; j < ITEMS; j++)
        if (j < nvalid) tagg = (T)op(tagg, x[j]);
    const bool tvalid = nvalid > 0;

    // wave inclusive scan of thread aggregates
    T winc = full ? wave_inclusive_scan(tagg, op) : wave_inclusive_scan_partial(tagg, tvalid, op);
    unsigned long long vmask = ballot(tvalid);
    const bool wvalid = vmask != 0;
    const unsigned wlast = wvalid ? 63 - (unsigned)__builtin_clzll(vmask) : 0;
    if (lane == wlast) {
        s_wtot[wave] = winc;
        s_wvalid[wave] = wvalid ? 1 : 0;
    }
    // lane-exclusive prefix inside the wave
    T wexcl = shfl_up(winc, 1);
    __syncthreads();

    // prefix of earlier waves and the tile aggregate
    MaybeT<T> wpre;
    wpre.has = false;
    MaybeT<T> tot;
    tot.has = false;
#pragma unroll
    for (int w = 0; w < WAVES; w++) {
        if (!s_wvalid[w]) continue;
        T t = s_wtot[w];
        if ((unsigned)w < wave) {
            wpre.v = wpre.has ? (T)op(wpre.v, t) : t;
            wpre.has = true;
        }
        tot.v = tot.has ? (T)op(tot.v, t) : t;
        tot.has = true;
    }

    // tile prefix P: the sequence's first tile has none (inclusive) or init
    // (exclusive); a sequence of one tile publishes nothing (lb.incl null)
    MaybeT<T> P;
    P.has = false;
    if (tile == first) {
        if (exclusive) {
            P.v = init;
            P.has = true;
            if (threadIdx.x == 0 && lb.incl) publish<T>(lb.incl, tile, (T)op(init, tot.v));
        } else if (threadIdx.x == 0 && lb.incl) {
            publish<T>(lb.incl, tile, tot.v);
        }
    } else {
        if (threadIdx.x == 0) publish<T>(lb.agg, tile, tot.v);
        if (wave == 0) {
            T pre = lookback<T>(lb, tile, op, first);
            if (lane == 0) {
                publish<T>(lb.incl, tile, (T)op(pre, tot.v));
                s_prefix = pre;
            }
        }
        __syncthreads();
        P.v = s_prefix;
        P.has = true;
    }

    // running prefix for this thread: P op (earlier waves) op (earlier lanes)
    MaybeT<T> run = P;
    if (wpre.has) {
        run.v = run.has ? (T)op(run.v, wpre.v) : wpre.v;
        run.has = true;
    }
    if (lane > 0 && tvalid) {
        run.v = run.has ? (T)op(run.v, wexcl) : wexcl;
        run.has = true;
    }

    T y[ITEMS];
    if (exclusive) {
        // run.has is always true here (P exists in exclusive mode)
#pragma unroll
        for (int j = 0; j < ITEMS; j++) {
            if (j < nvalid) {
                y[j] = run.v;
                run.v = (T)op(run.v, x[j]);
            }
        }
    } else {
#pragma unroll
        for (int j = 0; j < ITEMS; j++) {
            if (j < nvalid) {
                run.v = run.has ? (T)op(run.v, x[j]) : x[j];
                run.has = true;
                y[j] = run.v;
            }
        }
    }

    // blocked -> LDS -> striped stores (full cache lines per store instruction)
    wave_sync();
#pragma unroll
    for (int j = 0; j < ITEMS; j++)
        if (j < nvalid) lds[WaveTile<ITEMS>::pad((int)lane * ITEMS + j)] = y[j];
    wave_sync();
    typedef typename Out::value_t OutT;
    if (full) {
#pragma unroll
        for (int k = 0; k < ITEMS; k++) {
            int i = k * 64 + (int)lane;
            out[wave_base + i] = (OutT)lds[WaveTile<ITEMS>::pad(i)];
        }
    } else {
#pragma unroll
        for (int k = 0; k < ITEMS; k++) {
            int i = k * 64 + (int)lane;
            if (wave_base + i < n) out[wave_base + i] = (OutT)lds[WaveTile<ITEMS>::pad(i)];
        }
    }
}

// One tile of a whole-input scan (tile ids from the atomic counter).
template <int BLOCK, int ITEMS, class T, class Src, class Out, class Op>
TRTC_DEV void scan_tile(Src& src, Out& out, Op& op, const T& init, const ScanArgs& a)
{
    __shared__ unsigned s_tile;
    if (threadIdx.x == 0) s_tile = atomicAdd(a.lb.tile_counter, 1u);
    __syncthreads();
    const size_t tile = s_tile;
    scan_tile_at<BLOCK, ITEMS, T>(src, out, op, init, a.exclusive, a.lb, tile, 0, tile * (size_t)BLOCK * ITEMS, a.n,
                                  a.n);
}

// One tile of a batch of independent scans: segments of more than one tile
// use the launch-wide lookback arrays (slots = launch tile ids), each lookback
// stopping at its segment's first tile.
template <int BLOCK, int ITEMS, class T, class Op>
TRTC_DEV void seg_scan_tile(Op& op, const T& init, const SegScanArgs& a)
{
    constexpr size_t TILE = (size_t)BLOCK * ITEMS;
    __shared__ unsigned s_tile;
    if (threadIdx.x == 0) s_tile = atomicAdd(a.lb.tile_counter, 1u);
    __syncthreads();
    const size_t tile = s_tile;
    const SegScanDesc& d = a.descs[a.tile_seg[tile]];
    PtrView<const T> src{(const T*)d.in};
    PtrView<T> out{(T*)d.out};
    LookbackArgs lb = a.lb;
    if (d.n <= TILE) lb.incl = nullptr;
    // tile 0 holds TILE - lead elements, so later tiles start on 128-byte lines
    const size_t lt = tile - d.first_tile;
    const size_t base = lt ? lt * TILE - d.lead : 0;
    const size_t end = (lt + 1) * TILE - d.lead < (size_t)d.n ? (lt + 1) * TILE - d.lead : (size_t)d.n;
    scan_tile_at<BLOCK, ITEMS, T>(src, out, op, init, a.exclusive, lb, tile, (size_t)d.first_tile, base, end,
                                  (size_t)d.n);
}

// Items per thread of the whole-input scan (512 threads): 128 bytes per lane
// for 4- and 8-byte elements.  On 2^29 int32, 32 items run 1.18 -> 0.88 ms
// against 16 (2 workgroups/CU at 89 VGPRs and 67 KB LDS); the host's tile
// size follows the same table.
template <class T> struct ScanItems
{
    static constexpr int value = sizeof(T) <= 4 ? 32 : (sizeof(T) <= 8 ? 16 : 4);
};

}  // namespace trtc
