// Single-pass selections on the compaction machinery of copy_if (decoupled
// lookback over 64-bit tile counts, kept items packed in LDS and written as
// one contiguous run per tile):
//
//   unique_tile        Unique / Unique_Copy / Unique_By_Key[_Copy]
//                      (unique.cpp:5-95): item i is kept when i == 0 or
//                      !pred(key[i], key[i-1]); by-key forms move the values
//                      of the kept keys along.
//   partition_tile     Partition[_Copy][_Stencil] (partition.cpp:28-130):
//                      items whose predicate holds go, in order, to the
//                      "true" output, the others, in order, to the "false"
//                      output at (index - trues before it).
//   reduce_by_key_tile Reduce_By_Key (reduce.cpp:36-62), fused: the segmented
//                      scan of Scan_By_Key (segments break where
//                      !pred(key[i-1], key[i])) and the compaction of segment
//                      tails (i == n-1 or !pred(key[i], key[i+1])) in one
//                      pass; the carried state is (flag, value, tail count).
//
// Every output position is at or before the index of the item written there,
// and a tile writes only after every earlier tile has published its count
// (which it does after loading its inputs), so the in-place forms (output ==
// input) are race-free without a copy of the input.
#pragma once
#include "trtc_prims.hpp"
#include "trtc_scan_by_key.hpp"

namespace trtc {

typedef unsigned long long SelC;

// Offsets of a selection in which this thread selected `cnt` items: `local` =
// its first slot in the tile, `tot` = the tile's count, `P` = the count of all
// earlier tiles (lookback).  Publishes the tile's count; the last tile stores
// the grand total in *a.count.
template <int BLOCK>
TRTC_DEV void select_offsets(unsigned cnt, size_t tile, const CopyIfArgs& a, SelC* s_wtot, SelC* s_prefix,
                             SelC& local, SelC& tot, SelC& P)
{
    constexpr int WAVES = BLOCK / 64;
    const unsigned wave = threadIdx.x / 64;
    const unsigned lane = lane_id();
    Plus plus;
    SelC winc = wave_inclusive_scan((SelC)cnt, plus);
    if (lane == 63) s_wtot[wave] = winc;
    __syncthreads();
    SelC wpre = 0;
    tot = 0;
#pragma unroll
    for (int w = 0; w < WAVES; w++) {
        SelC t = s_wtot[w];
        if ((unsigned)w < wave) wpre += t;
        tot += t;
    }
    local = wpre + winc - cnt;
    P = 0;
    if (tile == 0) {
        if (threadIdx.x == 0) publish<SelC>(a.lb.incl, 0, tot);
    } else {
        if (threadIdx.x == 0) publish<SelC>(a.lb.agg, tile, tot);
        if (wave == 0) {
            SelC pre = lookback<SelC>(a.lb, tile, plus);
            if (lane == 0) {
                publish<SelC>(a.lb.incl, tile, pre + tot);
                *s_prefix = pre;
            }
        }
        __syncthreads();
        P = *s_prefix;
    }
    if (tile == a.num_tiles - 1 && threadIdx.x == 0) *a.count = P + tot;
}

// Pack the flagged items of x into pack[local...] (tile order), then write the
// tile's `cnt` packed items to out[base...], never past out.size().
template <int BLOCK, int ITEMS, class T, class Out>
TRTC_DEV void write_packed(T* pack, const T (&x)[ITEMS], unsigned flags, SelC local, SelC cnt, Out& out, SelC base)
{
    __syncthreads();  // pack may alias buffers other waves were reading
    unsigned s = (unsigned)local;
#pragma unroll
    for (int j = 0; j < ITEMS; j++)
        if ((flags >> j) & 1u) pack[s++] = x[j];
    __syncthreads();
    typedef typename Out::value_t OutT;
    const size_t cap = out.size();
    for (unsigned i = threadIdx.x; i < (unsigned)cnt; i += BLOCK)
        if (base + i < cap) out[base + i] = (OutT)pack[i];
}

template <int BLOCK, int ITEMS, bool VALS, class K, class V, class KIn, class VIn, class KOut, class VOut, class Pred>
TRTC_DEV void unique_tile(KIn& kin, VIn& vin, KOut& kout, VOut& vout, Pred& pred, const CopyIfArgs& a)
{
    constexpr int WAVES = BLOCK / 64;
    constexpr int WELEMS = 64 * ITEMS;
    constexpr size_t TILE = (size_t)BLOCK * ITEMS;
    constexpr int ESZ = (VALS && sizeof(V) > sizeof(K)) ? sizeof(V) : sizeof(K);
    __shared__ __attribute__((aligned(16))) unsigned char s_raw[WAVES][WaveTile<ITEMS>::padded * ESZ];
    __shared__ SelC s_wtot[WAVES];
    __shared__ SelC s_prefix;
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
    const int nvalid = lane_valid<ITEMS>(wave_base, n, full);

    K x[ITEMS];
    load_blocked<ITEMS>(kin, wave_base, n, full, nvalid, (K*)s_raw[wave], x);
    // key before this lane's first item: the previous lane's last (full
    // whenever this lane has an item), or the item before the wave
    K prev = shfl_up(x[ITEMS - 1], 1);
    // (a wave that starts at or past n has no items and reads nothing)
    if (lane == 0 && wave_base > 0 && wave_base < n) prev = (K)kin[wave_base - 1];
    V v[VALS ? ITEMS : 1];
    if constexpr (VALS) load_blocked<ITEMS>(vin, wave_base, n, full, nvalid, (V*)s_raw[wave], v);

    const size_t g0 = wave_base + (size_t)lane * ITEMS;
    unsigned flags = 0, cnt = 0;
#pragma unroll
    for (int j = 0; j < ITEMS; j++) {
        if (j < nvalid) {
            bool keep = (g0 + j == 0) || !(bool)pred(x[j], j ? x[j - 1] : prev);
            flags |= (unsigned)keep << j;
            cnt += keep;
        }
    }
    SelC local, tot, P;
    select_offsets<BLOCK>(cnt, tile, a, s_wtot, &s_prefix, local, tot, P);
    write_packed<BLOCK, ITEMS>((K*)&s_raw[0][0], x, flags, local, tot, kout, P);
    if constexpr (VALS) write_packed<BLOCK, ITEMS>((V*)&s_raw[0][0], v, flags, local, tot, vout, P);
}

template <int BLOCK, int ITEMS, class T, class In, class Sten, class TOut, class FOut, class Pred>
TRTC_DEV void partition_tile(In& in, Sten& sten, TOut& tout, FOut& fout, Pred& pred, const CopyIfArgs& a)
{
    constexpr int WAVES = BLOCK / 64;
    constexpr int WELEMS = 64 * ITEMS;
    constexpr size_t TILE = (size_t)BLOCK * ITEMS;
    typedef typename NoConst<typename Sten::value_t>::type ST;
    constexpr int ESZ = sizeof(T) > sizeof(ST) ? sizeof(T) : sizeof(ST);
    __shared__ __attribute__((aligned(16))) unsigned char s_raw[WAVES][WaveTile<ITEMS>::padded * ESZ];
    __shared__ SelC s_wtot[WAVES];
    __shared__ SelC s_prefix;
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
    const unsigned valid = full ? (unsigned)TILE : (unsigned)(n - tile_base);
    const int nvalid = lane_valid<ITEMS>(wave_base, n, full);

    T x[ITEMS];
    load_blocked<ITEMS>(in, wave_base, n, full, nvalid, (T*)s_raw[wave], x);
    unsigned flags = 0, cnt = 0;
    if constexpr (!_IsNoStencil<Sten>::value) {
        ST s[ITEMS];
        load_blocked<ITEMS>(sten, wave_base, n, full, nvalid, (ST*)s_raw[wave], s);
#pragma unroll
        for (int j = 0; j < ITEMS; j++)
            if (j < nvalid) {
                bool f = (bool)pred(s[j]);
                flags |= (unsigned)f << j;
                cnt += f;
            }
    } else {
#pragma unroll
        for (int j = 0; j < ITEMS; j++)
            if (j < nvalid) {
                bool f = (bool)pred(x[j]);
                flags |= (unsigned)f << j;
                cnt += f;
            }
    }
    SelC local, tot, P;
    select_offsets<BLOCK>(cnt, tile, a, s_wtot, &s_prefix, local, tot, P);
    write_packed<BLOCK, ITEMS>((T*)&s_raw[0][0], x, flags, local, tot, tout, P);
    // rejected items: this thread's first slot = its first position minus the
    // selected items before it; the tile's run starts at tile_base - P
    const unsigned vmask = nvalid >= 32 ? ~0u : ((1u << nvalid) - 1u);
    const SelC p0 = (SelC)wave * WELEMS + (SelC)lane * ITEMS;
    write_packed<BLOCK, ITEMS>((T*)&s_raw[0][0], x, ~flags & vmask, nvalid ? p0 - local : 0, (SelC)valid - tot,
                               fout, (SelC)tile_base - P);
}

// Reduce_By_Key carried state: segmented value, "a segment starts here" flag,
// number of segment tails.
template <class V> struct RbkState
{
    V v;
    unsigned flag;
    unsigned tails;
};

template <class V, class Op> struct RbkOp
{
    Op& op;
    TRTC_DEV RbkState<V> operator()(const RbkState<V>& a, const RbkState<V>& b) const
    {
        RbkState<V> r;
        r.flag = a.flag | b.flag;
        r.v = b.flag ? b.v : (V)op(a.v, b.v);
        r.tails = a.tails + b.tails;
        return r;
    }
};

template <int BLOCK, int ITEMS, class V, class K, class Keys, class Vals, class KOut, class VOut, class Pred, class Op>
TRTC_DEV void reduce_by_key_tile(Keys& keys, Vals& vals, KOut& kout, VOut& vout, Pred& pred, Op& op,
                                 const CopyIfArgs& a)
{
    constexpr int WAVES = BLOCK / 64;
    constexpr int WELEMS = 64 * ITEMS;
    constexpr size_t TILE = (size_t)BLOCK * ITEMS;
    typedef RbkState<V> S;
    constexpr int ESZ = sizeof(V) > sizeof(K) ? sizeof(V) : sizeof(K);
    __shared__ __attribute__((aligned(16))) unsigned char s_raw[WAVES][WaveTile<ITEMS>::padded * ESZ];
    __shared__ S s_wtot[WAVES];
    __shared__ int s_wvalid[WAVES];
    __shared__ S s_prefix;
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
    const int nvalid = lane_valid<ITEMS>(wave_base, n, full);
    RbkOp<V, Op> sop{op};

    K kx[ITEMS];
    V x[ITEMS];
    load_blocked<ITEMS>(keys, wave_base, n, full, nvalid, (K*)s_raw[wave], kx);
    // neighbour keys: before this lane's first item and after its last
    K kprev = shfl_up(kx[ITEMS - 1], 1);
    K knext = shfl_down(kx[0], 1);
    if (lane == 0 && wave_base > 0 && wave_base < n) kprev = (K)keys[wave_base - 1];
    if (lane == 63 && wave_base + WELEMS < n) knext = (K)keys[wave_base + WELEMS];
    load_blocked<ITEMS>(vals, wave_base, n, full, nvalid, (V*)s_raw[wave], x);

    const size_t g0 = wave_base + (size_t)lane * ITEMS;
    unsigned heads = 0, tails = 0;
#pragma unroll
    for (int j = 0; j < ITEMS; j++) {
        if (j < nvalid) {
            bool head = (g0 + j == 0) || !(bool)pred(j ? kx[j - 1] : kprev, kx[j]);
            bool tail = (g0 + j == n - 1) || !(bool)pred(kx[j], j + 1 < ITEMS ? kx[j + 1] : knext);
            heads |= (unsigned)head << j;
            tails |= (unsigned)tail << j;
        }
    }

    // thread aggregate
    S t;
    t.v = x[0];
    t.flag = heads & 1u;
    t.tails = (unsigned)__builtin_popcount(tails);
#pragma unroll
    for (int j = 1; j < ITEMS; j++) {
        if (j < nvalid) {
            if ((heads >> j) & 1u) {
                t.v = x[j];
                t.flag = 1;
            } else {
                t.v = (V)op(t.v, x[j]);
            }
        }
    }
    const bool tvalid = nvalid > 0;
    S winc = full ? wave_inclusive_scan(t, sop) : wave_inclusive_scan_partial(t, tvalid, sop);
    unsigned long long vmask = ballot(tvalid);
    const bool wvalid = vmask != 0;
    const unsigned wlast = wvalid ? 63 - (unsigned)__builtin_clzll(vmask) : 0;
    if (lane == wlast) {
        s_wtot[wave] = winc;
        s_wvalid[wave] = wvalid ? 1 : 0;
    }
    S wexcl = shfl_up(winc, 1);
    __syncthreads();

    S wpre, tot;
    bool wpre_has = false, tot_has = false;
#pragma unroll
    for (int w = 0; w < WAVES; w++) {
        if (!s_wvalid[w]) continue;
        S q = s_wtot[w];
        if ((unsigned)w < wave) {
            wpre = wpre_has ? sop(wpre, q) : q;
            wpre_has = true;
        }
        tot = tot_has ? sop(tot, q) : q;
        tot_has = true;
    }

    S P;
    bool P_has = false;
    if (tile == 0) {
        if (threadIdx.x == 0) publish<S>(a.lb.incl, 0, tot);
    } else {
        if (threadIdx.x == 0) publish<S>(a.lb.agg, tile, tot);
        if (wave == 0) {
            S pre = lookback<S>(a.lb, tile, sop);
            if (lane == 0) {
                publish<S>(a.lb.incl, tile, sop(pre, tot));
                s_prefix = pre;
            }
        }
        __syncthreads();
        P = s_prefix;
        P_has = true;
    }
    const SelC Pc = P_has ? (SelC)P.tails : 0;
    if (tile == a.num_tiles - 1 && threadIdx.x == 0) *a.count = Pc + tot.tails;

    S run = P;
    bool run_has = P_has;
    if (wpre_has) {
        run = run_has ? sop(run, wpre) : wpre;
        run_has = true;
    }
    if (lane > 0 && tvalid) {
        run = run_has ? sop(run, wexcl) : wexcl;
        run_has = true;
    }
    // this thread's first tail slot within the tile
    const SelC local = (wpre_has ? (SelC)wpre.tails : 0) + ((lane > 0 && tvalid) ? (SelC)wexcl.tails : 0);

    // inclusive segmented values; a tail's value is its segment's reduction
#pragma unroll
    for (int j = 0; j < ITEMS; j++) {
        if (j < nvalid) {
            if (((heads >> j) & 1u) || !run_has) run.v = x[j];
            else run.v = (V)op(run.v, x[j]);
            run_has = true;
            x[j] = run.v;
        }
    }
    write_packed<BLOCK, ITEMS>((K*)&s_raw[0][0], kx, tails, local, tot.tails, kout, Pc);
    write_packed<BLOCK, ITEMS>((V*)&s_raw[0][0], x, tails, local, tot.tails, vout, Pc);
}

}  // namespace trtc
