// Single-pass segmented scan (Inclusive/Exclusive_Scan_By_Key) for gfx950.
//
// Semantics follow scan_by_key.cpp:4-53 and general_scan_by_key.cpp:8-143: a
// segment starts at i == 0 and wherever !pred(key[i-1], key[i]); the scan of
// the values restarts at every segment head.  Inclusive: out[i] = v[head] op
// ... op v[i].  Exclusive: out[head] = init and out[i] = init op v[head] op
// ... op v[i-1] (the reference's inclusive scan of src(i) = head ? init :
// v[i-1], regrouped by associativity so that a tile never reads another
// tile's values — in-place calls stay safe).
//
// The carried state is the pair (flag, value) with the segmented operator
//   (fa, a) . (fb, b) = (fa | fb, fb ? b : a op b)
// which is associative, so the same tile / wave / lookback machinery as the
// plain scan applies.
#pragma once
#include "trtc_prims.hpp"

namespace trtc {

template <class V> struct Seg
{
    V v;
    unsigned flag;
};

template <class V, class Op> struct SegOp
{
    Op& op;
    TRTC_DEV Seg<V> operator()(const Seg<V>& a, const Seg<V>& b) const
    {
        Seg<V> r;
        r.flag = a.flag | b.flag;
        r.v = b.flag ? b.v : (V)op(a.v, b.v);
        return r;
    }
};

template <int BLOCK, int ITEMS, class V, class K, class Keys, class Vals, class Out, class Pred, class Op>
TRTC_DEV void scan_by_key_tile(Keys& keys, Vals& vals, Out& out, Pred& pred, Op& op, const V& init,
                               const ScanByKeyArgs& a)
{
    constexpr int WAVES = BLOCK / 64;
    constexpr int WELEMS = 64 * ITEMS;
    constexpr size_t TILE = (size_t)BLOCK * ITEMS;
    typedef Seg<V> S;
    // one transpose buffer per wave, used for the keys, the values, then the
    // outputs (each load syncs the wave before it overwrites the buffer)
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
    V* ldv = (V*)s_raw[wave];
    K* ldk = (K*)s_raw[wave];
    SegOp<V, Op> sop{op};

    // blocked loads (16-byte vector loads for contiguous views); the key
    // before this lane's first item is the previous lane's last, or for lane
    // 0 the item before the wave (keys are never written by the scan)
    const int nvalid = lane_valid<ITEMS>(wave_base, n, full);
    K kx[ITEMS];
    V x[ITEMS];
    load_blocked<ITEMS>(keys, wave_base, n, full, nvalid, ldk, kx);
    load_blocked<ITEMS>(vals, wave_base, n, full, nvalid, ldv, x);
    K kprev = shfl_up(kx[ITEMS - 1], 1);
    // (a wave that starts at or past n has no items and reads nothing)
    if (lane == 0 && wave_base > 0 && wave_base < n) kprev = (K)keys[wave_base - 1];
    unsigned heads = 0;
#pragma unroll
    for (int j = 0; j < ITEMS; j++) {
        if (j < nvalid) {
            bool head = (wave_base + (size_t)lane * ITEMS + j == 0) || !(bool)pred(j ? kx[j - 1] : kprev, kx[j]);
            heads |= (unsigned)head << j;
        }
    }

    // thread aggregate (flag, value)
    S t;
    t.v = x[0];
    t.flag = heads & 1u;
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

    V y[ITEMS];
#pragma unroll
    for (int j = 0; j < ITEMS; j++) {
        if (j < nvalid) {
            const bool head = (heads >> j) & 1u;
            if (a.exclusive) y[j] = head ? init : (V)op(init, run.v);
            if (head || !run_has) run.v = x[j];
            else run.v = (V)op(run.v, x[j]);
            run_has = true;
            if (!a.exclusive) y[j] = run.v;
        }
    }

    wave_sync();
#pragma unroll
    for (int j = 0; j < ITEMS; j++)
        if (j < nvalid) ldv[WaveTile<ITEMS>::pad((int)lane * ITEMS + j)] = y[j];
    wave_sync();
    typedef typename Out::value_t OutT;
#pragma unroll
    for (int k = 0; k < ITEMS; k++) {
        int i = k * 64 + (int)lane;
        if (full || wave_base + i < n) out[wave_base + i] = (OutT)ldv[WaveTile<ITEMS>::pad(i)];
    }
}

// Items per thread of the by-key scan (512 threads): 128 bytes per lane for
// 4- and 8-byte elements, as the compaction tiles.
template <class V> struct ScanByKeyItems
{
    static constexpr int value = sizeof(V) <= 4 ? 32 : (sizeof(V) <= 8 ? 16 : 4);
};

}  // namespace trtc
