// Deterministic single-launch reduction for gfx950.
//
// Semantics follow reduce.cpp:5-31 / general_reduce.cpp:55-95: the result is
// v0 op v1 op ... op v(n-1), with init (when given) as element 0 of the
// sequence; the reduction runs in the vector's element type.  The combination
// order is fixed by n and the grid size alone (not by scheduling): every wave
// owns one contiguous range, reduces it chunk by chunk in index order (each
// chunk: coalesced loads, LDS transpose, sequential thread reduce, ordered wave
// reduce), the workgroup combines its waves in order, and the last workgroup
// to finish combines the per-workgroup partials in order.  The op is never
// assumed commutative, and repeated calls give bit-identical results.
#pragma once
#include "trtc_prims.hpp"

namespace trtc {

template <class T> struct Maybe
{
    T v;
    unsigned has;
};

template <class T, class Op> TRTC_DEV void fold(Maybe<T>& acc, const T& v, Op& op)
{
    acc.v = acc.has ? (T)op(acc.v, v) : v;
    acc.has = 1;
}

// Ordered reduce of one value per lane where valid lanes form a prefix;
// the result (if any lane is valid) is returned in every lane.
template <class T, class Op> TRTC_DEV Maybe<T> wave_reduce_prefix(const T& x, bool valid, Op& op)
{
    unsigned long long vm = ballot(valid);
    Maybe<T> r;
    r.has = vm != 0;
    if (!r.has) return r;
    T inc = wave_inclusive_scan_partial(x, valid, op);
    unsigned last = 63 - (unsigned)__builtin_clzll(vm);
    r.v = shfl(inc, (int)last);
    return r;
}

// Ordered reduce of src[beg, end) by one wave, chunk by chunk in index order
// (coalesced loads, LDS transpose, sequential thread reduce, ordered wave
// reduce).  The result is valid in every lane.
template <int ITEMS, class T, class Src, class Op>
TRTC_DEV Maybe<T> wave_range_reduce(Src& src, size_t beg, size_t end, T* lds, Op& op)
{
    constexpr int CHUNK = 64 * ITEMS;
    const unsigned lane = lane_id();
    Maybe<T> acc;
    acc.has = 0;
    for (size_t base = beg; base < end; base += CHUNK) {
        const bool full = base + CHUNK <= end;
#pragma unroll
        for (int k = 0; k < ITEMS; k++) {
            int i = k * 64 + (int)lane;
            if (full || base + i < end) lds[WaveTile<ITEMS>::pad(i)] = (T)src[base + i];
        }
        wave_sync();
        int nvalid = ITEMS;
        if (!full) {
            size_t first = base + (size_t)lane * ITEMS;
            nvalid = first >= end ? 0 : (int)((end - first) < (size_t)ITEMS ? (end - first) : ITEMS);
        }
        T t = lds[WaveTile<ITEMS>::pad((int)lane * ITEMS)];
#pragma unroll
        for (int j = 1; j < ITEMS; j++)
            if (j < nvalid) t = (T)op(t, lds[WaveTile<ITEMS>::pad((int)lane * ITEMS + j)]);
        wave_sync();
        Maybe<T> c;
        if (full) {
            T inc = wave_inclusive_scan(t, op);
            c.v = shfl(inc, 63);
            c.has = 1;
        } else {
            c = wave_reduce_prefix(t, nvalid > 0, op);
        }
        if (c.has) fold(acc, c.v, op);
    }
    return acc;
}

template <int BLOCK, int ITEMS, class T, class Src, class Op>
TRTC_DEV void reduce_body(Src& src, Op& op, const T& init, const ReduceArgs& a)
{
    constexpr int WAVES = BLOCK / 64;
    __shared__ T s_x[WAVES][WaveTile<ITEMS>::padded];
    __shared__ Maybe<T> s_w[WAVES];
    __shared__ bool s_last;

    const unsigned wave = threadIdx.x / 64;
    const unsigned lane = lane_id();
    const size_t n = a.n;
    const size_t gw = (size_t)blockIdx.x * WAVES + wave;
    size_t beg = gw * a.per_wave;
    size_t end = beg + a.per_wave;
    if (beg > n) beg = n;
    if (end > n) end = n;
    T* lds = s_x[wave];

    Maybe<T> acc = wave_range_reduce<ITEMS, T>(src, beg, end, lds, op);
    if (lane == 0) s_w[wave] = acc;
    __syncthreads();

    if (threadIdx.x == 0) {
        Maybe<T> b;
        b.has = 0;
        for (int w = 0; w < WAVES; w++)
            if (s_w[w].has) fold(b, s_w[w].v, op);
        publish<Maybe<T>>(a.partials, blockIdx.x, b);
        unsigned t = __hip_atomic_fetch_add(a.done, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        s_last = (t == gridDim.x - 1);
    }
    __syncthreads();
    if (!s_last || wave != 0) return;

    // Last workgroup, wave 0: combine the partials in workgroup order.  Each
    // partial carries its own valid bits, so spin until every one is visible.
    Maybe<T> tot;
    tot.has = 0;
    const unsigned G = gridDim.x;
    for (unsigned base = 0; base < G; base += 64) {
        unsigned b = base + lane;
        Maybe<T> p;
        p.has = 0;
        if (b < G) {
            while (!try_read<Maybe<T>>(a.partials, b, p)) __builtin_amdgcn_s_sleep(1);
        }
        // non-empty partials form a prefix (ranges are filled from the front)
        Maybe<T> c = wave_reduce_prefix(p.v, b < G && p.has, op);
        if (c.has) fold(tot, c.v, op);
    }
    if (lane == 0) {
        T r;
        if (a.has_init) r = tot.has ? (T)op(init, tot.v) : init;
        else r = tot.v;
        *(T*)a.result = r;
    }
}

// Batch of independent reductions (no init): segment s is cut into tiles of
// TILE elements, each wave of a tile reduces a contiguous quarter-range in
// order, the tile combines its waves in order and publishes a partial, and
// the last tile of the segment to finish combines the segment's partials in
// tile order.  The order depends only on the segment's length, so a
// segment's result is the same whatever else is in the batch.
template <int BLOCK, int ITEMS, int TILE, class T, class Op>
TRTC_DEV void seg_reduce_tile(Op& op, const SegReduceArgs& a)
{
    constexpr int WAVES = BLOCK / 64;
    constexpr size_t PER_WAVE = (size_t)TILE / WAVES;
    __shared__ T s_x[WAVES][WaveTile<ITEMS>::padded];
    __shared__ Maybe<T> s_w[WAVES];
    __shared__ bool s_last;

    const unsigned t = blockIdx.x;  // grid = num_tiles
    const SegReduceDesc& d = a.descs[a.tile_seg[t]];
    const unsigned wave = threadIdx.x / 64;
    const unsigned lane = lane_id();
    PtrView<const T> src{(const T*)d.in};
    const size_t n = d.n;
    size_t beg = (size_t)(t - d.first_tile) * TILE + wave * PER_WAVE;
    size_t end = beg + PER_WAVE;
    if (beg > n) beg = n;
    if (end > n) end = n;
    Maybe<T> acc = wave_range_reduce<ITEMS, T>(src, beg, end, s_x[wave], op);
    if (lane == 0) s_w[wave] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        Maybe<T> b;
        b.has = 0;
        for (int w = 0; w < WAVES; w++)
            if (s_w[w].has) fold(b, s_w[w].v, op);
        publish<Maybe<T>>(a.partials, t, b);
        unsigned v = __hip_atomic_fetch_add(a.seg_done + a.tile_seg[t], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        s_last = (v == d.num_tiles - 1);
    }
    __syncthreads();
    if (!s_last || wave != 0) return;
    Maybe<T> tot;
    tot.has = 0;
    for (unsigned base = 0; base < d.num_tiles; base += 64) {
        unsigned b = base + lane;
        Maybe<T> p;
        p.has = 0;
        if (b < d.num_tiles) {
            while (!try_read<Maybe<T>>(a.partials, d.first_tile + b, p)) __builtin_amdgcn_s_sleep(1);
        }
        Maybe<T> c = wave_reduce_prefix(p.v, b < d.num_tiles && p.has, op);
        if (c.has) fold(tot, c.v, op);
    }
    if (lane == 0) *(T*)d.result = tot.has ? tot.v : (T)0;
}

template <class T> struct ReduceItems
{
    static constexpr int value = sizeof(T) <= 4 ? 16 : (sizeof(T) <= 8 ? 8 : 4);
};

}  // namespace trtc
