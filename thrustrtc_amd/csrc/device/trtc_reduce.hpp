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

// ---------------------------------------------------------------------------
// Streaming path for built-in commutative ops (Plus, Maximum, Minimum,
// Multiplies) on contiguous arithmetic data.  The reference's own tree
// (general_reduce.cpp:64-75) combines element i with element i+s, i.e. it
// already relies on commutativity; here the order is fixed by n, the grid
// and the data's alignment: each thread folds its grid-strided 16-byte
// vectors into UNROLL accumulators, then threads, waves and workgroups are
// combined in a fixed tree.  Loads are nontemporal and UNROLL deep per
// thread, so the kernel streams at close to copy bandwidth.
// ---------------------------------------------------------------------------
template <class T> struct IsArith { static constexpr bool value = false; };
template <> struct IsArith<int8_t> { static constexpr bool value = true; };
template <> struct IsArith<uint8_t> { static constexpr bool value = true; };
template <> struct IsArith<int16_t> { static constexpr bool value = true; };
template <> struct IsArith<uint16_t> { static constexpr bool value = true; };
template <> struct IsArith<int32_t> { static constexpr bool value = true; };
template <> struct IsArith<uint32_t> { static constexpr bool value = true; };
template <> struct IsArith<int64_t> { static constexpr bool value = true; };
template <> struct IsArith<uint64_t> { static constexpr bool value = true; };
template <> struct IsArith<float> { static constexpr bool value = true; };
template <> struct IsArith<double> { static constexpr bool value = true; };

template <class Op> struct IsCommutative { static constexpr bool value = false; };
template <> struct IsCommutative<Plus> { static constexpr bool value = true; };
template <> struct IsCommutative<Maximum> { static constexpr bool value = true; };
template <> struct IsCommutative<Minimum> { static constexpr bool value = true; };
template <> struct IsCommutative<Multiplies> { static constexpr bool value = true; };

template <class Src, class Op, class T> struct StreamingReduce { static constexpr bool value = false; };
template <class Op, class T> struct StreamingReduce<VectorView<T>, Op, T>
{
    static constexpr bool value = IsArith<T>::value && IsCommutative<Op>::value;
};

// Butterfly-free ordered tree over lanes: lane l takes lane l+d for d = 32..1;
// the result is in lane 0.
template <class T, class Op> TRTC_DEV Maybe<T> wave_tree_reduce(Maybe<T> x, Op& op)
{
    const unsigned l = lane_id();
#pragma unroll
    for (unsigned d = 32; d >= 1; d >>= 1) {
        T yv = shfl_down(x.v, d);
        unsigned yh = shfl_down(x.has, d);
        if (l + d < 64 && yh) {
            x.v = x.has ? (T)op(x.v, yv) : yv;
            x.has = 1;
        }
    }
    return x;
}

template <int BLOCK, class T, class Op>
TRTC_DEV void reduce_streaming(const T* p, Op& op, const T& init, const ReduceArgs& a)
{
    constexpr int WAVES = BLOCK / 64;
    constexpr int V = 16 / sizeof(T);
    constexpr int UNROLL = 4;
    typedef unsigned v4u __attribute__((ext_vector_type(4)));
    __shared__ Maybe<T> s_w[WAVES];
    __shared__ bool s_last;

    const size_t n = a.n;
    const size_t tid = (size_t)blockIdx.x * BLOCK + threadIdx.x;
    const size_t nthreads = (size_t)gridDim.x * BLOCK;
    // elements before the first 16-byte boundary, whole vectors, the tail
    size_t head = ((16 - ((size_t)p & 15)) & 15) / sizeof(T);
    if (((size_t)p % sizeof(T)) != 0 || head > n) head = n;  // no aligned vectors: all scalar
    const size_t nv = (n - head) / V;
    const size_t tail_begin = head + nv * V;
    const v4u* q = (const v4u*)(p + head);

    Maybe<T> acc[UNROLL];
#pragma unroll
    for (int u = 0; u < UNROLL; u++) acc[u].has = 0;
    // each wave reads UNROLL KiB contiguously per step (one 1 KiB piece per
    // load instruction); the waves take the chunks grid-stride
    const size_t wave_g = tid / 64, nwaves = nthreads / 64, lane64 = tid & 63;
    for (size_t c0 = wave_g * (UNROLL * 64) + lane64; c0 < nv; c0 += nwaves * (UNROLL * 64)) {
        v4u x[UNROLL];
#pragma unroll
        for (int u = 0; u < UNROLL; u++) {
            size_t v = c0 + (size_t)u * 64;
            if (v < nv) x[u] = __builtin_nontemporal_load(q + v);
        }
#pragma unroll
        for (int u = 0; u < UNROLL; u++) {
            size_t v = c0 + (size_t)u * 64;
            if (v < nv) {
                T e[V];
                __builtin_memcpy(e, &x[u], 16);
#pragma unroll
                for (int k = 0; k < V; k++) fold(acc[u], e[k], op);
            }
        }
    }
    // head and tail elements (grid-strided scalars)
    for (size_t i = tid; i < head; i += nthreads) fold(acc[0], p[i], op);
    for (size_t i = tail_begin + tid; i < n; i += nthreads) fold(acc[0], p[i], op);
    Maybe<T> t = acc[0];
#pragma unroll
    for (int u = 1; u < UNROLL; u++)
        if (acc[u].has) fold(t, acc[u].v, op);
    t = wave_tree_reduce(t, op);
    const unsigned wave = threadIdx.x / 64;
    if (lane_id() == 0) s_w[wave] = t;
    __syncthreads();
    if (threadIdx.x == 0) {
        Maybe<T> b;
        b.has = 0;
        for (int w = 0; w < WAVES; w++)
            if (s_w[w].has) fold(b, s_w[w].v, op);
        publish<Maybe<T>>(a.partials, blockIdx.x, b);
        unsigned d = __hip_atomic_fetch_add(a.done, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        s_last = (d == gridDim.x - 1);
    }
    __syncthreads();
    if (!s_last || wave != 0) return;
    Maybe<T> tot;
    tot.has = 0;
    const unsigned G = gridDim.x;
    for (unsigned base = 0; base < G; base += 64) {
        unsigned b = base + lane_id();
        Maybe<T> pp;
        pp.has = 0;
        if (b < G) {
            while (!try_read<Maybe<T>>(a.partials, b, pp)) __builtin_amdgcn_s_sleep(1);
        }
        Maybe<T> c = wave_tree_reduce(pp, op);
        if (lane_id() == 0 && c.has) fold(tot, c.v, op);
    }
    if (lane_id() == 0) {
        T r;
        if (a.has_init) r = tot.has ? (T)op(init, tot.v) : init;
        else r = tot.v;
        *(T*)a.result = r;
    }
}

// ---------------------------------------------------------------------------
// Min/Max_Element on a contiguous arithmetic vector (extrema.cpp:5-80 ask for
// the position of the first smallest / first largest element under `comp`).
// The kernel streams like reduce_streaming: every lane keeps the best
// (index, value) pair of the elements it reads, in increasing index order per
// accumulator, so a later element replaces the best one only when it is
// strictly better.  Pairs from different lanes, waves and workgroups meet in
// no fixed order, so their combination breaks ties by the lower index: for a
// comparator that is a strict weak order the result is the first extreme
// element whatever the grid.
// ---------------------------------------------------------------------------
template <class T> struct ExtPair
{
    unsigned long long i;  // at offset 0: the host reads the index from there
    T v;
};

template <class T, class Comp, bool MAX> struct ExtCombine
{
    Comp comp;
    TRTC_DEV bool better(const T& x, const T& y) { return MAX ? (bool)comp(y, x) : (bool)comp(x, y); }
    TRTC_DEV ExtPair<T> operator()(const ExtPair<T>& p, const ExtPair<T>& q)
    {
        const bool qb = better(q.v, p.v);
        const bool pb = better(p.v, q.v);
        return (qb || (!pb && q.i < p.i)) ? q : p;
    }
};

template <int BLOCK, class T, class Comp, bool MAX>
TRTC_DEV void extreme_streaming(const T* p, Comp& comp, const ReduceArgs& a)
{
    typedef ExtPair<T> P;
    constexpr int WAVES = BLOCK / 64;
    constexpr int V = 16 / sizeof(T);
    constexpr int UNROLL = 4;
    typedef unsigned v4u __attribute__((ext_vector_type(4)));
    __shared__ Maybe<P> s_w[WAVES];
    __shared__ bool s_last;
    ExtCombine<T, Comp, MAX> op{comp};

    const size_t n = a.n;
    const size_t tid = (size_t)blockIdx.x * BLOCK + threadIdx.x;
    const size_t nthreads = (size_t)gridDim.x * BLOCK;
    size_t head = ((16 - ((size_t)p & 15)) & 15) / sizeof(T);
    if (((size_t)p % sizeof(T)) != 0 || head > n) head = n;
    const size_t nv = (n - head) / V;
    const size_t tail_begin = head + nv * V;
    const v4u* q = (const v4u*)(p + head);

    // accumulator u sees increasing indices only: strictly better replaces
    T bv[UNROLL];
    unsigned long long bi[UNROLL];
    bool has[UNROLL];
#pragma unroll
    for (int u = 0; u < UNROLL; u++) {
        has[u] = false;
        bv[u] = T();
        bi[u] = 0;
    }
    const size_t wave_g = tid / 64, nwaves = nthreads / 64, lane64 = tid & 63;
    for (size_t c0 = wave_g * (UNROLL * 64) + lane64; c0 < nv; c0 += nwaves * (UNROLL * 64)) {
        v4u x[UNROLL];
#pragma unroll
        for (int u = 0; u < UNROLL; u++) {
            size_t v = c0 + (size_t)u * 64;
            if (v < nv) x[u] = __builtin_nontemporal_load(q + v);
        }
#pragma unroll
        for (int u = 0; u < UNROLL; u++) {
            size_t v = c0 + (size_t)u * 64;
            if (v < nv) {
                T e[V];
                __builtin_memcpy(e, &x[u], 16);
                const unsigned long long base = head + v * V;
#pragma unroll
                for (int k = 0; k < V; k++) {
                    const bool take = !has[u] || op.better(e[k], bv[u]);
                    bv[u] = take ? e[k] : bv[u];
                    bi[u] = take ? base + k : bi[u];
                    has[u] = true;
                }
            }
        }
    }
    Maybe<P> t;
    t.has = 0;
#pragma unroll
    for (int u = 0; u < UNROLL; u++)
        if (has[u]) fold(t, P{bi[u], bv[u]}, op);
    for (size_t i = tid; i < head; i += nthreads) fold(t, P{i, p[i]}, op);
    for (size_t i = tail_begin + tid; i < n; i += nthreads) fold(t, P{i, p[i]}, op);
    t = wave_tree_reduce(t, op);
    const unsigned wave = threadIdx.x / 64;
    if (lane_id() == 0) s_w[wave] = t;
    __syncthreads();
    if (threadIdx.x == 0) {
        Maybe<P> b;
        b.has = 0;
        for (int w = 0; w < WAVES; w++)
            if (s_w[w].has) fold(b, s_w[w].v, op);
        publish<Maybe<P>>(a.partials, blockIdx.x, b);
        unsigned d = __hip_atomic_fetch_add(a.done, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        s_last = (d == gridDim.x - 1);
    }
    __syncthreads();
    if (!s_last || wave != 0) return;
    Maybe<P> tot;
    tot.has = 0;
    const unsigned G = gridDim.x;
    for (unsigned base = 0; base < G; base += 64) {
        unsigned b = base + lane_id();
        Maybe<P> pp;
        pp.has = 0;
        if (b < G) {
            while (!try_read<Maybe<P>>(a.partials, b, pp)) __builtin_amdgcn_s_sleep(1);
        }
        Maybe<P> c = wave_tree_reduce(pp, op);
        if (lane_id() == 0 && c.has) fold(tot, c.v, op);
    }
    if (lane_id() == 0) *(P*)a.result = tot.v;  // n >= 1: some element was read
}

template <int BLOCK, class Src, class Op, class T> struct ReduceDispatch
{
    template <int ITEMS> TRTC_DEV static void run(Src& src, Op& op, const T& init, const ReduceArgs& a)
    {
        reduce_body<BLOCK, ITEMS, T>(src, op, init, a);
    }
};
template <int BLOCK, class Op, class T> struct ReduceDispatch<BLOCK, VectorView<T>, Op, T>
{
    template <int ITEMS> TRTC_DEV static void run(VectorView<T>& src, Op& op, const T& init, const ReduceArgs& a)
    {
        if constexpr (StreamingReduce<VectorView<T>, Op, T>::value)
            reduce_streaming<BLOCK, T>(src._data, op, init, a);
        else
            reduce_body<BLOCK, ITEMS, T>(src, op, init, a);
    }
};

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
