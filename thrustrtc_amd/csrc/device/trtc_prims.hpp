// Wave64 / workgroup primitives and the single-pass decoupled-lookback protocol
// shared by the scan, scan-by-key, copy-if and reduce kernels (CDNA4, gfx950).
//
// Memory model.  Tiles publish their aggregate / inclusive prefix in 64-bit
// "status words" {payload:32 | valid:32}.  A value of any trivially copyable
// type T is split into ceil(sizeof(T)/4) such words, written with relaxed
// agent-scope atomic stores and read with relaxed agent-scope atomic loads.
// Because every word carries its own valid bit, a reader never needs a fence:
// it only accepts a value once all its words read back valid.  Aggregates and
// inclusive prefixes live in separate arrays so that a reader never mixes the
// two.  These atomics are coherent across the 8 XCDs (agent scope), which the
// per-XCD L2s are not for plain stores.
#pragma once
#include "trtc_builtin.hpp"
#include "trtc_args.h"

namespace trtc {

constexpr int kWave = 64;

// Global-memory pointers.  Buffers whose addresses come from memory (a sort's
// pass plan, a batch's segment table) are generic pointers to the compiler,
// and generic accesses compile to flat loads and stores, which count on the
// LDS counter as well: every LDS wait of the kernel then also waits for the
// loads and stores in flight.  Accesses through these pointer types compile
// to global_load / global_store (2^28 u32/u32, long-tile radix pass: 0.930 ->
// 0.904 ms per pass, bit-identical output).
template <class T> using gmem_t = __attribute__((address_space(1))) T;
template <class T> TRTC_DEV gmem_t<T>* gmem(T* p) { return (gmem_t<T>*)p; }

TRTC_DEV unsigned lane_id()
{
    return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
}

TRTC_DEV unsigned long long ballot(bool p) { return __ballot(p); }

TRTC_DEV void wave_sync()
{
    // Orders this wave's LDS accesses: everything before is visible to every
    // lane of the wave after.
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <class T> struct Words
{
    static constexpr int n = (int)((sizeof(T) + 3) / 4);
};

// Read lane `src` of a value of any type (word-wise ds_bpermute).
template <class T> TRTC_DEV T shfl(const T& v, int src)
{
    constexpr int W = Words<T>::n;
    unsigned w[W];
    w[W - 1] = 0;
    __builtin_memcpy(w, &v, sizeof(T));
    unsigned r[W];
#pragma unroll
    for (int i = 0; i < W; i++) r[i] = (unsigned)__builtin_amdgcn_ds_bpermute(src << 2, (int)w[i]);
    T out;
    __builtin_memcpy(&out, r, sizeof(T));
    return out;
}

template <class T> TRTC_DEV T shfl_up(const T& v, unsigned delta)
{
    int l = (int)lane_id();
    int src = l - (int)delta;
    return shfl(v, src < 0 ? l : src);
}

template <class T> TRTC_DEV T shfl_down(const T& v, unsigned delta)
{
    int l = (int)lane_id();
    int src = l + (int)delta;
    return shfl(v, src > 63 ? l : src);
}

// Inclusive scan across the 64 lanes, earlier lanes always the left operand
// (correct for any associative op, commutative or not).
template <class T, class Op> TRTC_DEV T wave_inclusive_scan(T x, Op& op)
{
    unsigned l = lane_id();
#pragma unroll
    for (unsigned d = 1; d < 64; d <<= 1) {
        T y = shfl_up(x, d);
        if (l >= d) x = (T)op(y, x);
    }
    return x;
}

// Segmented variant with validity: lanes whose `valid` is false hold no value;
// they act as the identity.  Valid lanes always form a prefix of the wave.
template <class T, class Op> TRTC_DEV T wave_inclusive_scan_partial(T x, bool valid, Op& op)
{
    unsigned l = lane_id();
    unsigned long long vm = ballot(valid);
#pragma unroll
    for (unsigned d = 1; d < 64; d <<= 1) {
        T y = shfl_up(x, d);
        bool yv = l >= d && ((vm >> (l - d)) & 1ull);
        if (yv && valid) x = (T)op(y, x);
    }
    return x;
}

// ---------------------------------------------------------------------------
// Relaxed agent-scope atomics on status words
// ---------------------------------------------------------------------------
TRTC_DEV unsigned long long ld_status(const unsigned long long* p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
TRTC_DEV void st_status(unsigned long long* p, unsigned long long v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
TRTC_DEV unsigned ld_status32(const unsigned* p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
TRTC_DEV void st_status32(unsigned* p, unsigned v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Publish value v of tile `tile` in array `arr` (Words<T>::n words per tile).
template <class T> TRTC_DEV void publish(unsigned long long* arr, size_t tile, const T& v)
{
    constexpr int W = Words<T>::n;
    unsigned w[W];
    w[W - 1] = 0;
    __builtin_memcpy(w, &v, sizeof(T));
#pragma unroll
    for (int i = 0; i < W; i++)
        st_status(arr + tile * W + i, (1ull << 32) | (unsigned long long)w[i]);
}

// Try to read tile `tile`'s value; true if every word is valid.
template <class T> TRTC_DEV bool try_read(const unsigned long long* arr, size_t tile, T& out)
{
    constexpr int W = Words<T>::n;
    unsigned w[W];
    bool ok = true;
#pragma unroll
    for (int i = 0; i < W; i++) {
        unsigned long long s = ld_status(arr + tile * W + i);
        ok = ok && (s >> 32) != 0;
        w[i] = (unsigned)s;
    }
    __builtin_memcpy(&out, w, sizeof(T));
    return ok;
}

// Combine lanes 0..last (lane 0 = nearest predecessor tile, so higher lanes
// are earlier tiles and go on the left).  Returns the result in every lane.
template <class T, class Op> TRTC_DEV T wave_reverse_reduce(T x, unsigned last, Op& op)
{
    unsigned l = lane_id();
#pragma unroll
    for (unsigned d = 1; d < 64; d <<= 1) {
        T y = shfl_down(x, d);  // lane l + d: an earlier tile
        if (l + d <= last && (l & (2 * d - 1)) == 0) x = (T)op(y, x);
    }
    return shfl(x, 0);
}

// Decoupled lookback run by one full wave for tile `tile` > `first` (the
// sequence's first tile): returns the inclusive prefix of the tiles
// first .. tile-1.  The window of 64 predecessors
// is read in parallel (lane k reads tile - 1 - k).  A predecessor's
// inclusive prefix ends the walk; otherwise its aggregate is folded in and
// the window moves back by 64.  Every tile publishes its aggregate before it
// looks back and tile ids are handed out in launch order by an atomic counter,
// so every awaited tile is already resident and the wait ends.  The sequence's
// first tile publishes only an inclusive prefix, so the walk ends there at the
// latest; should it ever pass that tile (status left over from another launch
// in memory that was not cleared), it stops with what it folded instead of
// walking on forever.
template <class T, class Op>
TRTC_DEV T lookback(const LookbackArgs& lb, size_t tile, Op& op, size_t first = 0)
{
    unsigned l = lane_id();
    T acc;
    bool have = false;
    long long top = (long long)tile - 1;
    while (true) {
        long long j = top - (long long)l;
        T v;
        int st = 3;  // 3: before the sequence's first tile (not part of the window)
        if (j >= (long long)first) {
            // issue both reads before testing either (one latency, not two)
            T vi, va;
            bool oki = try_read<T>(lb.incl, (size_t)j, vi);
            bool oka = try_read<T>(lb.agg, (size_t)j, va);
            st = oki ? 2 : (oka ? 1 : 0);
            v = oki ? vi : va;
        }
        unsigned long long incl_mask = ballot(st == 2);
        unsigned long long wait_mask = ballot(st == 0);
        unsigned last;
        if (incl_mask) {
            last = (unsigned)__builtin_ctzll(incl_mask);
            unsigned long long below = (last == 63) ? ~0ull : ((2ull << last) - 1);
            if (wait_mask & below) {
                __builtin_amdgcn_s_sleep(2);
                continue;
            }
        } else {
            if (wait_mask) {
                __builtin_amdgcn_s_sleep(2);
                continue;
            }
            unsigned long long in_window = ballot(st != 3);
            if (!in_window) break;  // the window lies wholly before the first tile
            last = 63 - (unsigned)__builtin_clzll(in_window);
        }
        T w = wave_reverse_reduce(v, last, op);
        acc = have ? (T)op(w, acc) : w;
        have = true;
        if (incl_mask) break;
        top -= 64;
    }
    return acc;
}

// Raw-pointer view of one segment of a batch (device memory whose address
// comes from the batch's descriptor table: accessed as global memory).
template <class T> struct PtrView
{
    typedef T value_t;
    T* p;
    TRTC_DEV gmem_t<T>& operator[](size_t i) const { return gmem(p)[i]; }
};

// Raw element pointer of contiguous views (VectorView, PtrView), for the
// direct vector-load paths.
template <class T> struct NoConst { typedef T type; };
template <class T> struct NoConst<const T> { typedef T type; };
template <class A, class B> struct SameType { static constexpr bool value = false; };
template <class A> struct SameType<A, A> { static constexpr bool value = true; };

template <class View> struct Contig { static constexpr bool ok = false; typedef void elem; };
template <class U> struct Contig<VectorView<U>>
{
    static constexpr bool ok = true;
    typedef U elem;
    static TRTC_DEV U* ptr(const VectorView<U>& v) { return v._data; }
};
template <class U> struct Contig<PtrView<U>>
{
    static constexpr bool ok = true;
    typedef U elem;
    static TRTC_DEV U* ptr(const PtrView<U>& v) { return v.p; }
};

// True when a view holds T contiguously and T packs 16-byte vectors (4- and
// 8-byte scalars): a lane can then load its ITEMS consecutive elements with
// ITEMS*sizeof(T)/16 vector loads instead of a transpose through LDS.
template <class View, class T, int ITEMS> struct DirectVec
{
    static constexpr bool value = Contig<View>::ok &&
                                  SameType<typename NoConst<typename Contig<View>::elem>::type, T>::value &&
                                  (sizeof(T) == 4 || sizeof(T) == 8) && (ITEMS * sizeof(T)) % 16 == 0;
};

typedef unsigned trtc_v4u __attribute__((ext_vector_type(4)));

// ---------------------------------------------------------------------------
// Wave-local striped <-> blocked transposition through LDS.
// Striped: lane l owns elements k*64 + l.  Blocked: lane l owns elements
// l*ITEMS + j.  One pad word every 32 words keeps the blocked accesses
// bank-conflict-light.
// ---------------------------------------------------------------------------
template <int ITEMS> struct WaveTile
{
    static constexpr int elems = 64 * ITEMS;
    static constexpr int padded = elems + elems / 32 + 1;
    TRTC_DEV static int pad(int i) { return i + (i >> 5); }
};

// Valid items of this lane in a blocked layout (wave_base + lane*ITEMS + j).
template <int ITEMS> TRTC_DEV int lane_valid(size_t wave_base, size_t n, bool full)
{
    if (full) return ITEMS;
    size_t first = wave_base + (size_t)lane_id() * ITEMS;
    return first >= n ? 0 : (int)((n - first) < (size_t)ITEMS ? (n - first) : ITEMS);
}

// One wave's 64*ITEMS consecutive elements at p (16-byte aligned) in the
// blocked layout: striped 16-byte loads (lane l reads piece k*64 + l, so each
// load instruction covers 1 KiB contiguously) transposed through the wave's
// LDS buffer.  On 2^29 int32 the single-pass scan runs 1.36 -> 1.18 ms with
// these in place of each lane loading its own run (a 64-byte stride between
// the lanes of one load instruction).
template <int ITEMS, class T> TRTC_DEV void load_striped_vec(const T* p, T* lds, T (&x)[ITEMS])
{
    constexpr int PER = 16 / (int)sizeof(T);
    constexpr int NV = ITEMS / PER;
    static_assert(NV * PER == ITEMS, "whole 16-byte pieces per lane");
    const unsigned lane = lane_id();
    trtc_v4u v[NV];
#pragma unroll
    for (int k = 0; k < NV; k++) v[k] = __builtin_nontemporal_load(gmem((const trtc_v4u*)p) + k * 64 + lane);
    wave_sync();  // the buffer's previous contents are consumed
#pragma unroll
    for (int k = 0; k < NV; k++) {
        T t[PER];
        __builtin_memcpy(t, &v[k], 16);
        const int i = PER * (k * 64 + (int)lane);
#pragma unroll
        for (int e = 0; e < PER; e++) lds[WaveTile<ITEMS>::pad(i + e)] = t[e];
    }
    wave_sync();
#pragma unroll
    for (int j = 0; j < ITEMS; j++) x[j] = lds[WaveTile<ITEMS>::pad((int)lane * ITEMS + j)];
}

// As load_striped_vec for a p aligned to its element size but not to 16
// bytes: the 16-byte pieces covering [p, p + 64*ITEMS) are loaded (one piece
// more, read by lane 0) and each element lands in the LDS buffer at its
// offset from p.  The elements read before p and after the range share a
// 16-byte piece, hence a page, with an element of the range; they are
// dropped.  Sequences starting one element past a 16-byte boundary took
// twice as long with element loads (a 3.9e8-int32 scan 0.74 -> 1.42 ms).
template <int ITEMS, class T> TRTC_DEV void load_striped_vec_shifted(const T* p, T* lds, T (&x)[ITEMS])
{
    constexpr int PER = 16 / (int)sizeof(T);
    constexpr int NV = ITEMS / PER;
    constexpr int ELEMS = 64 * ITEMS;
    static_assert(NV * PER == ITEMS, "whole 16-byte pieces per lane");
    const unsigned lane = lane_id();
    const int m = (int)(((size_t)p & 15) / sizeof(T));  // 1 .. PER-1 elements past the boundary
    const gmem_t<const trtc_v4u>* q = gmem((const trtc_v4u*)((const char*)p - (size_t)m * sizeof(T)));
    trtc_v4u v[NV];
#pragma unroll
    for (int k = 0; k < NV; k++) v[k] = __builtin_nontemporal_load(q + k * 64 + lane);
    trtc_v4u last = {0u, 0u, 0u, 0u};
    if (lane == 0) last = __builtin_nontemporal_load(q + NV * 64);
    wave_sync();  // the buffer's previous contents are consumed
#pragma unroll
    for (int k = 0; k < NV; k++) {
        T t[PER];
        __builtin_memcpy(t, &v[k], 16);
        const int i = PER * (k * 64 + (int)lane) - m;
#pragma unroll
        for (int e = 0; e < PER; e++)
            if (i + e >= 0) lds[WaveTile<ITEMS>::pad(i + e)] = t[e];
    }
    if (lane == 0) {
        T t[PER];
        __builtin_memcpy(t, &last, 16);
        const int i = ELEMS - m;
#pragma unroll
        for (int e = 0; e < PER; e++)
            if (e < m) lds[WaveTile<ITEMS>::pad(i + e)] = t[e];
    }
    wave_sync();
#pragma unroll
    for (int j = 0; j < ITEMS; j++) x[j] = lds[WaveTile<ITEMS>::pad((int)lane * ITEMS + j)];
}

// Blocked load of a full wave range through 16-byte pieces when the view
// holds 4- or 8-byte elements contiguously (aligned or shifted); returns
// false, having loaded nothing, for any other view, a partial range or a
// pointer not aligned to its element size.
template <int ITEMS, class T, class View>
TRTC_DEV bool load_blocked_vec(const View& v, size_t wave_base, bool full, T* lds, T (&x)[ITEMS])
{
    if constexpr (DirectVec<View, T, ITEMS>::value) {
        if (!full) return false;
        const T* p = Contig<View>::ptr(v) + wave_base;
        if ((((size_t)p) & 15) == 0) {
            load_striped_vec<ITEMS>(p, lds, x);
            return true;
        }
        if ((((size_t)p) % sizeof(T)) == 0) {
            load_striped_vec_shifted<ITEMS>(p, lds, x);
            return true;
        }
    }
    return false;
}

// Blocked load of a wave's 64*ITEMS items: 16-byte pieces for contiguous
// views of 4/8-byte elements, else coalesced striped element loads, both
// transposed through the wave's LDS buffer `lds`.
template <int ITEMS, class T, class View>
TRTC_DEV void load_blocked(View& v, size_t wave_base, size_t n, bool full, int nvalid, T* lds, T (&x)[ITEMS])
{
    const unsigned lane = lane_id();
    if (load_blocked_vec<ITEMS>(v, wave_base, full, lds, x)) return;
    wave_sync();  // the buffer's previous contents are consumed
#pragma unroll
    for (int k = 0; k < ITEMS; k++) {
        int i = k * 64 + (int)lane;
        // past the end a clamped index: a load inside the bounds test would
        // wait before the next one could issue
        const size_t g = wave_base + i;
        const T e = (T)v[full || g < n ? g : n - 1];
        if (full || g < n) lds[WaveTile<ITEMS>::pad(i)] = e;
    }
    wave_sync();
#pragma unroll
    for (int j = 0; j < ITEMS; j++)
        if (j < nvalid) x[j] = lds[WaveTile<ITEMS>::pad((int)lane * ITEMS + j)];
}

}  // namespace trtc
