// Onesweep LSD radix sort for gfx950: 8-bit digits, one kernel per digit pass.
//
// Order.  The reference sorts by diff = d_u(x) - d_u(min) (radix_sort.cpp:47-66,
// built_in.h:430-458): integers in their natural order, floats with -0.0 and
// +0.0 as one key, every pass a stable split, and "Greater" the stable
// descending order.  The same total order comes from mapping each key to an
// unsigned "sort key" in registers:
//   unsigned: bits          signed: bits ^ SIGN
//   float:    sign ? SIGN - (bits & ~SIGN) : bits | SIGN     (-0.0 -> SIGN = +0.0)
// and, for Greater, complementing it.  Only sort keys are ranked; the original
// key bits are what gets moved, so -0.0 stays -0.0 and keeps its input order
// relative to +0.0.
//
// Pipeline per sort:  fill(zero) -> radix_hist (all digit histograms in one
// read of the keys; the last workgroup turns them into exclusive digit
// offsets) -> one radix_pass per digit.  Each pass tile:
//   1. takes a tile id from an atomic counter (in-order => lookback progress),
//   2. loads BLOCK*ITEMS keys (and values) with coalesced striped loads,
//   3. ranks every key inside its wave with 8 ballots (a wave64 "match"),
//      counting per-wave digit totals in LDS,
//   4. publishes the tile's 256 digit counts and looks back (one thread per
//      digit) for the counts of all earlier tiles,
//   5. stages keys (then values) in LDS in digit order and writes each digit
//      run to its global position with consecutive threads on consecutive
//      addresses.
#pragma once
#include "trtc_prims.hpp"

namespace trtc {

template <int KB> struct KeyWord;
template <> struct KeyWord<1> { typedef uint8_t U; };
template <> struct KeyWord<2> { typedef uint16_t U; };
template <> struct KeyWord<4> { typedef uint32_t U; };
template <> struct KeyWord<8> { typedef uint64_t U; };

template <int VB> struct ValWord;
template <> struct ValWord<0> { typedef uint8_t U; };
template <> struct ValWord<1> { typedef uint8_t U; };
template <> struct ValWord<2> { typedef uint16_t U; };
template <> struct ValWord<4> { typedef uint32_t U; };
template <> struct ValWord<8> { typedef uint64_t U; };

enum KeyKind { kUnsigned = 0, kSigned = 1, kFloat = 2 };

template <class U> TRTC_DEV U sort_key(U bits, int kind, U flip)
{
    constexpr U SIGN = (U)((U)1 << (sizeof(U) * 8 - 1));
    U k;
    if (kind == kUnsigned) k = bits;
    else if (kind == kSigned) k = (U)(bits ^ SIGN);
    else k = (bits & SIGN) ? (U)(SIGN - (U)(bits & (U)~SIGN)) : (U)(bits | SIGN);
    return (U)(k ^ flip);
}

// Status words of the digit lookback: two flag bits on top of a count.
template <class S> struct StatusBits
{
    static constexpr int shift = (int)sizeof(S) * 8 - 2;
    static constexpr S agg = (S)1 << shift;
    static constexpr S incl = (S)2 << shift;
    static constexpr S mask = ((S)1 << shift) - 1;
};

TRTC_DEV unsigned ld_relaxed(const unsigned* p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
TRTC_DEV unsigned long long ld_relaxed(const unsigned long long* p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
TRTC_DEV void st_relaxed(unsigned* p, unsigned v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
TRTC_DEV void st_relaxed(unsigned long long* p, unsigned long long v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// One read of the keys for all digit histograms.  Also clears the passes'
// lookback state (grid-stride), so no separate clearing launch is needed.
template <int BLOCK, int KB>
TRTC_DEV void radix_hist(const RadixHistArgs& a)
{
    typedef typename KeyWord<KB>::U U;
    __shared__ unsigned s_h[KB][256];
    __shared__ bool s_last;
    for (int i = threadIdx.x; i < KB * 256; i += BLOCK) (&s_h[0][0])[i] = 0;
    __syncthreads();

    const U flip = a.descending ? (U)~(U)0 : (U)0;
    const U* keys = (const U*)a.keys;
    const size_t stride = (size_t)gridDim.x * BLOCK;
    for (size_t i = (size_t)blockIdx.x * BLOCK + threadIdx.x; i < a.n; i += stride) {
        U k = sort_key<U>(keys[i], a.kind, flip);
#pragma unroll
        for (int p = 0; p < KB; p++)
            __hip_atomic_fetch_add(&s_h[p][(k >> (8 * p)) & 255], 1u, __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_WORKGROUP);
    }
    for (size_t i = (size_t)blockIdx.x * BLOCK + threadIdx.x; i < a.zero_words; i += stride)
        a.zero_ptr[i] = 0;
    __syncthreads();

    for (int i = threadIdx.x; i < KB * 256; i += BLOCK) {
        unsigned c = (&s_h[0][0])[i];
        if (c) __hip_atomic_fetch_add(a.hist + i, (unsigned long long)c, __ATOMIC_RELAXED,
                                      __HIP_MEMORY_SCOPE_AGENT);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        unsigned t = __hip_atomic_fetch_add(a.done, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        s_last = (t == gridDim.x - 1);
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    }
    __syncthreads();
    if (!s_last) return;
    // Last workgroup: exclusive scan of each 256-bin histogram.
    if (threadIdx.x < 64) {
        for (int p = 0; p < KB; p++) {
            unsigned long long run = 0;
            for (int c = 0; c < 4; c++) {
                int d = c * 64 + (int)threadIdx.x;
                unsigned long long v = ld_relaxed(a.hist + p * 256 + d);
                Plus plus;
                unsigned long long inc = wave_inclusive_scan(v, plus);
                a.digit_start[p * 256 + d] = run + inc - v;
                run += shfl(inc, 63);
            }
        }
    }
}

template <int BLOCK, int ITEMS, int KB, int VB, class S>
TRTC_DEV void radix_pass(const RadixPassArgs& a)
{
    typedef typename KeyWord<KB>::U U;
    typedef typename ValWord<VB>::U V;
    constexpr int WAVES = BLOCK / 64;
    constexpr int TILE = BLOCK * ITEMS;
    constexpr int STAGE_BYTES = TILE * (KB > VB ? KB : VB);
    typedef StatusBits<S> SB;
    typedef S P;  // positions fit the status count field
    static_assert(BLOCK >= 256, "one thread per digit");

    __shared__ unsigned s_wcnt[WAVES][256];       // per-wave digit counts -> tile offsets
    __shared__ P s_base[256];                     // global position of local slot 0 per digit
    __shared__ unsigned s_dscan[WAVES];
    __shared__ unsigned s_tile;
    __shared__ __attribute__((aligned(16))) unsigned char s_stage[STAGE_BYTES];
    U* s_keys = (U*)s_stage;
    V* s_vals = (V*)s_stage;

    for (int i = threadIdx.x; i < WAVES * 256; i += BLOCK) (&s_wcnt[0][0])[i] = 0;
    if (threadIdx.x == 0) s_tile = atomicAdd(a.tile_counter, 1u);
    __syncthreads();

    const size_t tile = s_tile;
    const unsigned wave = threadIdx.x / 64;
    const unsigned lane = lane_id();
    const size_t n = a.n;
    const size_t tile_base = tile * (size_t)TILE;
    const bool full = tile_base + TILE <= n;
    const unsigned valid = full ? (unsigned)TILE : (unsigned)(n - tile_base);
    const U flip = a.descending ? (U)~(U)0 : (U)0;
    const U* kin = (const U*)a.keys_in;
    const V* vin = (const V*)a.vals_in;
    const int wpos0 = (int)wave * 64 * ITEMS;

    // 1. striped loads: item k of this lane sits at tile position wpos0 + k*64 + lane
    U key[ITEMS];
    V val[ITEMS];
    unsigned dig[ITEMS];
#pragma unroll
    for (int k = 0; k < ITEMS; k++) {
        unsigned p = (unsigned)(wpos0 + k * 64) + lane;
        if (full || p < valid) {
            key[k] = kin[tile_base + p];
            if (VB) val[k] = vin[tile_base + p];
        }
    }
#pragma unroll
    for (int k = 0; k < ITEMS; k++) {
        unsigned p = (unsigned)(wpos0 + k * 64) + lane;
        // positions past the end rank as the largest sort key (digit 255) and
        // are never written
        U sk = (full || p < valid) ? sort_key<U>(key[k], a.kind, flip) : (U)~(U)0;
        dig[k] = (unsigned)(sk >> a.shift) & 255u;
    }

    // 2. in-wave ranking: peers = lanes with the same digit in this slot
    const unsigned long long lt_mask = (1ull << lane) - 1ull;
    unsigned rank[ITEMS];
#pragma unroll
    for (int k = 0; k < ITEMS; k++) {
        const unsigned d = dig[k];
        unsigned long long peers = ~0ull;
#pragma unroll
        for (int b = 0; b < 8; b++) {
            const bool bit = (d >> b) & 1u;
            const unsigned long long bal = ballot(bit);
            peers &= bit ? bal : ~bal;
        }
        const unsigned leader = (unsigned)__builtin_ctzll(peers);
        const unsigned cnt = (unsigned)__builtin_popcountll(peers);
        unsigned old = 0;
        if (lane == leader)
            old = __hip_atomic_fetch_add(&s_wcnt[wave][d], cnt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        old = (unsigned)__builtin_amdgcn_ds_bpermute((int)(leader << 2), (int)old);
        rank[k] = old + (unsigned)__builtin_popcountll(peers & lt_mask);
    }
    __syncthreads();

    // 3. per digit (thread d): wave offsets, tile count, tile-local digit start
    unsigned tcount = 0;
    if (threadIdx.x < 256) {
        const unsigned d = threadIdx.x;
        unsigned run = 0;
#pragma unroll
        for (int w = 0; w < WAVES; w++) {
            unsigned c = s_wcnt[w][d];
            s_wcnt[w][d] = run;
            run += c;
        }
        tcount = run;
        // publish this tile's real count (padding keys all sit in digit 255)
        unsigned real = (d == 255 && !full) ? run - (unsigned)(TILE - valid) : run;
        S* st = (S*)a.status;
        if (tile == 0) st_relaxed(st + d, (S)(SB::incl | (S)real));
        else st_relaxed(st + tile * 256 + d, (S)(SB::agg | (S)real));
    }
    // exclusive scan of tcount over the 256 digits (threads 0..255 = 4 waves)
    unsigned dstart = 0;
    if (threadIdx.x < 256) {
        Plus plus;
        unsigned inc = wave_inclusive_scan(tcount, plus);
        if (lane == 63) s_dscan[wave] = inc;
        dstart = inc - tcount;
    }
    __syncthreads();
    if (threadIdx.x < 256) {
        const unsigned d = threadIdx.x;
        for (unsigned w = 0; w < wave; w++) dstart += s_dscan[w];
#pragma unroll
        for (int w = 0; w < WAVES; w++) s_wcnt[w][d] += dstart;

        // 4. lookback over earlier tiles for digit d
        unsigned long long excl = 0;
        if (tile > 0) {
            S* st = (S*)a.status;
            long long j = (long long)tile - 1;
            while (true) {
                // read up to 4 predecessors at once
                S w0 = ld_relaxed(st + (size_t)j * 256 + d);
                S w1 = j >= 1 ? ld_relaxed(st + (size_t)(j - 1) * 256 + d) : (S)0;
                S w2 = j >= 2 ? ld_relaxed(st + (size_t)(j - 2) * 256 + d) : (S)0;
                S w3 = j >= 3 ? ld_relaxed(st + (size_t)(j - 3) * 256 + d) : (S)0;
                S ws[4] = {w0, w1, w2, w3};
                bool done = false;
                int used = 0;
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    if (done || used != q) break;
                    S w = ws[q];
                    if ((w & ~SB::mask) == 0) break;  // not yet published
                    excl += (unsigned long long)(w & SB::mask);
                    used = q + 1;
                    if (w & SB::incl) done = true;
                }
                if (done) break;
                if (used == 0) __builtin_amdgcn_s_sleep(1);
                j -= used;
            }
            unsigned real = (d == 255 && !full) ? tcount - (unsigned)(TILE - valid) : tcount;
            st_relaxed(st + tile * 256 + d, (S)(SB::incl | (S)(excl + real)));
        }
        s_base[d] = (P)(a.digit_start[d] + excl - dstart);
    }
    __syncthreads();

    // 5. stage keys in digit order, then write each run to its global slot
    unsigned lpos[ITEMS];
#pragma unroll
    for (int k = 0; k < ITEMS; k++) {
        lpos[k] = s_wcnt[wave][dig[k]] + rank[k];
        s_keys[lpos[k]] = key[k];
    }
    __syncthreads();
    U* kout = (U*)a.keys_out;
    P gpos[ITEMS];
#pragma unroll
    for (int r = 0; r < ITEMS; r++) {
        unsigned j = (unsigned)(r * BLOCK) + threadIdx.x;
        U kb = s_keys[j];
        U sk = sort_key<U>(kb, a.kind, flip);
        unsigned d = (unsigned)(sk >> a.shift) & 255u;
        gpos[r] = s_base[d] + j;
        // gpos < n holds by construction; the check keeps a logic error from
        // ever writing outside the output
        if ((full || j < valid) && gpos[r] < (P)n) kout[gpos[r]] = kb;
    }
    if (VB) {
        __syncthreads();
#pragma unroll
        for (int k = 0; k < ITEMS; k++) s_vals[lpos[k]] = val[k];
        __syncthreads();
        V* vout = (V*)a.vals_out;
#pragma unroll
        for (int r = 0; r < ITEMS; r++) {
            unsigned j = (unsigned)(r * BLOCK) + threadIdx.x;
            if ((full || j < valid) && gpos[r] < (P)n) vout[gpos[r]] = s_vals[j];
        }
    }
}

}  // namespace trtc
