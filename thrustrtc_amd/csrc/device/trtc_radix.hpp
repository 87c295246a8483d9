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
template <> struct KeyWord<1> { typedef uint8_t U; typedef int8_t S; };
template <> struct KeyWord<2> { typedef uint16_t U; typedef int16_t S; };
template <> struct KeyWord<4> { typedef uint32_t U; typedef int32_t S; };
template <> struct KeyWord<8> { typedef uint64_t U; typedef int64_t S; };

template <int VB> struct ValWord;
template <> struct ValWord<0> { typedef uint8_t U; };
template <> struct ValWord<1> { typedef uint8_t U; };
template <> struct ValWord<2> { typedef uint16_t U; };
template <> struct ValWord<4> { typedef uint32_t U; };
template <> struct ValWord<8> { typedef uint64_t U; };

enum KeyKind { kUnsigned = 0, kSigned = 1, kFloat = 2 };

// Order-preserving unsigned image of a key's bits:
//   unsigned: bits;  signed: bits ^ SIGN;
//   float: non-negative -> bits | SIGN, negative -> SIGN - magnitude = -bits
//   (so -0.0 and +0.0 map to the same image and keep their input order).
// Written without branches, (bits ^ (s | xm)) - s with s the float sign mask,
// so the per-key loops stay one basic block and their LDS reads can be
// issued together.
template <class U> TRTC_DEV U sort_key(U bits, int kind, U flip)
{
    typedef typename KeyWord<(int)sizeof(U)>::S SU;
    constexpr U SIGN = (U)((U)1 << (sizeof(U) * 8 - 1));
    const U fm = kind == kFloat ? (U)~(U)0 : (U)0;
    const U xm = kind == kUnsigned ? (U)0 : SIGN;
    const U s = (U)((U)((SU)bits >> (sizeof(U) * 8 - 1)) & fm);
    return (U)((U)((U)(bits ^ (U)(s | xm)) - s) ^ flip);
}

// sort_key with the key kind fixed at compile time (KIND >= 0: the mapping
// costs at most one xor for unsigned keys), or read at run time (KIND < 0).
template <int KIND, class U> TRTC_DEV U sort_key_k(U bits, int kind, U flip)
{
    constexpr U SIGN = (U)((U)1 << (sizeof(U) * 8 - 1));
    if constexpr (KIND == kUnsigned) return (U)(bits ^ flip);
    else if constexpr (KIND == kSigned) return (U)(bits ^ (U)(SIGN ^ flip));
    else if constexpr (KIND == kFloat) return sort_key<U>(bits, kFloat, flip);
    else return sort_key<U>(bits, kind, flip);
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

// The passes that run read and write alternately the ping-pong and the
// caller's buffers so that the last one writes the caller's buffers: out of
// place (source != result) the first pass reads the source and the parity is
// chosen backwards from the last pass; in place, the first pass must leave
// the source, so an odd number of sorting passes takes one copy, done by the
// launch of a pass whose digit is constant (bit p of copy_mask: pass p copies
// kin -> kout instead of sorting), which exists whenever KB is even.  With
// every digit constant an out-of-place sort copies the source in pass 0's
// launch.  Only a 1-byte-key sort in place has no spare slot: its copy is
// left to radix_copyback (copy_ksrc / copy_kdst).
template <int KB> TRTC_DEV void write_radix_plan(const RadixHistArgs& a, unsigned constant_mask, unsigned rank_lds)
{
    RadixPlan& P = *a.plan;
    P.rank_lds = rank_lds;
    const unsigned long long ks = (unsigned long long)a.keys, vs = (unsigned long long)a.vals_src;
    const unsigned long long kd = (unsigned long long)a.keys_dst, vd = (unsigned long long)a.vals_dst;
    const unsigned long long kt = (unsigned long long)a.keys_tmp, vt = (unsigned long long)a.vals_tmp;
    const bool in_place = ks == kd;
    int runs = 0;
    for (int p = 0; p < KB; p++) runs += (constant_mask >> p) & 1u ? 0 : 1;
    // the slot of the parity copy: the first constant digit (in place with an
    // odd number of sorting passes), pass 0 (out of place, every digit constant)
    int copy_slot = -1;
    if (in_place ? (runs & 1) != 0 : runs == 0) {
        for (int p = 0; p < KB && copy_slot < 0; p++)
            if ((constant_mask >> p) & 1u) copy_slot = p;
    }
    const int ops = runs + (copy_slot >= 0 ? 1 : 0);
    unsigned long long ck = ks, cv = vs;  // current location of the data
    int i = 0;
    P.copy_mask = 0;
    for (int p = 0; p < 8; p++) {
        P.kin[p] = P.kout[p] = P.vin[p] = P.vout[p] = 0;
        if (p >= KB || (((constant_mask >> p) & 1u) && p != copy_slot)) continue;
        // operation number i of `ops`: write the result buffers when the
        // operations left after this one are even in number (out of place), or
        // alternate starting with the ping-pong buffers (in place)
        const bool to_dst = in_place ? (i & 1) != 0 : ((ops - 1 - i) & 1) == 0;
        P.kin[p] = ck;
        P.vin[p] = cv;
        P.kout[p] = to_dst ? kd : kt;
        P.vout[p] = to_dst ? vd : vt;
        if (p == copy_slot) P.copy_mask |= 1u << p;
        ck = P.kout[p];
        cv = P.vout[p];
        i++;
    }
    const bool need_copy = ck != kd;
    P.copy_ksrc = need_copy ? ck : 0;
    P.copy_kdst = need_copy ? kd : 0;
    P.copy_vsrc = need_copy && a.val_bytes ? cv : 0;
    P.copy_vdst = need_copy && a.val_bytes ? vd : 0;
}

// Counts keys[0, n) into bin-major LDS counter columns (hcol = the
// calling lane's column).  The interior that starts on a 128-byte line (a
// wave's load instruction then covers 8 lines, not 9) is read in chunks of
// UNR * 64 16-byte vectors, wave w of nthreads / 64 taking chunks w, w + the
// wave count, ...: each step of a wave reads UNR KiB contiguously (2^28 u32:
// 0.184 against 0.192 ms for chunks a grid apart), its next chunk's loads
// issued before the current one is counted (one chunk at a time leaves the
// read latency-bound).  The keys before the first line boundary and after
// the last 16-byte piece are counted one per thread.  nthreads is a whole
// number of waves and at least 128 / sizeof(key).
template <int KB, int UNR, int C = (KB <= 4 ? 32 : 16)>
TRTC_DEV void hist_count_keys(const gmem_t<typename KeyWord<KB>::U>* keys, size_t n, size_t tid, size_t nthreads,
                              unsigned* hcol, int kind, typename KeyWord<KB>::U flip)
{
    typedef typename KeyWord<KB>::U U;
    typedef unsigned v4u __attribute__((ext_vector_type(4)));
    constexpr int PER_VEC = 16 / KB;
    auto count1 = [&](U key) {
        const U k = sort_key<U>(key, kind, flip);
#pragma unroll
        for (int p = 0; p < KB; p++)
            __hip_atomic_fetch_add(hcol + (p * 256 + ((k >> (8 * p)) & 255)) * C, 1u, __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_WORKGROUP);
    };
    auto countv = [&](const v4u& x) {
        U ks[PER_VEC];
        __builtin_memcpy(ks, &x, 16);
#pragma unroll
        for (int e = 0; e < PER_VEC; e++) count1(ks[e]);
    };
    // keys are aligned to their size, so the head is a whole number of keys
    constexpr size_t LINE_KEYS = 128 / sizeof(U);
    const size_t mis = ((size_t)keys & 127) / sizeof(U);
    size_t head = mis ? LINE_KEYS - mis : 0;
    if (head > n) head = n;
    if (tid < head) count1(keys[tid]);
    const gmem_t<const v4u>* kv = (const gmem_t<const v4u>*)(keys + head);
    const size_t nvec = (n - head) / PER_VEC;
    const v4u zero = {0u, 0u, 0u, 0u};
    constexpr size_t CH = (size_t)UNR * 64;  // vectors per chunk
    const size_t lane = tid & 63, nw = nthreads >> 6;
    v4u cur[UNR];
#pragma unroll
    for (int u = 0; u < UNR; u++) {
        const size_t i = (tid >> 6) * CH + (size_t)u * 64 + lane;
        cur[u] = i < nvec ? __builtin_nontemporal_load(kv + i) : zero;
    }
    for (size_t c = tid >> 6; c * CH < nvec; c += nw) {
        const size_t vn = (c + nw) * CH + lane;
        v4u nxt[UNR];
#pragma unroll
        for (int u = 0; u < UNR; u++) {
            const size_t i = vn + (size_t)u * 64;
            nxt[u] = i < nvec ? __builtin_nontemporal_load(kv + i) : zero;
        }
#pragma unroll
        for (int u = 0; u < UNR; u++)
            if (c * CH + (size_t)u * 64 + lane < nvec) countv(cur[u]);
#pragma unroll
        for (int u = 0; u < UNR; u++) cur[u] = nxt[u];
    }
    const size_t t0 = head + nvec * PER_VEC;  // fewer than PER_VEC keys follow
    if (t0 + tid < n) count1(keys[t0 + tid]);
}

// One read of the keys for all digit histograms (16-byte nontemporal loads).
// Also clears the passes' lookback state (grid-stride), so no separate
// clearing launch is needed.  The last workgroup to finish turns the global
// histograms into exclusive digit offsets.
//
// LDS histograms: C counter columns per bin, lane l counting into column
// l % C, so the 32 lanes served together by an atomic instruction never share
// an address or a bank, whatever their digits.  (One copy per wave put up to
// ~22 lanes on one address for keys whose bits are 1 with probability 1/8:
// 2^28 such keys took 1.39 ms against 0.33 ms for uniform keys; 32/KB copies
// KB*256+1 words apart, 0.29 / 0.45 ms, still spent 71% of their LDS cycles
// on bank conflicts.)
template <int BLOCK, int KB>
TRTC_DEV void radix_hist(const RadixHistArgs& a)
{
    typedef typename KeyWord<KB>::U U;
    static_assert(BLOCK % 64 == 0, "whole waves count the keys");
    // Counters bin-major with C columns: lane l counts into column l % C, so
    // the 32 lanes an LDS instruction serves at once hit 32 different banks
    // whatever their digits (C = 16 for 8-byte keys: 128 KB of LDS at most).
    constexpr int C = KB <= 4 ? 32 : 16;
    constexpr int BINS = KB * 256;
    __shared__ unsigned s_h[BINS * C];
    __shared__ bool s_last;
    for (int i = threadIdx.x; i < BINS * C; i += BLOCK) s_h[i] = 0;
    __syncthreads();

    unsigned* const hcol = s_h + (lane_id() & (unsigned)(C - 1));
    const U flip = a.descending ? (U)~(U)0 : (U)0;
    const size_t stride = (size_t)gridDim.x * BLOCK;
    const size_t gid = (size_t)blockIdx.x * BLOCK + threadIdx.x;
    hist_count_keys<KB, 4>(gmem((const U*)a.keys), a.n, gid, stride, hcol, a.kind, flip);
    {
        // clear the passes' lookback state, 16 bytes per store when aligned
        size_t z0 = 0;
        if ((((size_t)a.zero_ptr) & 15) == 0) {
            typedef unsigned v4z __attribute__((ext_vector_type(4)));
            v4z* zp = (v4z*)a.zero_ptr;
            const size_t nz = a.zero_words / 4;
            const v4z zero = {0u, 0u, 0u, 0u};
            for (size_t i = gid; i < nz; i += stride) zp[i] = zero;
            z0 = nz * 4;
        }
        for (size_t i = z0 + gid; i < a.zero_words; i += stride) a.zero_ptr[i] = 0;
    }
    __syncthreads();

    for (int i = threadIdx.x; i < BINS; i += BLOCK) {
        // row i, columns in a lane-skewed order: neighbouring lanes read
        // different banks at every step
        unsigned c = 0;
#pragma unroll
        for (int j = 0; j < C; j++) c += s_h[i * C + ((j + i) & (C - 1))];
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
    // Last workgroup: exclusive scan of each 256-bin histogram; a pass whose
    // digit has one bin holding all n keys is constant (reference
    // radix_sort.cpp:323-325, 351-353: all-equal keys and constant bits).
    // Also the chance that two keys share the pass's digit, sum (h_d / n)^2:
    // below 1/32 the long-tile pass ranks with LDS words (a uniform digit:
    // 1/256), above it with ballots (the LDS words serialise the lanes of a
    // slot that share a digit: keys that AND three random words, ~0.14).
    if (threadIdx.x < 64) {
        unsigned constant_mask = 0, rank_lds = 0;
        const double inv_n = 1.0 / (double)a.n;
        // every count of the lane first: one round trip instead of one per
        // (pass, quarter) of the scan
        unsigned long long hv[KB * 4];
#pragma unroll
        for (int i = 0; i < KB * 4; i++) hv[i] = ld_relaxed(a.hist + i * 64 + (int)threadIdx.x);
        if (a.self_clear) {
            // every other workgroup has added its counts and its completion:
            // the block is left zero for the stream's next sort
#pragma unroll
            for (int i = 0; i < KB * 4; i++) st_relaxed(a.hist + i * 64 + (int)threadIdx.x, 0ull);
            if (threadIdx.x == 0) st_relaxed(a.done, 0u);
        }
#pragma unroll
        for (int p = 0; p < KB; p++) {
            unsigned long long run = 0;
            bool whole = false;
            double sq = 0.0;
#pragma unroll
            for (int c = 0; c < 4; c++) {
                int d = c * 64 + (int)threadIdx.x;
                unsigned long long v = hv[p * 4 + c];
                whole = whole || v == (unsigned long long)a.n;
                const double f = (double)v * inv_n;
                sq += f * f;
                Plus plus;
                unsigned long long inc = wave_inclusive_scan(v, plus);
                a.digit_start[p * 256 + d] = run + inc - v;
                run += shfl(inc, 63);
            }
            if (ballot(whole)) constant_mask |= 1u << p;
            Plus plus;
            sq = shfl(wave_inclusive_scan(sq, plus), 63);
            if (sq < 1.0 / 32) rank_lds |= 1u << p;
        }
        if (a.plan && threadIdx.x == 0) write_radix_plan<KB>(a, constant_mask, rank_lds);
    }
}

// Barrier that orders LDS only: outstanding global loads and stores stay in
// flight across it (a full __syncthreads would wait for all of them).
TRTC_DEV void lds_barrier()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

template <class T> TRTC_DEV T stream_load(const T* p) { return __builtin_nontemporal_load(p); }

// Buffer resource for `count` 4-byte words from p: raw buffer loads through it
// return 0 at and past the end (no clamped addresses for a partial tile).
TRTC_DEV __amdgpu_buffer_rsrc_t tile_rsrc(const void* p, unsigned count)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, (int)(count * 4u), 0x00020000);
}

// A grid-stride copy of `bytes` bytes, 16 per step when both buffers and the
// length allow, else 4 or 1: the copy a sort's plan asks for (in a pass slot
// whose digit is constant, or after the passes of a 1-byte-key sort).
TRTC_DEV void radix_copy_span(const void* src, void* dst, size_t bytes)
{
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    const size_t gid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t al = (size_t)src | (size_t)dst | bytes;
    if ((al & 15) == 0) {
        typedef unsigned v4u __attribute__((ext_vector_type(4)));
        const gmem_t<const v4u>* s = gmem((const v4u*)src);
        gmem_t<v4u>* d = gmem((v4u*)dst);
        for (size_t i = gid; i < bytes / 16; i += stride) d[i] = s[i];
    } else if ((al & 3) == 0) {
        const gmem_t<const unsigned>* s = gmem((const unsigned*)src);
        gmem_t<unsigned>* d = gmem((unsigned*)dst);
        for (size_t i = gid; i < bytes / 4; i += stride) d[i] = s[i];
    } else {
        const gmem_t<const unsigned char>* s = gmem((const unsigned char*)src);
        gmem_t<unsigned char>* d = gmem((unsigned char*)dst);
        for (size_t i = gid; i < bytes; i += stride) d[i] = s[i];
    }
}

// One digit pass.  Tile = BLOCK*ITEMS keys; item k of a lane sits at tile
// position wave*64*ITEMS + k*64 + lane (coalesced striped loads), and ranks
// are assigned in that position order, which keeps the pass stable.
//   1. load keys and values (values stay in flight until they are staged),
//   2. digits,
//   3. rank inside each wave with 8 ballots per key, counting per-wave digit
//      totals in LDS,
//   4. issue the first lookback reads, sum the per-wave totals into the
//      tile's 256 digit counts and publish them at once ("aggregate"; the
//      first tile publishes them as inclusive prefixes), then do the local
//      digit scan and stage keys and values in LDS in digit order while the
//      lookback reads are in flight,
//   5. finish the lookback (one thread per digit, LBB predecessors per read),
//      publish the inclusive prefix,
//   6. write keys, then values, each digit run to consecutive addresses.
// The tile counts come from the ranking's per-wave totals rather than from a
// separate LDS histogram: 16 fewer LDS atomics per thread (which collide
// heavily on skewed digits) and one barrier fewer, at the price of a later
// publication; 2^28 u32/u32 pass 1.347 -> 1.154 ms (uniform keys), 1.216 ->
// 1.024 ms (low-entropy keys).
// What one workgroup works on in a digit pass: tile `tile` of one key
// sequence (the whole input of a sort, or one segment of a batch).
template <class U, class V, class S> struct RadixJob
{
    const U* kin;
    const V* vin;
    U* kout;
    V* vout;
    size_t n;                             // keys in the sequence
    size_t tile;                          // tile index inside the sequence
    S* status;                            // lookback rows of the sequence; null for a one-tile sequence
    const unsigned long long* dstart64;   // [256] digit offsets of the pass, or
    const unsigned* dstart32;             // [256] (null both: one-tile sequence)
    unsigned lead = 0;                    // keys tile 0 gives up: tile t >= 1 starts at t*TILE - lead
};

// CLAMP_TAIL: the loads of a partial tile read a clamped index (all issued
// together) instead of testing each item.  Both give the same results; which
// one the compiler turns into the faster full-tile code differs by caller:
// clamped, the batched pass over 128 segments of 2^20 u32/u32 pairs runs
// 0.588 -> 0.550 ms and a single 2^27-pair pass 0.577 -> 0.563 ms, but a
// single 2^25-pair pass 0.143 -> 0.153 ms.
template <int BLOCK, int ITEMS, int KB, int VB, class S, int KIND, bool CLAMP_TAIL, class MakeJob>
TRTC_DEV void radix_pass_core(unsigned* tile_counter, const MakeJob& make_job, int shift, int kind, int descending)
{
    typedef typename KeyWord<KB>::U U;
    typedef typename ValWord<VB>::U V;
    constexpr int WAVES = BLOCK / 64;
    constexpr int TILE = BLOCK * ITEMS;
    constexpr int VSTAGE_BYTES = VB ? TILE * VB : 16;
    constexpr int LBB = 4;
    typedef StatusBits<S> SB;
    typedef S P;  // positions fit the status count field
    static_assert(BLOCK >= 256, "one thread per digit");

    __shared__ unsigned s_wcnt[WAVES][256];       // per-wave digit counts -> tile offsets
    __shared__ P s_base[256];                     // global position of local slot 0 per digit
    __shared__ unsigned s_dscan[4];
    __shared__ unsigned s_tile;
    // keys and values are staged in separate buffers, both before the
    // lookback, so the write-out needs no further barrier
    __shared__ __attribute__((aligned(16))) unsigned char s_kstage[TILE * KB];
    __shared__ __attribute__((aligned(16))) unsigned char s_vstage[VSTAGE_BYTES];
    U* s_keys = (U*)s_kstage;
    V* s_vals = (V*)s_vstage;

    for (int i = threadIdx.x; i < WAVES * 256; i += BLOCK) (&s_wcnt[0][0])[i] = 0;
    if (threadIdx.x == 0) s_tile = atomicAdd(tile_counter, 1u);
    lds_barrier();

    const RadixJob<U, V, S> job = make_job((unsigned)s_tile);
    const size_t tile = job.tile;
    const unsigned wave = threadIdx.x / 64;
    const unsigned lane = lane_id();
    const size_t n = job.n;
    // tile 0 holds TILE - lead keys, so that with lead = the sequence's
    // offset from a 128-byte line in keys every later tile starts on a line
    // (a 256-byte load instruction then covers two lines, not three); a tile
    // past the end (the tile count allows for the largest lead of any pass)
    // holds no keys and publishes zero counts
    const size_t tb0 = tile ? tile * (size_t)TILE - job.lead : 0;
    const size_t te0 = (tile + 1) * (size_t)TILE - job.lead;
    const size_t tile_base = tb0 < n ? tb0 : n;
    const size_t tile_end = te0 < n ? te0 : n;
    const unsigned valid = (unsigned)(tile_end - tile_base);
    const bool full = valid == (unsigned)TILE;
    const U flip = descending ? (U)~(U)0 : (U)0;
    const gmem_t<const U>* kin = gmem(job.kin + tile_base);
    const gmem_t<const V>* vin = gmem(job.vin + tile_base);
    const int wpos0 = (int)wave * 64 * ITEMS;

    // 1. loads
    U key[ITEMS];
    V val[ITEMS];
    unsigned dig[ITEMS];
    if (full) {
#pragma unroll
        for (int k = 0; k < ITEMS; k++) key[k] = stream_load(kin + (wpos0 + k * 64) + lane);
        if (VB) {
#pragma unroll
            for (int k = 0; k < ITEMS; k++) val[k] = stream_load(vin + (wpos0 + k * 64) + lane);
        }
    } else {
#pragma unroll
        for (int k = 0; k < ITEMS; k++) {
            unsigned p = (unsigned)(wpos0 + k * 64) + lane;
            if (CLAMP_TAIL) {
                if (valid) {
                    const unsigned q = p < valid ? p : valid - 1;
                    key[k] = kin[q];
                    if (VB) val[k] = vin[q];
                } else {
                    key[k] = 0;
                    if (VB) val[k] = 0;
                }
            } else if (p < valid) {
                key[k] = kin[p];
                if (VB) val[k] = vin[p];
            }
        }
    }
    // 2. digits (positions past the end rank as the largest sort key, digit
    //    255, and are never written)
#pragma unroll
    for (int k = 0; k < ITEMS; k++) {
        unsigned p = (unsigned)(wpos0 + k * 64) + lane;
        U sk = (full || p < valid) ? sort_key_k<KIND, U>(key[k], kind, flip) : (U)~(U)0;
        dig[k] = (unsigned)(sk >> shift) & 255u;
    }
    S* st = job.status;
    const bool head = tile == 0;
    unsigned real = 0;

    // 3. in-wave ranking: peers = lanes with the same digit in this slot;
    //    the rank is packed above the digit (rank < 64*ITEMS).  The lowest
    //    peer adds the peer count to the wave's digit counter and every peer
    //    reads the counter back right after: a wave's LDS operations execute
    //    in issue order, so the read sees this slot's add and no later one.
    {
#pragma unroll
        for (int k = 0; k < ITEMS; k++) {
            const unsigned d = dig[k];
            // lanes whose digit equals this lane's in every bit: per bit, keep
            // the lanes where the ballot agrees with this lane's bit spread to
            // all 32 bits (m = 0 or ~0): acc & ~(ballot ^ m), one v_bitop3
            // (truth table 0x90) per half.  The empty asm hides that m is a bit
            // field, so the ballot stays one compare on m rather than a second
            // shift-and-compare of the digit.  All eight ballots come first, so
            // no bitop3 waits on the compare just before it.
            int m[8];
            unsigned long long bal[8];
#pragma unroll
            for (int b = 0; b < 8; b++) {
                m[b] = __builtin_amdgcn_sbfe((int)d, b, 1);
                __asm__("" : "+v"(m[b]));
                bal[b] = ballot(m[b] != 0);
            }
            unsigned plo = ~0u, phi = ~0u;
#pragma unroll
            for (int b = 0; b < 8; b++) {
                plo = __builtin_amdgcn_bitop3_b32(plo, (unsigned)bal[b], (unsigned)m[b], 0x90);
                phi = __builtin_amdgcn_bitop3_b32(phi, (unsigned)(bal[b] >> 32), (unsigned)m[b], 0x90);
            }
            // peers in lower lanes, and all peers
            const unsigned below = __builtin_amdgcn_mbcnt_hi(phi, __builtin_amdgcn_mbcnt_lo(plo, 0u));
            const unsigned cnt = (unsigned)__builtin_popcount(plo) + (unsigned)__builtin_popcount(phi);
            unsigned* ctr = &s_wcnt[wave][d];
            if (below == 0u) __hip_atomic_fetch_add(ctr, cnt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            const unsigned after = __hip_atomic_load(ctr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            dig[k] |= (after - cnt + below) << 8;
        }
    }
    lds_barrier();

    // 4. first lookback reads in flight; local digit scan
    S w[LBB];
    unsigned dstart = 0;
    if (threadIdx.x < 256) {
        const unsigned d = threadIdx.x;
#pragma unroll
        for (int b = 0; b < LBB; b++) {
            long long j = (long long)tile - 1 - b;
            w[b] = (!head && j >= 0) ? ld_relaxed(st + (size_t)j * 256 + d) : (S)0;
        }
        unsigned run = 0;
#pragma unroll
        for (int ww = 0; ww < WAVES; ww++) {
            unsigned c = s_wcnt[ww][d];
            s_wcnt[ww][d] = run;
            run += c;
        }
        // the tile's digit count (the column sum of the per-wave counts of
        // the ranking) is published as soon as it is known
        real = (d == 255 && !full) ? run - (unsigned)(TILE - valid) : run;
        if (st) st_relaxed(st + tile * 256 + d, (S)((head ? SB::incl : SB::agg) | (S)real));
        Plus plus;
        unsigned inc = wave_inclusive_scan(run, plus);
        if (lane == 63) s_dscan[wave] = inc;
        dstart = inc - run;
    }
    lds_barrier();
    if (threadIdx.x < 256) {
        const unsigned d = threadIdx.x;
        for (unsigned ww = 0; ww < wave; ww++) dstart += s_dscan[ww];
#pragma unroll
        for (int ww = 0; ww < WAVES; ww++) s_wcnt[ww][d] += dstart;
    }
    lds_barrier();
#pragma unroll
    for (int k = 0; k < ITEMS; k++) {
        const unsigned lpos = s_wcnt[wave][dig[k] & 255u] + (dig[k] >> 8);
        s_keys[lpos] = key[k];
        if (VB) s_vals[lpos] = val[k];
    }

    // 5. finish the lookback for digit d
    if (threadIdx.x < 256) {
        const unsigned d = threadIdx.x;
        unsigned long long excl = 0;
        if (!head) {
            long long j = (long long)tile - 1;
            bool first = true;
            while (true) {
                S ws[LBB];
#pragma unroll
                for (int b = 0; b < LBB; b++)
                    ws[b] = first ? w[b] : ((j - b >= 0) ? ld_relaxed(st + (size_t)(j - b) * 256 + d) : (S)0);
                first = false;
                bool done = false;
                int used = 0;
#pragma unroll
                for (int b = 0; b < LBB; b++) {
                    if (done || used != b) break;
                    S x = ws[b];
                    if ((x & ~SB::mask) == 0) break;  // not yet published
                    excl += (unsigned long long)(x & SB::mask);
                    used = b + 1;
                    if (x & SB::incl) done = true;
                }
                if (done) break;
                if (used == 0) __builtin_amdgcn_s_sleep(1);
                j -= used;
            }
            st_relaxed(st + tile * 256 + d, (S)(SB::incl | (S)(excl + real)));
        }
        // digit offset of the sequence: from the histogram pre-pass, or for a
        // sequence of one tile, the tile's own digit scan
        const unsigned long long ds = job.dstart64 ? job.dstart64[d]
                                    : (job.dstart32 ? (unsigned long long)job.dstart32[d] : (unsigned long long)dstart);
        s_base[d] = (P)(ds + excl - dstart);
    }
    lds_barrier();

    // 6. write each digit run to its global slot.  All LDS reads are issued
    //    before the first store (separate loops), so they overlap instead of
    //    waiting one key at a time.
    gmem_t<U>* kout = gmem(job.kout);
    U kb[ITEMS];
    P gpos[ITEMS];
#pragma unroll
    for (int r = 0; r < ITEMS; r++) kb[r] = s_keys[(unsigned)(r * BLOCK) + threadIdx.x];
#pragma unroll
    for (int r = 0; r < ITEMS; r++) {
        const unsigned d = (unsigned)(sort_key_k<KIND, U>(kb[r], kind, flip) >> shift) & 255u;
        gpos[r] = s_base[d] + (unsigned)(r * BLOCK) + threadIdx.x;
    }
    // gpos < n holds by construction; the check keeps a logic error from
    // ever writing outside the output
#pragma unroll
    for (int r = 0; r < ITEMS; r++) {
        const unsigned j = (unsigned)(r * BLOCK) + threadIdx.x;
        if ((full || j < valid) && gpos[r] < (P)n) kout[gpos[r]] = kb[r];
    }
    if (VB) {
        gmem_t<V>* vout = gmem(job.vout);
        V vb[ITEMS];
#pragma unroll
        for (int r = 0; r < ITEMS; r++) vb[r] = s_vals[(unsigned)(r * BLOCK) + threadIdx.x];
#pragma unroll
        for (int r = 0; r < ITEMS; r++) {
            const unsigned j = (unsigned)(r * BLOCK) + threadIdx.x;
            if ((full || j < valid) && gpos[r] < (P)n) vout[gpos[r]] = vb[r];
        }
    }
}


// One digit pass over the whole input.  With a plan, the buffers come from
// it and a pass over a constant digit returns at once.
template <int BLOCK, int ITEMS, int KB, int VB, class S, int KIND = -1>
TRTC_DEV void radix_pass(const RadixPassArgs& a)
{
    typedef typename KeyWord<KB>::U U;
    typedef typename ValWord<VB>::U V;
    const void* kin = a.keys_in;
    const void* vin = a.vals_in;
    void* kout = a.keys_out;
    void* vout = a.vals_out;
    if (a.plan) {
        const RadixPlan& P = *a.plan;
        if (!P.kin[a.pass]) return;
        kin = (const void*)P.kin[a.pass];
        vin = (const void*)P.vin[a.pass];
        kout = (void*)P.kout[a.pass];
        vout = (void*)P.vout[a.pass];
        if ((P.copy_mask >> a.pass) & 1u) {
            radix_copy_span(kin, kout, a.n * (size_t)KB);
            if (VB) radix_copy_span(vin, vout, a.n * (size_t)VB);
            return;
        }
    }
    auto make = [&](unsigned t) {
        RadixJob<U, V, S> j;
        j.kin = (const U*)kin;
        j.vin = (const V*)vin;
        j.kout = (U*)kout;
        j.vout = (V*)vout;
        j.n = a.n;
        j.tile = t;
        j.status = (S*)a.status;
        j.dstart64 = a.digit_start;
        j.dstart32 = nullptr;
        return j;
    };
    radix_pass_core<BLOCK, ITEMS, KB, VB, S, KIND, false>(a.tile_counter, make, a.shift, a.kind, a.descending);
}

// What one workgroup of the long-tile pass works on: tile `tile` of one key
// sequence (the whole input of a sort, or one segment of a batch).
template <class S> struct LongJob
{
    const unsigned* kin;
    const unsigned* vin;
    unsigned* kout;
    unsigned* vout;
    size_t n;                             // keys in the sequence
    size_t tile;                          // tile index inside the sequence
    S* status;                            // lookback rows of the sequence; null for a one-tile sequence
    const unsigned long long* dstart64;   // [256] digit offsets of the pass, or
    const unsigned* dstart32;             // [256] (null both: one-tile sequence)
    unsigned lead;                        // keys tile 0 gives up: tile t >= 1 starts at t*TILE - lead
    bool lds_rank;                        // rank with LDS count|mask words (else ballots)
};

// Long-tile digit pass for large sorts of 4-byte keys with 4-byte or no
// values (single sorts, and the large segments of a batch): BLOCK x ITEMS =
// 1024 x 32 keys per tile (128-key digit runs on average, so
// fewer partially written 128-byte lines at run ends than the 64-key runs of
// 16384-key tiles), ranked in 32-lane units:
//   * unit u = threadIdx.x / 32 owns the 32*ITEMS consecutive tile positions
//     from u*32*ITEMS; item k of lane l (l = lane % 32) sits at u*32*ITEMS +
//     k*32 + l, so a load instruction covers two whole lines, one per unit;
//   * each unit has 256 LDS words {count:32 | peer mask:32}.  Flat digits
//     (the plan's rank_lds bit): every lane adds (1 << 32) | (1 << l) to its
//     digit's word and reads it back; the mask gives the lanes of this slot
//     with the same digit, the count all such keys so far: rank = count -
//     popcount(mask) + popcount(mask below l); the lowest such lane clears the
//     mask.  A wave's LDS operations execute in issue order, so the read sees
//     this slot's adds and no later one.  Skewed digits: the 8-ballot match of
//     radix_pass_core on the unit's half of the wave, the lowest peer adding
//     the peer count to the word's count (LDS adds to one address serialise,
//     and with ~a third of the keys in one bin the ballots are cheaper).
//   * ranks are 16-bit, two per register, and digits are recomputed from the
//     keys where needed, so 32 keys fit 128 VGPRs with the keys in registers;
//     the rank words live in the staging buffer until the local positions
//     are known.
// Tiles go to XCDs in runs of 8 consecutive tiles (per-XCD counters, XCD =
// blockIdx.x mod 8, the order in which the hardware spreads workgroups): the
// partial lines that neighbouring tiles write at the two ends of a digit run
// then meet in one L2 instead of reaching HBM separately.  2^28 u32/u32, one
// pass, uniform keys: 1.105 ms (1024 x 24, wave ranking, arrival order) ->
// 0.928 ms; low-entropy keys 1.080 -> 1.009 ms.
//   1. take a tile (the workgroup loops until no tile is left; the grid is a
//      multiple of 8 so that every virtual XCD has workgroups),
//   2. load keys, rank in units,
//   3. publish the tile's digit counts, local digit scan, first lookback reads,
//   4. write keys to LDS at their digit-sorted slot; load the values,
//   5. finish the lookback; write the keys out,
//   6. stage the values in the same buffer and write them out.
// Stable: a unit's ranks follow its positions (slot, then lane) and units are
// scanned in position order.

template <int BLOCK, int ITEMS, int VB, class S, int KIND, class MakeJob>
TRTC_DEV void radix_pass_long_core(unsigned* tile_counter, size_t num_tiles, const MakeJob& make_job, int shift,
                                   int kind, int descending)
{
    typedef unsigned U;
    typedef unsigned V;
    typedef S P;
    constexpr int UNITS = BLOCK / 32;
    constexpr int TILE = BLOCK * ITEMS;
    constexpr int UITEMS = 32 * ITEMS;
    constexpr int LBB = 4;
    constexpr unsigned XRUN = 8;  // consecutive tiles per XCD
    typedef StatusBits<S> SB;
    static_assert(VB == 0 || VB == 4, "4-byte values or none");
    static_assert(TILE * 4 >= UNITS * 256 * 8, "rank words fit the staging buffer");
    static_assert(TILE <= 65536 && ITEMS % 4 == 0, "16-bit local positions, digits packed by four");

    __shared__ __attribute__((aligned(16))) unsigned s_stage[TILE];
    __shared__ unsigned short s_off[UNITS][256];  // unit's first local slot per digit
    __shared__ P s_base[256];
    __shared__ unsigned s_dscan[BLOCK / 64];
    __shared__ unsigned s_real[256];              // the tile's digit counts
    __shared__ unsigned s_dstart[256];            // the tile's local digit offsets
    __shared__ unsigned s_tile;
    unsigned long long* const s_rank = (unsigned long long*)s_stage;

    // the next tile of this workgroup: runs of XRUN consecutive tiles go to
    // virtual XCD blockIdx.x mod 8 (thread 0; read by all after a barrier)
    auto claim = [&]() {
        const unsigned x = blockIdx.x & 7u;
        const unsigned k = atomicAdd(tile_counter + x, 1u);
        s_tile = (k / XRUN) * (8u * XRUN) + x * XRUN + k % XRUN;
    };
    // Tiles until none is left: a grid of one workgroup per CU loops over them
    // (no workgroup launch between tiles), a grid of one per tile takes one
    // each.  Each tile claims its successor while it stages its values (keys
    // only: once its lookback is done), so that the claim's round trip is off
    // the critical path; a claimed tile only waits for its workgroup's current
    // tile, which is lower, so every wait still points to a lower tile.  The
    // thread index is made opaque per tile so that per-item addresses are
    // recomputed instead of held across the loop.
    if (threadIdx.x == 0) claim();
    for (;;) {
        unsigned tid = threadIdx.x;
        __asm__ volatile("" : "+v"(tid));
        {
            typedef unsigned v4u __attribute__((ext_vector_type(4)));
            v4u* z = (v4u*)s_rank;
            const v4u zero = {0u, 0u, 0u, 0u};
            for (int i = tid; i < UNITS * 256 / 2; i += BLOCK) z[i] = zero;
        }
        lds_barrier();
        if ((size_t)s_tile >= num_tiles) break;  // no tiles left (or padding of the grid)

        const LongJob<S> job = make_job((unsigned)s_tile);
        const size_t tile = job.tile;
        const size_t n = job.n;
        // tile 0 holds TILE - lead keys, so that with lead = the sequence's
        // offset from a 128-byte line in keys every later tile starts on a
        // line; a tile past the end (room for the largest lead of any pass)
        // holds no keys and publishes zero counts
        const size_t tb0 = tile ? tile * (size_t)TILE - job.lead : 0;
        const size_t te0 = (tile + 1) * (size_t)TILE - job.lead;
        const size_t tile_base = tb0 < n ? tb0 : n;
        const size_t tile_end = te0 < n ? te0 : n;
        const unsigned valid = (unsigned)(tile_end - tile_base);
        const bool full = valid == (unsigned)TILE;
        const unsigned wave = tid / 64;
        const unsigned lane = lane_id();
        const unsigned l32 = lane & 31u;
        const unsigned unit = tid >> 5;
        const gmem_t<const U>* kin = gmem(job.kin + tile_base);
        const gmem_t<const V>* vin = gmem(job.vin + tile_base);
        const int upos0 = (int)unit * UITEMS;
        const U flip = descending ? ~0u : 0u;
        S* st = job.status;
        const bool head = tile == 0;
        const bool lds_rank = job.lds_rank;
        // 1. keys (registers)
        // Full tiles: one address per lane and constant offsets (k * 128 bytes)
        // in the load instructions.  The partial last tile: buffer loads, which
        // return 0 past its end, so that both paths issue the same 32 loads and
        // the ranking can wait for them one at a time (clamped per-item indices
        // made the compiler merge both paths into one set of loads with 32
        // separate 64-bit addresses; loads under a branch made it wait for all
        // 32 before the first rank).
        U key[ITEMS];
        unsigned in_range = ~0u;  // bit k: item k lies inside the input
        if (full) {
            const gmem_t<const U>* const kl = kin + (upos0 + (int)l32);
    #pragma unroll
            for (int k = 0; k < ITEMS; k++) key[k] = kl[k * 32];
        } else {
            const __amdgpu_buffer_rsrc_t rs = tile_rsrc(job.kin + tile_base, valid);
            const unsigned off = ((unsigned)upos0 + l32) * 4u;
            in_range = 0;
    #pragma unroll
            for (int k = 0; k < ITEMS; k++) {
                key[k] = __builtin_amdgcn_raw_buffer_load_b32(rs, (int)(off + k * 128), 0, 0);
                in_range |= ((unsigned)(upos0 + k * 32) + l32 < valid ? 1u : 0u) << k;
            }
        }
        // digit of item k, recomputed from the key wherever it is needed
        // (positions past the end rank as digit 255 and are never written)
        auto digit_of = [&](int k) -> unsigned {
            U sk = sort_key_k<KIND, U>(key[k], kind, flip);
            if (!full && !((in_range >> k) & 1u)) sk = ~0u;
            return (sk >> shift) & 255u;
        };
        unsigned rk2[ITEMS / 2];  // ranks inside the unit, two 16-bit halves per register
        // the asm keeps each rank's arithmetic inside its slot rather than holding
        // every slot's LDS words until the end of the loop
        auto put_rank = [&](int k, unsigned r) {
            if (k & 1) rk2[k / 2] |= r << 16;
            else rk2[k / 2] = r;
            __asm__ volatile("" : "+v"(rk2[k / 2]));
        };

        // 2. ranks inside the unit
        unsigned long long* const words = s_rank + unit * 256;
        const unsigned below_mask = (1u << l32) - 1u;
        if (lds_rank) {
            const unsigned long long add = (1ull << 32) | (1ull << l32);
    #pragma unroll
            for (int k = 0; k < ITEMS; k++) {
                unsigned long long* const w = words + digit_of(k);
                __hip_atomic_fetch_add(w, add, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                const unsigned long long x = __hip_atomic_load(w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                const unsigned mask = (unsigned)x, count = (unsigned)(x >> 32);
                const unsigned below = (unsigned)__builtin_popcount(mask & below_mask);
                if (below == 0u) __hip_atomic_store((unsigned*)w, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                put_rank(k, count - (unsigned)__builtin_popcount(mask) + below);
                __asm__ volatile("" ::: "memory");
            }
        } else {
            const bool upper = lane >= 32u;
    #pragma unroll
            for (int k = 0; k < ITEMS; k++) {
                const unsigned d = digit_of(k);
                int m[8];
                unsigned long long bal[8];
    #pragma unroll
                for (int b = 0; b < 8; b++) {
                    m[b] = __builtin_amdgcn_sbfe((int)d, b, 1);
                    __asm__("" : "+v"(m[b]));
                    bal[b] = ballot(m[b] != 0);
                }
                unsigned plo = ~0u, phi = ~0u;
    #pragma unroll
                for (int b = 0; b < 8; b++) {
                    plo = __builtin_amdgcn_bitop3_b32(plo, (unsigned)bal[b], (unsigned)m[b], 0x90);
                    phi = __builtin_amdgcn_bitop3_b32(phi, (unsigned)(bal[b] >> 32), (unsigned)m[b], 0x90);
                }
                const unsigned own = upper ? phi : plo;  // peers in this lane's unit
                const unsigned below = (unsigned)__builtin_popcount(own & below_mask);
                const unsigned cnt = (unsigned)__builtin_popcount(own);
                unsigned* const ctr = (unsigned*)(words + d) + 1;
                if (below == 0u) __hip_atomic_fetch_add(ctr, cnt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                const unsigned after = __hip_atomic_load(ctr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                put_rank(k, after - cnt + below);
                if (k % 8 == 7) __builtin_amdgcn_sched_barrier(0);
            }
        }
        // the keys look new to the compiler here, so later digits are recomputed
        // from them rather than kept live from the ranking
    #pragma unroll
        for (int k = 0; k < ITEMS; k++) __asm__ volatile("" : "+v"(key[k]));
        lds_barrier();

        // 3. counts and local digit offsets, four threads per digit (adjacent
        //    lanes, eight units each); the first lookback reads in flight (one
        //    thread per digit)
        static_assert(BLOCK == 1024 && UNITS == 32, "four threads per digit, eight units each");
        S w[LBB];
        if (tid < 256) {
#pragma unroll
            for (int b = 0; b < LBB; b++) {
                long long j = (long long)tile - 1 - b;
                w[b] = (!head && j >= 0) ? ld_relaxed(st + (size_t)j * 256 + tid) : (S)0;
            }
        }
        {
            const unsigned d = tid >> 2, q = tid & 3u;
            unsigned pre[8];  // offsets inside this thread's eight units
            unsigned part = 0;
#pragma unroll
            for (int i = 0; i < 8; i++) {
                pre[i] = part;
                part += ((const unsigned*)(s_rank + (8 * q + i) * 256 + d))[1];
            }
            // exclusive prefix over the digit's four threads, and the digit's
            // count (the tile histogram) from its fourth thread
            unsigned incl = part;
            {
                const unsigned y1 = (unsigned)__shfl_up((int)incl, 1, 4);
                if (q >= 1) incl += y1;
                const unsigned y2 = (unsigned)__shfl_up((int)incl, 2, 4);
                if (q >= 2) incl += y2;
            }
            const unsigned total = (unsigned)__shfl((int)incl, (int)(lane | 3u), 64);
            const unsigned excl_q = incl - part;
            // published at once ("aggregate"; the first tile publishes its
            // counts as inclusive prefixes)
            const unsigned real = (d == 255 && !full) ? total - (unsigned)(TILE - valid) : total;
            if (q == 0) {
                if (st) st_relaxed(st + tile * 256 + d, (S)((head ? SB::incl : SB::agg) | (S)real));
                s_real[d] = real;
            }
            // local digit scan: 16 digits per wave (their totals in the lanes
            // with q == 0), then across the waves
            Plus plus;
            const unsigned own = q == 0 ? total : 0u;
            const unsigned wincl = wave_inclusive_scan(own, plus);
            if (lane == 63) s_dscan[wave] = wincl;
            lds_barrier();
            // the digit's exclusive offset inside the wave, from its q == 0 lane
            unsigned ds = (unsigned)__shfl((int)(wincl - own), (int)(lane & ~3u), 64);
            for (unsigned ww = 0; ww < wave; ww++) ds += s_dscan[ww];
            if (q == 0) s_dstart[d] = ds;
#pragma unroll
            for (int i = 0; i < 8; i++) s_off[8 * q + i][d] = (unsigned short)(pre[i] + excl_q + ds);
        }
        lds_barrier();
        const unsigned dstart = tid < 256 ? s_dstart[tid] : 0u;
        const unsigned real = tid < 256 ? s_real[tid] : 0u;

        // 4. local positions replace the ranks; keys into the staging buffer (the
        //    rank words are dead)
        unsigned* const lp2 = rk2;
    #pragma unroll
        for (int k = 0; k < ITEMS; k++) lp2[k / 2] += (unsigned)s_off[unit][digit_of(k)] << (16 * (k & 1));
    #pragma unroll
        for (int k = 0; k < ITEMS; k++) s_stage[(lp2[k / 2] >> (16 * (k & 1))) & 0xFFFFu] = key[k];
        // the value staging recomputes its addresses from the packed positions
    #pragma unroll
        for (int i = 0; i < ITEMS / 2; i++) __asm__ volatile("" : "+v"(lp2[i]));
        __asm__ volatile("" ::: "memory");

        // values in flight during the lookback (the load positions recomputed, not
        // kept live from the key loads)
        V val[VB ? ITEMS : 1];
        if (VB) {
            unsigned vpos0 = (unsigned)upos0 + l32;
            __asm__ volatile("" : "+v"(vpos0));
            if (full) {
                const gmem_t<const V>* const vl = vin + vpos0;
    #pragma unroll
                for (int k = 0; k < ITEMS; k++) val[k] = vl[k * 32];
            } else {
                const __amdgpu_buffer_rsrc_t rs = tile_rsrc(job.vin + tile_base, valid);
    #pragma unroll
                for (int k = 0; k < ITEMS; k++) val[k] = __builtin_amdgcn_raw_buffer_load_b32(rs, (int)(vpos0 * 4u + k * 128), 0, 0);
            }
        }
        // 5. finish the lookback for digit d
        if (tid < 256) {
            const unsigned d = tid;
            unsigned long long excl = 0;
            if (!head) {
                long long j = (long long)tile - 1;
                bool first = true;
                while (true) {
                    S ws[LBB];
    #pragma unroll
                    for (int b = 0; b < LBB; b++)
                        ws[b] = first ? w[b] : ((j - b >= 0) ? ld_relaxed(st + (size_t)(j - b) * 256 + d) : (S)0);
                    first = false;
                    bool done = false;
                    int used = 0;
    #pragma unroll
                    for (int b = 0; b < LBB; b++) {
                        if (done || used != b) break;
                        S x = ws[b];
                        if ((x & ~SB::mask) == 0) break;  // not yet published
                        excl += (unsigned long long)(x & SB::mask);
                        used = b + 1;
                        if (x & SB::incl) done = true;
                    }
                    if (done) break;
                    if (used == 0) __builtin_amdgcn_s_sleep(1);
                    j -= used;
                }
                st_relaxed(st + tile * 256 + d, (S)(SB::incl | (S)(excl + real)));
            }
            // digit offset of the sequence: from the histogram pre-pass, or for a
            // sequence of one tile, the tile's own digit scan
            const unsigned long long ds = job.dstart64 ? job.dstart64[d]
                                        : (job.dstart32 ? (unsigned long long)job.dstart32[d] : (unsigned long long)dstart);
            s_base[d] = (P)(ds + excl - dstart);
        }
        if (!VB && tid == 0) claim();  // keys only: the next tile, while this one writes out
        lds_barrier();

        // keys out (digits remembered, four per register)
        gmem_t<U>* kout = gmem(job.kout);
        gmem_t<V>* vout = gmem(job.vout);
        unsigned dg4[ITEMS / 4];
        // groups of WB items: their LDS reads issue together, then their stores
        // (one item at a time, each store's branch kept the next item's LDS
        // reads behind it)
        constexpr int WB = 8;
    #pragma unroll
        for (int r0 = 0; r0 < ITEMS; r0 += WB) {
            U kb[WB];
            P gp[WB];
    #pragma unroll
            for (int i = 0; i < WB; i++) kb[i] = s_stage[(unsigned)((r0 + i) * BLOCK) + tid];
    #pragma unroll
            for (int i = 0; i < WB; i++) {
                const int r = r0 + i;
                const unsigned d = (sort_key_k<KIND, U>(kb[i], kind, flip) >> shift) & 255u;
                if (r % 4 == 0) dg4[r / 4] = d;
                else dg4[r / 4] |= d << (8 * (r % 4));
                gp[i] = s_base[d] + (P)((unsigned)(r * BLOCK) + tid);
            }
    #pragma unroll
            for (int i = 0; i < WB; i++) {
                const unsigned j = (unsigned)((r0 + i) * BLOCK) + tid;
                // gpos < n by construction; the check keeps a logic error from writing outside the output
                if ((full || j < valid) && gp[i] < (P)n) kout[gp[i]] = kb[i];
            }
            __builtin_amdgcn_sched_barrier(0);
        }
        if (!VB) {
            lds_barrier();  // the staging buffer is read until here
            continue;
        }
        // 6. values staged in the same buffer and written out
        lds_barrier();
        unsigned spos0 = (unsigned)upos0 + l32;
        __asm__ volatile("" : "+v"(spos0));
    #pragma unroll
        for (int k = 0; k < ITEMS; k++) {
            unsigned p = spos0 + (unsigned)(k * 32);
            if (full || p < valid) s_stage[(lp2[k / 2] >> (16 * (k & 1))) & 0xFFFFu] = val[k];
        }
        if (tid == 0) claim();  // the next tile, while this one writes its values out
        lds_barrier();
        // slot indices and digit addresses recomputed here, not kept live from the key write-out
        unsigned tid2 = tid;
        __asm__ volatile("" : "+v"(tid2));
    #pragma unroll
        for (int i = 0; i < ITEMS / 4; i++) __asm__ volatile("" : "+v"(dg4[i]));
    #pragma unroll
        for (int r0 = 0; r0 < ITEMS; r0 += WB) {
            V vb[WB];
            P gp[WB];
    #pragma unroll
            for (int i = 0; i < WB; i++) vb[i] = s_stage[(unsigned)((r0 + i) * BLOCK) + tid2];
    #pragma unroll
            for (int i = 0; i < WB; i++) {
                const int r = r0 + i;
                const unsigned d = (dg4[r / 4] >> (8 * (r % 4))) & 255u;
                gp[i] = s_base[d] + (P)((unsigned)(r * BLOCK) + tid2);
            }
    #pragma unroll
            for (int i = 0; i < WB; i++) {
                const unsigned j = (unsigned)((r0 + i) * BLOCK) + tid2;
                if ((full || j < valid) && gp[i] < (P)n) vout[gp[i]] = vb[i];
            }
            __builtin_amdgcn_sched_barrier(0);
        }

        lds_barrier();  // the staging buffer is read until here
    }
}

// One long-tile digit pass over the whole input of a sort (buffers and
// ranking mode from the plan; a pass over a constant digit returns at once).
template <int BLOCK, int ITEMS, int VB, class S, int KIND>
TRTC_DEV void radix_pass_long(const RadixPassArgs& a)
{
    const void* kin = a.keys_in;
    const void* vin = a.vals_in;
    void* kout = a.keys_out;
    void* vout = a.vals_out;
    bool lds_rank = false;
    if (a.plan) {
        const RadixPlan& PL = *a.plan;
        if (!PL.kin[a.pass]) return;
        kin = (const void*)PL.kin[a.pass];
        vin = (const void*)PL.vin[a.pass];
        kout = (void*)PL.kout[a.pass];
        vout = (void*)PL.vout[a.pass];
        if ((PL.copy_mask >> a.pass) & 1u) {
            radix_copy_span(kin, kout, a.n * 4);
            if (VB) radix_copy_span(vin, vout, a.n * (size_t)VB);
            return;
        }
        lds_rank = (PL.rank_lds >> a.pass) & 1u;
    }
    constexpr size_t TILE = (size_t)BLOCK * ITEMS;
    auto make = [&](unsigned t) {
        LongJob<S> j;
        j.kin = (const unsigned*)kin;
        j.vin = (const unsigned*)vin;
        j.kout = (unsigned*)kout;
        j.vout = (unsigned*)vout;
        j.n = a.n;
        j.tile = t;
        j.status = (S*)a.status;
        j.dstart64 = a.digit_start;
        j.dstart32 = nullptr;
        j.lead = 0;
        j.lds_rank = lds_rank;
        return j;
    };
    radix_pass_long_core<BLOCK, ITEMS, VB, S, KIND>(a.tile_counter, (a.n + TILE - 1) / TILE, make, a.shift, a.kind,
                                                     a.descending);
}

TRTC_DEV void radix_copyback(const RadixCopyArgs& a)
{
    const RadixPlan& P = *a.plan;
    if (!P.copy_ksrc) return;
    radix_copy_span((const void*)P.copy_ksrc, (void*)P.copy_kdst, a.key_bytes);
    if (a.val_bytes && P.copy_vsrc) radix_copy_span((const void*)P.copy_vsrc, (void*)P.copy_vdst, a.val_bytes);
}

// One digit pass over a batch of segments.  Pass 0 reads the segment's
// source, later even passes its own buffers; even passes write the ping-pong
// buffers and odd passes the segment's buffers, so an even number of passes
// ends there.  A one-tile segment has no lookback and
// no pre-pass histogram, and sorts in place (its tile holds every key in
// registers before it writes).
template <int BLOCK, int ITEMS, int KB, int VB, int KIND = -1>
TRTC_DEV void seg_radix_pass(const SegRadixArgs& a)
{
    typedef typename KeyWord<KB>::U U;
    typedef typename ValWord<VB>::U V;
    constexpr int TILE = BLOCK * ITEMS;
    auto make = [&](unsigned t) {
        const SegSortDesc& d = a.descs[a.tile_seg[t]];
        const bool odd = a.pass & 1;
        RadixJob<U, V, unsigned> j;
        j.kin = (const U*)(a.pass == 0 ? d.src_keys : (odd ? d.tmp_keys : d.keys));
        j.vin = (const V*)(a.pass == 0 ? d.src_vals : (odd ? d.tmp_vals : d.vals));
        j.kout = (U*)(odd ? d.keys : d.tmp_keys);
        j.vout = (V*)(odd ? d.vals : d.tmp_vals);
        j.n = d.n;
        j.tile = t - d.first_tile;
        // this pass's input's offset from a 128-byte line, in keys
        j.lead = d.n > (unsigned long long)TILE ? (unsigned)(((size_t)j.kin & 127) / KB) : 0u;
        j.dstart64 = nullptr;
        if (d.n > (unsigned long long)TILE) {
            j.status = a.status + d.status_row * 256;
            j.dstart32 = a.seg_dstart + ((size_t)d.multi * KB + a.pass) * 256;
        } else {
            j.status = nullptr;
            j.dstart32 = nullptr;
        }
        return j;
    };
    radix_pass_core<BLOCK, ITEMS, KB, VB, unsigned, KIND, true>(a.tile_counter, make, a.shift, a.kind,
                                                                a.descending);
}

// The long-tile pass (radix_pass_long_core) over the segments of a batch
// class: the tile map gives the segment, lookback stops at its first tile,
// digit offsets and the ranking mode come from the segment histograms; buffers
// as seg_radix_pass.  One-tile segments sort in place with their own digit
// scan and ballot ranking.
template <int BLOCK, int ITEMS, int VB, int KIND>
TRTC_DEV void seg_radix_pass_long(const SegRadixArgs& a)
{
    constexpr int TILE = BLOCK * ITEMS;
    auto make = [&](unsigned t) {
        const SegSortDesc& d = a.descs[a.tile_seg[t]];
        const bool odd = a.pass & 1;
        LongJob<unsigned> j;
        j.kin = (const unsigned*)(a.pass == 0 ? d.src_keys : (odd ? d.tmp_keys : d.keys));
        j.vin = (const unsigned*)(a.pass == 0 ? d.src_vals : (odd ? d.tmp_vals : d.vals));
        j.kout = (unsigned*)(odd ? d.keys : d.tmp_keys);
        j.vout = (unsigned*)(odd ? d.vals : d.tmp_vals);
        j.n = d.n;
        j.tile = t - d.first_tile;
        // this pass's input's offset from a 128-byte line, in keys
        j.lead = d.n > (unsigned long long)TILE ? (unsigned)(((size_t)j.kin & 127) / 4) : 0u;
        j.dstart64 = nullptr;
        if (d.n > (unsigned long long)TILE) {
            j.status = a.status + d.status_row * 256;
            j.dstart32 = a.seg_dstart + ((size_t)d.multi * 4 + a.pass) * 256;
            j.lds_rank = (a.seg_rank_lds[d.multi] >> a.pass) & 1u;
        } else {
            j.status = nullptr;
            j.dstart32 = nullptr;
            j.lds_rank = false;
        }
        return j;
    };
    radix_pass_long_core<BLOCK, ITEMS, VB, unsigned, KIND>(a.tile_counter, a.num_tiles, make, a.shift, a.kind,
                                                            a.descending);
}

// Digit histograms of the multi-tile segments of a batch: one workgroup per
// chunk of hist_chunk_keys keys (all KB digits from one read, bank-conflict
// free LDS counters as in radix_hist, one global add per bin and workgroup); the last chunk of a
// segment to finish turns the segment's histograms into exclusive digit
// offsets.  Counters and histograms are zeroed by the batch driver.
template <int BLOCK, int KB>
TRTC_DEV void seg_radix_hist(const SegRadixArgs& a)
{
    static_assert(BLOCK % 64 == 0, "whole waves count the keys");
    typedef typename KeyWord<KB>::U U;
    // bin-major LDS counters in C columns, lane l counting into column l % C,
    // as in radix_hist
    constexpr int C = KB <= 4 ? 32 : 16;
    constexpr int BINS = KB * 256;
    __shared__ unsigned s_h[BINS * C];
    __shared__ bool s_last;
    for (int i = threadIdx.x; i < BINS * C; i += BLOCK) s_h[i] = 0;
    __syncthreads();
    unsigned* const hcol = s_h + (lane_id() & (unsigned)(C - 1));

    const unsigned c = blockIdx.x;  // grid = num_hist_chunks
    const unsigned seg = a.hist_chunk[2 * c];
    const SegSortDesc& d = a.descs[seg];
    const size_t ck = a.hist_chunk_keys;
    const size_t base = (size_t)a.hist_chunk[2 * c + 1] * ck;
    const size_t cnt = (d.n - base) < ck ? (size_t)(d.n - base) : ck;
    const U flip = a.descending ? (U)~(U)0 : (U)0;
    hist_count_keys<KB, 4>(gmem((const U*)d.src_keys + base), cnt, threadIdx.x, BLOCK, hcol, a.kind, flip);
    __syncthreads();
    unsigned* hist = a.seg_hist + (size_t)d.multi * KB * 256;
    for (int i = threadIdx.x; i < BINS; i += BLOCK) {
        unsigned t = 0;
#pragma unroll
        for (int j = 0; j < C; j++) t += s_h[i * C + ((j + i) & (C - 1))];  // lane-skewed columns
        if (t) __hip_atomic_fetch_add(hist + i, t, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        unsigned v = __hip_atomic_fetch_add(a.seg_done + d.multi, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        s_last = (v == d.hist_chunks - 1);
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    }
    __syncthreads();
    if (!s_last) return;
    unsigned* ds = a.seg_dstart + (size_t)d.multi * KB * 256;
    // digit offsets, and the ranking mode of the long-tile pass per digit (as
    // radix_hist: LDS rank words when sum (h_d / n)^2 < 1/32)
    if (threadIdx.x < 64) {
        unsigned rank_lds = 0;
        const double inv_n = 1.0 / (double)d.n;
        unsigned hv[KB * 4];  // the lane's counts, loaded together
#pragma unroll
        for (int i = 0; i < KB * 4; i++) hv[i] = ld_relaxed(hist + i * 64 + (int)threadIdx.x);
#pragma unroll
        for (int p = 0; p < KB; p++) {
            unsigned run = 0;
            double sq = 0.0;
#pragma unroll
            for (int q = 0; q < 4; q++) {
                int dd = q * 64 + (int)threadIdx.x;
                unsigned x = hv[p * 4 + q];
                const double f = (double)x * inv_n;
                sq += f * f;
                Plus plus;
                unsigned inc = wave_inclusive_scan(x, plus);
                ds[p * 256 + dd] = run + inc - x;
                run += shfl(inc, 63);
            }
            Plus plus;
            sq = shfl(wave_inclusive_scan(sq, plus), 63);
            if (sq < 1.0 / 32) rank_lds |= 1u << p;
        }
        if (threadIdx.x == 0 && a.seg_rank_lds) a.seg_rank_lds[d.multi] = rank_lds;
    }
}

}  // namespace trtc
