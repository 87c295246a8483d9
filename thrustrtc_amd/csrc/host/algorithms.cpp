// Host drivers of the primitive core: argument checks, kernel selection (an
// ahead-of-time instantiation when one exists, else hipRTC), scratch sizing and
// launches.  Entry points keep the reference's C++ API (sort.h, scan.h,
// reduce.h, copy.h) and defaults (sort.cpp:30-65, scan.cpp:12-39,
// scan_by_key.cpp:4-53, reduce.cpp:5-31, copy.cpp:16-28).
#include <algorithm>
#include <deque>

#include "copy.h"
#include "count.h"
#include "fake_vectors/DVTransform.h"
#include "partition.h"
#include "reduce.h"
#include "remove.h"
#include "transform_reduce.h"
#include "transform_scan.h"
#include "unique.h"
#include "scan.h"
#include "sort.h"
#include "trtc_internal.h"
#include "../device/trtc_args.h"

using namespace trtc_host;

namespace {

constexpr unsigned kBlock = 256;
constexpr unsigned kScanBlock = 512;  // TRTC_SCAN_BLOCK: scans, selections, by-key kernels
constexpr unsigned kCopyIfBlock = 1024;  // TRTC_COPY_IF_BLOCK
constexpr unsigned kSelBlock = 1024;     // TRTC_SEL_BLOCK: Unique without values

size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

// Items per thread of the reduce chunks (must match ReduceItems).
int items_for(size_t elem_size) { return elem_size <= 4 ? 16 : (elem_size <= 8 ? 8 : 4); }
// Items per thread of the single-pass tiles -- whole-input scan, compaction,
// by-key scan and reduction: 128 bytes per lane (must match ScanItems,
// CopyIfItems and ScanByKeyItems in the device headers).
int tile_items_for(size_t elem_size) { return elem_size <= 4 ? 32 : (elem_size <= 8 ? 16 : 4); }
// Block of Scan_By_Key from the wider of the key and value types (WideBlock
// in the device header).
unsigned wide_block(size_t elem_size) { return elem_size == 8 ? 1024u : kScanBlock; }
size_t words_for(size_t bytes) { return (bytes + 3) / 4; }

std::string vcls(const DeviceViewable& v) { return canon_type(v.name_view_cls()); }

// Kernel arguments: each argument's bytes are kept alive until launch.
struct Args
{
    std::deque<ViewBuf> bufs;
    std::vector<void*> ptrs;
    void view(const DeviceViewable& v)
    {
        bufs.push_back(v.view());
        ptrs.push_back(bufs.back().data());
    }
    void raw(const void* p, size_t n)
    {
        bufs.push_back(ViewBuf((const char*)p, (const char*)p + n));
        ptrs.push_back(bufs.back().data());
    }
    template <class T> void val(const T& v) { raw(&v, sizeof(T)); }
    void** data() { return ptrs.data(); }
};

// Scalar value of an init DeviceViewable converted to `to` through the
// reference's cast chain: init -> input element type -> output element type.
bool init_value(const DeviceViewable* init, const std::string& in_elem, const std::string& out_elem, ViewBuf& out)
{
    const ScalarInfo* so = scalar_info(out_elem);
    size_t out_size = so ? so->size : TRTC_Size_Of(out_elem.c_str());
    out.assign(out_size, 0);
    if (!init) return true;  // default: (T)0
    ViewBuf iv = init->view();
    const ScalarInfo* si = scalar_info(init->name_view_cls());
    const ScalarInfo* sm = scalar_info(in_elem);
    if (si && sm && so) {
        char mid[8];
        if (!convert_scalar(si->code, iv.data(), sm->code, mid)) return false;
        return convert_scalar(sm->code, mid, so->code, out.data());
    }
    if (canon_type(init->name_view_cls()) == canon_type(out_elem) && iv.size() >= out_size) {
        memcpy(out.data(), iv.data(), out_size);
        return true;
    }
    TRTC_ERR("cannot convert init of type %s to %s", init->name_view_cls().c_str(), out_elem.c_str());
    return false;
}

struct LookbackScratch
{
    PoolBuf buf;
    unsigned* counter = nullptr;
    unsigned long long* agg = nullptr;
    unsigned long long* incl = nullptr;
    size_t words = 0;
    bool make(size_t tiles, size_t value_bytes, hipStream_t s)
    {
        size_t W = words_for(value_bytes);
        size_t bytes = 16 + 2 * tiles * W * 8;
        new (&buf) PoolBuf(bytes, s);
        if (!buf.p) return false;
        counter = (unsigned*)buf.p;
        agg = (unsigned long long*)((char*)buf.p + 16);
        incl = agg + tiles * W;
        return fill_u32(buf.p, bytes / 4, 0, s);
    }
    // The same layout in caller-provided memory (nothing is allocated),
    // cleared with agent-scope atomic stores: the memory is the library's
    // static device array, and its clear must reach the scan's agent-scope
    // atomics (tile counter, status words) whatever its caching.
    bool use(void* mem, size_t capacity, size_t tiles, size_t value_bytes, hipStream_t s)
    {
        size_t W = words_for(value_bytes);
        size_t bytes = 16 + 2 * tiles * W * 8;
        if (bytes > capacity) {
            TRTC_ERR("scan: %zu tiles need %zu bytes of lookback scratch, more than the %zu available", tiles,
                     bytes, capacity);
            return false;
        }
        counter = (unsigned*)mem;
        agg = (unsigned long long*)((char*)mem + 16);
        incl = agg + tiles * W;
        return clear_u32_agent(mem, bytes / 4, s);
    }
};

bool radix_elem(const std::string& elem, int& key_bytes, int& kind) { return radix_key_info(elem, key_bytes, kind); }

// ---------------------------------------------------------------- radix sort
// Onesweep LSD radix sort of n keys (key_bytes wide) with optional values
// (val_bytes in {0,1,2,4,8}); result in keys/vals.  Stable.  With keys_src
// (and vals_src) the first pass reads the input from there instead
// (out-of-place sort, no extra traffic).
bool radix_sort(void* keys, void* vals, size_t n, int key_bytes, int kind, bool descending, int val_bytes,
                hipStream_t s, const void* keys_src = nullptr, const void* vals_src = nullptr)
{
    if (keys_src && n == 1) {
        if (!check_hip(hipMemcpyAsync(keys, keys_src, key_bytes, hipMemcpyDeviceToDevice, s), "hipMemcpyAsync"))
            return false;
        return !val_bytes ||
               check_hip(hipMemcpyAsync(vals, vals_src, val_bytes, hipMemcpyDeviceToDevice, s), "hipMemcpyAsync");
    }
    if (n <= 1) return true;
    if (!keys_src) {
        keys_src = keys;
        vals_src = vals;
    }
    const int KB = key_bytes;
    // tile shape of trtc_radix_pass_k (RadixItems in the device header):
    // 1024 threads once that gives every CU a tile.  u32/u32 pairs, per sort:
    // 2^22 0.189 -> 0.168 ms, 2^24 0.453 -> 0.414, 2^28 5.96 -> 4.94 against
    // 512 threads; equal at 2^21, 512 faster below (2^20 0.131 vs 0.140 ms).
    const int ITEMS = (KB > 4 || val_bytes > 4) ? 8 : 16;
    const size_t cus = (size_t)std::max(1, context().num_cus);
    const unsigned kRadixBlock = n / (1024 * (size_t)ITEMS) >= cus ? 1024u : 512u;
    // 4-byte keys with 4-byte or no values: the long-tile pass
    // (trtc_radix_pass_long_k, TRTC_LONG_BLOCK x TRTC_LONG_ITEMS = 1024 x 32
    // keys in the device header) once every CU gets a tile; one workgroup per
    // CU loops over the tiles (long_pass_grid).  u32/u32 sorts, long tiles
    // against 1024x16 (one box): 2^22 0.165 vs 0.134 ms, 2^23 0.187 vs 0.211,
    // 2^24 0.325 vs 0.369, 2^26 1.13 vs 1.28, 2^27 2.10 vs 2.46
    constexpr size_t kLongTile = 1024 * 32;
    const bool long_tiles = KB == 4 && (val_bytes == 0 || val_bytes == 4) && n / kLongTile >= cus;
    const size_t TILE = long_tiles ? kLongTile : (size_t)kRadixBlock * ITEMS;
    const size_t tiles = (n + TILE - 1) / TILE;
    const size_t grid = long_tiles ? long_pass_grid(tiles) : tiles;
    const bool wide = n >= (1ull << 30);
    const size_t sw = wide ? 8 : 4;

    PoolBuf ktmp(n * KB, s);
    PoolBuf vtmp(val_bytes ? n * val_bytes : 0, s);
    // scratch: hist [KB*256 u64] | done u32 (+pad) | tile counters [8 passes][8 u32] |
    // status [KB][tiles][256] | digit_start [KB*256 u64] | plan (the passes' buffers and
    // ranking modes, written by the histogram pre-pass)
    size_t off_hist = 0;
    size_t off_done = off_hist + KB * 256 * 8;
    size_t off_cnt = off_done + 16;
    size_t off_status = align_up(off_cnt + 8 * 8 * 4, 256);
    size_t status_bytes_per_pass = align_up(tiles * 256 * sw, 256);
    size_t off_dstart = off_status + KB * status_bytes_per_pass;
    size_t off_plan = align_up(off_dstart + KB * 256 * 8, 256);
    size_t scratch_bytes = off_plan + align_up(sizeof(trtc::RadixPlan), 256);
    PoolBuf scratch(scratch_bytes, s);
    if (!ktmp.p || (val_bytes && !vtmp.p) || !scratch.p) return false;
    char* sc = (char*)scratch.p;

    // The pre-pass accumulates into the stream's histogram block, which is
    // zero between sorts and cleared again by the pre-pass itself, so no
    // clearing launch precedes it (2^20 u32 sort: 6 -> 5 launches).  Under
    // graph capture the sort keeps its own counters in the scratch and clears
    // them first, so that the graph holds no state shared with other sorts.
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(s, &cap) != hipSuccess) {
        (void)hipGetLastError();
        cap = hipStreamCaptureStatusNone;
    }
    const bool in_capture = capturing() || cap != hipStreamCaptureStatusNone;
    char* hb = in_capture ? nullptr : (char*)radix_hist_block(s);
    if (!hb && !fill_u32(sc, off_cnt / 4, 0, s)) return false;  // hist + done

    KernelRef hk;
    std::string hname = "trtc_radix_hist_k<" + std::to_string(KB) + ">";
    if (!get_template_kernel(hname, "", hk)) return false;
    trtc::RadixHistArgs ha{};
    ha.keys = keys_src;
    ha.n = n;
    ha.kind = kind;
    ha.descending = descending ? 1 : 0;
    ha.passes = KB;
    ha.hist = (unsigned long long*)(hb ? hb : sc + off_hist);
    ha.digit_start = (unsigned long long*)(sc + off_dstart);
    ha.done = (unsigned*)(hb ? hb + 8 * 256 * 8 : sc + off_done);
    ha.self_clear = hb ? 1 : 0;
    ha.zero_ptr = (unsigned*)(sc + off_cnt);
    ha.zero_words = (off_dstart - off_cnt) / 4;
    // passes over constant digits (one bin holds all n keys) are skipped
    // through the device-side plan, without a host synchronisation
    trtc::RadixPlan* plan = (trtc::RadixPlan*)(sc + off_plan);
    ha.plan = plan;
    ha.vals_src = vals_src;
    ha.keys_dst = keys;
    ha.vals_dst = vals;
    ha.keys_tmp = ktmp.p;
    ha.vals_tmp = vtmp.p;
    ha.val_bytes = val_bytes;
    // 1024-thread workgroups (TRTC_HIST_BLOCK), at most one per CU for their
    // LDS, at least 16 keys per thread: every workgroup adds its KB*256 bins
    // into the same global counters, so a grid larger than the work needs
    // piles same-address atomics on them (2^20 keys: 48.7 us with 1024
    // 256-thread workgroups)
    constexpr size_t kHistBlock = 1024, kHistKeys = 16;
    unsigned hgrid = (unsigned)std::min<size_t>((n + kHistBlock * kHistKeys - 1) / (kHistBlock * kHistKeys),
                                                (size_t)std::max(1, context().num_cus));
    {
        void* args[] = {&ha};
        if (!launch_kernel(hk, dim3(hgrid), dim3((unsigned)kHistBlock), args, 0, s)) return false;
    }

    KernelRef pk;
    std::string pname =
        long_tiles ? "trtc_radix_pass_long_k<" + std::to_string(val_bytes) + ", " +
                         (wide ? "unsigned long long" : "unsigned int") + ", " + std::to_string(kind) + ">"
                   : "trtc_radix_pass_k<" + std::to_string(KB) + ", " + std::to_string(val_bytes) + ", " +
                         (wide ? "unsigned long long" : "unsigned int") + ", " + std::to_string(kRadixBlock) + ", " +
                         std::to_string(kRadixBlock == 1024 ? kind : -1) + ">";
    const unsigned pass_block = long_tiles ? 1024u : kRadixBlock;
    if (!get_template_kernel(pname, "", pk)) return false;
    trtc::RadixPassArgs pa{};
    void* kbuf[2] = {keys, ktmp.p};
    void* vbuf[2] = {vals, vtmp.p};
    for (int p = 0; p < KB; p++) {
        pa.keys_in = p == 0 ? keys_src : kbuf[p & 1];
        pa.keys_out = kbuf[(p + 1) & 1];
        pa.vals_in = p == 0 ? vals_src : vbuf[p & 1];
        pa.vals_out = vbuf[(p + 1) & 1];
        pa.n = n;
        pa.digit_start = (const unsigned long long*)(sc + off_dstart) + p * 256;
        pa.status = sc + off_status + p * status_bytes_per_pass;
        pa.tile_counter = (unsigned*)(sc + off_cnt) + p * 8;  // the long-tile pass uses all 8
        pa.shift = 8 * p;
        pa.kind = kind;
        pa.descending = descending ? 1 : 0;
        pa.plan = plan;
        pa.pass = p;
        void* args[] = {&pa};
        if (!launch_kernel(pk, dim3((unsigned)grid), dim3(pass_block), args, 0, s)) return false;
    }
    // The parity copy of an in-place sort with an odd number of sorting
    // passes runs in the launch of a pass whose digit is constant (the plan's
    // copy_mask), which exists for every even key width; only 1-byte keys
    // need a copy kernel after the pass (its workgroups return at once when
    // the plan needs no copy, so the grid is at most one workgroup per CU,
    // grid-stride when it copies).  Without it a sort is 6 launches: fill,
    // histograms, 4 passes (2^28 u32/u32: 11 us less per sort).
    if (KB != 1) return true;
    KernelRef ck;
    if (!get_template_kernel("trtc_radix_copyback_k", "", ck)) return false;
    trtc::RadixCopyArgs ca{plan, n * (size_t)KB, n * (size_t)val_bytes};
    {
        void* args[] = {&ca};
        const size_t vecs = (n * (size_t)KB + 15) / 16;
        const unsigned cgrid = (unsigned)std::min<size_t>((vecs + 255) / 256, (size_t)std::max(1, context().num_cus));
        if (!launch_kernel(ck, dim3(std::max(1u, cgrid)), dim3(256), args, 0, s)) return false;
    }
    return true;
}

// ---------------------------------------------------------------- merge sort
// Stable comparison sort with any comparator functor (merge_sort.cpp:7-280):
// tiles of 2048 keys sorted in LDS, then log2(n / 2048) merge passes; keys
// carry their input index when values have to follow.
constexpr size_t kMsortTile = 256 * 8;  // TRTC_MSORT_BLOCK * TRTC_MSORT_ITEMS

bool merge_sort(void* keys, const std::string& T, size_t esize, size_t n, const Functor& comp, void* vals,
                size_t vsize, hipStream_t s)
{
    if (n <= 1) return true;
    const std::string cc = vcls(comp);
    const std::string tname = "trtc_msort_tile_k<" + T + ", " + cc + ">";
    const std::string mname = "trtc_msort_merge_k<" + T + ", " + cc + ">";
    const std::string sname = "trtc_msort_split_k<" + T + ", " + cc + ">";
    KernelRef tk, mk, sk;
    if (!get_template_kernel(tname, tname, tk)) return false;
    const bool with_idx = vals != nullptr;
    PoolBuf ia(with_idx ? n * 4 : 0, s), ib(with_idx ? n * 4 : 0, s), kb(n * esize, s);
    if ((with_idx && (!ia.p || !ib.p)) || !kb.p) return false;
    if (with_idx && !iota_u32(ia.as<unsigned>(), n, s)) return false;
    const unsigned tiles = (unsigned)((n + kMsortTile - 1) / kMsortTile);
    PoolBuf splits(n > kMsortTile ? (size_t)(tiles + 1) * 8 : 0, s);
    if (n > kMsortTile && !splits.p) return false;
    trtc::MergeSortArgs a{n, kMsortTile, with_idx ? 1 : 0, nullptr};
    {
        Args args;
        args.val(keys);
        args.val(ia.p);
        args.view(comp);
        args.val(a);
        if (!launch_kernel(tk, dim3(tiles), dim3(256), args.data(), 0, s)) return false;
    }
    void* kbuf[2] = {keys, kb.p};
    void* ibuf[2] = {ia.p, ib.p};
    int cur = 0;
    if (n > kMsortTile) {
        if (!get_template_kernel(mname, mname, mk) || !get_template_kernel(sname, sname, sk)) return false;
        a.splits = splits.as<unsigned long long>();
        for (size_t w = kMsortTile; w < n; w *= 2) {
            a.width = w;
            {
                // every tile's split first (all searches of the pass at once)
                Args sargs;
                sargs.val(kbuf[cur]);
                sargs.view(comp);
                sargs.val(a);
                if (!launch_kernel(sk, dim3((tiles + 255) / 256), dim3(256), sargs.data(), 0, s)) return false;
            }
            Args args;
            args.val(kbuf[cur]);
            args.val(ibuf[cur]);
            args.val(kbuf[cur ^ 1]);
            args.val(ibuf[cur ^ 1]);
            args.view(comp);
            args.val(a);
            if (!launch_kernel(mk, dim3(tiles), dim3(256), args.data(), 0, s)) return false;
            cur ^= 1;
        }
    }
    if (cur == 1 &&
        !copy_bytes(kb.p, keys, n * esize, s))
        return false;
    if (!with_idx) return true;
    PoolBuf vcopy(n * vsize, s);
    if (!vcopy.p) return false;
    if (!copy_bytes(vals, vcopy.p, n * vsize, s))
        return false;
    return gather_bytes(vcopy.p, vals, (const unsigned*)ibuf[cur], n, vsize, s);
}

bool sort_dispatch(DVVectorLike& keys, DVVectorLike* values, const Functor& comp)
{
    if (!ensure_init()) return false;
    if (!contiguous_data(keys) || (values && !contiguous_data(*values))) {
        // computed views (permutations, reversals, zipped or custom vectors):
        // sort contiguous copies, then write the results back through the views
        size_t n = keys.size();
        if (!keys.is_readable() || !keys.is_writable() ||
            (values && (!values->is_readable() || !values->is_writable()))) {
            TRTC_ERR("Sort: vectors must be readable and writable");
            return false;
        }
        if (values && values->size() < n) {
            TRTC_ERR("Sort_By_Key: values (%zu) shorter than keys (%zu)", values->size(), n);
            return false;
        }
        if (n == 0) return true;
        hipStream_t s = cur_stream();
        PoolBuf kb(n * keys.elem_size(), s), vb(values ? n * values->elem_size() : 0, s);
        if (!kb.p || (values && !vb.p)) return false;
        DVVectorAdaptor kc(keys.name_elem_cls().c_str(), n, kb.p);
        DVVectorAdaptor vc(values ? values->name_elem_cls().c_str() : "int32_t", values ? n : 0, vb.p);
        DVRange vr(values ? *values : keys, 0, n);
        if (!TRTC_Copy(keys, kc) || (values && !TRTC_Copy(vr, vc))) return false;
        if (!sort_dispatch(kc, values ? &vc : nullptr, comp)) return false;
        if (!TRTC_Copy(kc, keys) || (values && !TRTC_Copy(vc, vr))) return false;
        return check_hip(hipStreamSynchronize(s), "hipStreamSynchronize");
    }
    const std::string& cn = comp.name_view_cls();
    int kb = 0, kind = 0;
    if ((cn == "Less" || cn == "Greater") && radix_elem(keys.name_elem_cls(), kb, kind)) {
        void* kp = contiguous_data(keys);
        if (!kp) {
            TRTC_ERR("Sort: keys must be a DVVector or DVVectorAdaptor");
            return false;
        }
        size_t n = keys.size();
        hipStream_t s = cur_stream();
        bool desc = cn == "Greater";
        if (!values) return radix_sort(kp, nullptr, n, kb, kind, desc, 0, s);
        void* vp = contiguous_data(*values);
        if (!vp) {
            TRTC_ERR("Sort_By_Key: values must be a DVVector or DVVectorAdaptor");
            return false;
        }
        if (values->size() < n) {
            TRTC_ERR("Sort_By_Key: values (%zu) shorter than keys (%zu)", values->size(), n);
            return false;
        }
        size_t vb = values->elem_size();
        if (vb == 1 || vb == 2 || vb == 4 || vb == 8) return radix_sort(kp, vp, n, kb, kind, desc, (int)vb, s);
        // other value sizes: sort an index payload, then gather the values
        PoolBuf idx(n * 4, s), vcopy(n * vb, s);
        if (!idx.p || !vcopy.p) return false;
        if (!iota_u32(idx.as<unsigned>(), n, s)) return false;
        if (!radix_sort(kp, idx.p, n, kb, kind, desc, 4, s)) return false;
        if (!copy_bytes(vp, vcopy.p, n * vb, s))
            return false;
        return gather_bytes(vcopy.p, vp, idx.as<unsigned>(), n, vb, s);
    }
    // any other comparator or key type: stable merge sort
    void* kp = contiguous_data(keys);
    if (!kp) {
        TRTC_ERR("Sort: keys must be a DVVector or DVVectorAdaptor");
        return false;
    }
    size_t n = keys.size();
    void* vp = nullptr;
    size_t vb = 0;
    if (values) {
        vp = contiguous_data(*values);
        if (!vp) {
            TRTC_ERR("Sort_By_Key: values must be a DVVector or DVVectorAdaptor");
            return false;
        }
        if (values->size() < n) {
            TRTC_ERR("Sort_By_Key: values (%zu) shorter than keys (%zu)", values->size(), n);
            return false;
        }
        vb = values->elem_size();
    }
    if (n >= (1ull << 32)) {
        TRTC_ERR("Sort: comparison sort of %zu keys (limit 2^32 - 1)", n);
        return false;
    }
    return merge_sort(kp, canon_type(keys.name_elem_cls()), keys.elem_size(), n, comp, vp, vb, cur_stream());
}

// ---------------------------------------------------------------- scan

// With fixed_scratch, the lookback state lives in that caller memory and the
// kernel must already exist (ahead-of-time, or compiled by an earlier call):
// the call then allocates nothing and compiles nothing.
bool general_scan(const DVVectorLike& vin, DVVectorLike& vout, const Functor& op, const DeviceViewable* init,
                  bool exclusive, void* fixed_scratch = nullptr, size_t fixed_bytes = 0)
{
    if (!ensure_init()) return false;
    size_t n = vin.size();
    if (n == 0) return true;
    if (vout.size() < n) {
        TRTC_ERR("Scan: output shorter than input");
        return false;
    }
    std::string T = canon_type(vout.name_elem_cls());
    size_t tsz = vout.elem_size();
    ViewBuf initv;
    if (!init_value(init, vin.name_elem_cls(), vout.name_elem_cls(), initv)) return false;
    std::string name = "trtc_scan_k<" + vcls(vin) + ", " + vcls(vout) + ", " + vcls(op) + ", " + T + ">";
    KernelRef k;
    if (fixed_scratch) {
        if (!find_template_kernel(name, name, k)) {
            TRTC_ERR("scan without allocation: %s is neither built in nor compiled by an earlier scan call of "
                     "this process", name.c_str());
            return false;
        }
    } else if (!get_template_kernel(name, name, k)) {
        return false;
    }
    hipStream_t s = cur_stream();
    const unsigned block = tsz == 8 ? 1024u : kScanBlock;  // ScanBlock<T> in the device header
    size_t tile = (size_t)block * tile_items_for(tsz);
    size_t tiles = (n + tile - 1) / tile;
    LookbackScratch lb;
    if (fixed_scratch ? !lb.use(fixed_scratch, fixed_bytes, tiles, tsz, s) : !lb.make(tiles, tsz, s)) return false;
    trtc::ScanArgs a{n, {lb.agg, lb.incl, lb.counter}, exclusive ? 1 : 0};
    Args args;
    args.view(vin);
    args.view(vout);
    args.view(op);
    args.raw(initv.data(), initv.size());
    args.val(a);
    return launch_kernel(k, dim3((unsigned)tiles), dim3(block), args.data(), 0, s);
}

// ---------------------------------------------------------------- reduce

bool general_reduce(const DVVectorLike& vec, const DeviceViewable* init, const Functor& op, ViewBuf& ret)
{
    if (!ensure_init()) return false;
    size_t n = vec.size();
    std::string T = canon_type(vec.name_elem_cls());
    size_t tsz = vec.elem_size();
    ViewBuf initv;
    if (!init_value(init, vec.name_elem_cls(), vec.name_elem_cls(), initv)) return false;
    ret.assign(tsz, 0);
    if (n == 0) {
        if (!init) return false;  // the reference's general_reduce rejects n < 1
        ret = initv;
        return true;
    }
    std::string name = "trtc_reduce_k<" + vcls(vec) + ", " + vcls(op) + ", " + T + ">";
    KernelRef k;
    if (!get_template_kernel(name, name, k)) return false;
    hipStream_t s = cur_stream();
    const size_t waves_per_block = kBlock / 64;
    const size_t chunk = 64 * (size_t)items_for(tsz);
    size_t chunks = (n + chunk - 1) / chunk;
    // four 256-thread workgroups per CU: every workgroup of either kernel
    // path (ordered or streaming) is resident at once
    size_t max_blocks = (size_t)std::max(1, context().num_cus) * 4;
    size_t blocks = std::min(max_blocks, (chunks + waves_per_block - 1) / waves_per_block);
    size_t waves = blocks * waves_per_block;
    size_t per_wave = (chunks + waves - 1) / waves * chunk;
    blocks = (n + per_wave * waves_per_block - 1) / (per_wave * waves_per_block);
    // partials: Maybe<T> = {T v; unsigned has} per block
    size_t maybe_bytes = align_up(tsz + 4, std::min<size_t>(std::max<size_t>(tsz, 4), 8));
    size_t W = words_for(maybe_bytes);
    size_t res_off = 16 + blocks * W * 8;
    PoolBuf scratch(res_off + align_up(tsz, 8), s);
    if (!scratch.p) return false;
    if (!fill_u32(scratch.p, res_off / 4, 0, s)) return false;
    trtc::ReduceArgs a{n, per_wave, (unsigned long long*)((char*)scratch.p + 16), (unsigned*)scratch.p,
                     (char*)scratch.p + res_off, init ? 1 : 0};
    Args args;
    args.view(vec);
    args.view(op);
    args.raw(initv.data(), initv.size());
    args.val(a);
    if (!launch_kernel(k, dim3((unsigned)blocks), dim3(kBlock), args.data(), 0, s)) return false;
    if (!check_hip(hipMemcpyAsync(ret.data(), a.result, tsz, hipMemcpyDeviceToHost, s), "hipMemcpyAsync"))
        return false;
    return check_hip(hipStreamSynchronize(s), "hipStreamSynchronize");
}

// Min/Max_Element on a plain arithmetic vector (extrema.cpp:5-80): one
// streaming launch that carries (index, value) pairs, index ties going to the
// lower index (trtc_extreme_k).  `handled` is false for other vectors, which
// take the ordered reduction of zipped (index, value) pairs instead.
bool general_extreme(const DVVectorLike& vec, const Functor& comp, bool max, size_t& id, bool& handled)
{
    handled = false;
    const ScalarInfo* si = scalar_info(vec.name_elem_cls());
    if (!si || si->code > TC_F64) return true;  // bool and other types: not here
    const std::string view = vcls(vec);
    if (view.compare(0, 11, "VectorView<") != 0) return true;
    handled = true;
    if (!ensure_init()) return false;
    const size_t n = vec.size();
    const size_t tsz = si->size;
    std::string name = "trtc_extreme_k<" + view + ", " + vcls(comp) + ", " + (max ? "1" : "0") + ">";
    KernelRef k;
    if (!get_template_kernel(name, name, k)) return false;
    hipStream_t s = cur_stream();
    // as the streaming reduce: four 256-thread workgroups per CU at most
    const size_t chunk = 64 * 16;
    size_t chunks = (n + chunk - 1) / chunk;
    size_t blocks = std::min<size_t>((size_t)std::max(1, context().num_cus) * 4, (chunks + 3) / 4);
    if (blocks == 0) blocks = 1;
    // partials: Maybe<ExtPair<T>> = {u64 index; T value; unsigned has} <= 24 bytes
    const size_t W = words_for(align_up(align_up(8 + tsz, 8) + 4, 8));
    const size_t res_off = 16 + blocks * W * 8;
    PoolBuf scratch(res_off + 16, s);
    if (!scratch.p) return false;
    if (!fill_u32(scratch.p, res_off / 4, 0, s)) return false;
    trtc::ReduceArgs a{n, 0, (unsigned long long*)((char*)scratch.p + 16), (unsigned*)scratch.p,
                     (char*)scratch.p + res_off, 0};
    Args args;
    args.view(vec);
    args.view(comp);
    args.val(a);
    if (!launch_kernel(k, dim3((unsigned)blocks), dim3(kBlock), args.data(), 0, s)) return false;
    uint64_t r = 0;
    if (!check_hip(hipMemcpyAsync(&r, a.result, 8, hipMemcpyDeviceToHost, s), "hipMemcpyAsync")) return false;
    if (!check_hip(hipStreamSynchronize(s), "hipStreamSynchronize")) return false;
    id = (size_t)r;
    return true;
}

// ---------------------------------------------------------------- copy_if

// The stencil argument of a selection: the stencil's view, or the one-byte
// _NoStencil marker (the predicate then reads the input).
std::string stencil_cls(const DVVectorLike* sten) { return sten ? vcls(*sten) : std::string("_NoStencil"); }
void stencil_arg(Args& args, const DVVectorLike* sten)
{
    if (sten) {
        args.view(*sten);
    } else {
        char none = 0;
        args.raw(&none, 1);
    }
}

// Launch a selection kernel (its last argument is CopyIfArgs) over n items in
// tiles of `tile`, lookback state of `state_bytes`; returns the selected count
// (synchronous: the count is a host value) or 0xFFFFFFFF.
uint32_t run_selection(const std::string& name, Args& args, size_t n, size_t tile, unsigned block,
                       size_t state_bytes, bool has_stencil)
{
    KernelRef k;
    if (!get_template_kernel(name, name, k)) return (uint32_t)-1;
    hipStream_t s = cur_stream();
    size_t tiles = (n + tile - 1) / tile;
    LookbackScratch lb;
    if (!lb.make(tiles, state_bytes, s)) return (uint32_t)-1;
    PoolBuf cnt(8, s);
    if (!cnt.p) return (uint32_t)-1;
    trtc::CopyIfArgs a{n, {lb.agg, lb.incl, lb.counter}, cnt.as<unsigned long long>(), (unsigned)tiles,
                       has_stencil ? 1 : 0};
    args.val(a);
    if (!launch_kernel(k, dim3((unsigned)tiles), dim3(block), args.data(), 0, s)) return (uint32_t)-1;
    unsigned long long count = 0;
    if (!check_hip(hipMemcpyAsync(&count, cnt.p, 8, hipMemcpyDeviceToHost, s), "hipMemcpyAsync")) return (uint32_t)-1;
    if (!check_hip(hipStreamSynchronize(s), "hipStreamSynchronize")) return (uint32_t)-1;
    if (count >= 0xFFFFFFFFull) {
        TRTC_ERR("selection: count does not fit the 32-bit return value");
        return (uint32_t)-1;
    }
    return (uint32_t)count;
}

uint32_t general_copy_if(const DVVectorLike& vin, const DVVectorLike* sten, DVVectorLike& vout, const Functor& pred)
{
    if (!ensure_init()) return (uint32_t)-1;
    size_t n = vin.size();
    if (n == 0) return 0;
    if (sten && sten->size() < n) {
        TRTC_ERR("Copy_If_Stencil: stencil shorter than input");
        return (uint32_t)-1;
    }
    std::string name = "trtc_copy_if_k<" + vcls(vin) + ", " + stencil_cls(sten) + ", " + vcls(vout) + ", " +
                       vcls(pred) + ">";
    Args args;
    args.view(vin);
    stencil_arg(args, sten);
    args.view(vout);
    args.view(pred);
    // items per thread follow the wider of the input and the stencil (kernel: CopyIfItems)
    const size_t esz = std::max(vin.elem_size(), sten ? sten->elem_size() : (size_t)1);
    return run_selection(name, args, n, (size_t)kCopyIfBlock * tile_items_for(esz), kCopyIfBlock, 8, sten != nullptr);
}

// Unique / Unique_By_Key (unique.cpp:5-95): keep item i when i == 0 or
// !pred(key[i], key[i-1]); values, when given, follow their keys.
uint32_t general_unique(const DVVectorLike& kin, const DVVectorLike* vin, DVVectorLike& kout, DVVectorLike* vout,
                        const Functor& pred)
{
    if (!ensure_init()) return (uint32_t)-1;
    size_t n = kin.size();
    if (n == 0) return 0;
    if (vin && vin->size() < n) {
        TRTC_ERR("Unique_By_Key: values shorter than keys");
        return (uint32_t)-1;
    }
    Args args;
    std::string name;
    size_t esz = kin.elem_size();
    if (vin) {
        name = "trtc_unique_by_key_k<" + vcls(kin) + ", " + vcls(*vin) + ", " + vcls(kout) + ", " + vcls(*vout) +
               ", " + vcls(pred) + ">";
        esz = std::max(esz, vin->elem_size());
        args.view(kin);
        args.view(*vin);
        args.view(kout);
        args.view(*vout);
    } else {
        name = "trtc_unique_k<" + vcls(kin) + ", " + vcls(kout) + ", " + vcls(pred) + ">";
        args.view(kin);
        args.view(kout);
    }
    args.view(pred);
    const unsigned block = vin ? kScanBlock : kSelBlock;
    return run_selection(name, args, n, (size_t)block * tile_items_for(esz), block, 8, false);
}

// Stable two-way split (partition.cpp:28-130): items whose predicate holds go
// to tout in order, the others to fout in order; returns the first count.
uint32_t general_partition_copy(const DVVectorLike& vin, const DVVectorLike* sten, DVVectorLike& tout,
                                DVVectorLike& fout, const Functor& pred)
{
    if (!ensure_init()) return (uint32_t)-1;
    size_t n = vin.size();
    if (n == 0) return 0;
    if (sten && sten->size() < n) {
        TRTC_ERR("Partition: stencil shorter than input");
        return (uint32_t)-1;
    }
    std::string name = "trtc_partition_k<" + vcls(vin) + ", " + stencil_cls(sten) + ", " + vcls(tout) + ", " +
                       vcls(fout) + ", " + vcls(pred) + ">";
    Args args;
    args.view(vin);
    stencil_arg(args, sten);
    args.view(tout);
    args.view(fout);
    args.view(pred);
    const size_t esz = std::max(vin.elem_size(), sten ? sten->elem_size() : (size_t)1);
    return run_selection(name, args, n, (size_t)kScanBlock * tile_items_for(esz), kScanBlock, 8, sten != nullptr);
}

// In-place partition: selected items move forward in place (each lands at or
// before its own index); the rest go to a temporary, then after them.
uint32_t general_partition(DVVectorLike& vec, const DVVectorLike* sten, const Functor& pred)
{
    if (!ensure_init()) return (uint32_t)-1;
    size_t n = vec.size();
    if (n == 0) return 0;
    hipStream_t s = cur_stream();
    PoolBuf tmp(n * vec.elem_size(), s);
    if (!tmp.p) return (uint32_t)-1;
    DVVectorAdaptor falses(vec.name_elem_cls().c_str(), n, tmp.p);
    uint32_t c = general_partition_copy(vec, sten, vec, falses, pred);
    if (c == (uint32_t)-1 || c == n) return c;
    DVVectorAdaptor fpart(falses, 0, n - c);
    DVRange dst(vec, c, n);
    if (!TRTC_Copy(fpart, dst)) return (uint32_t)-1;
    if (!check_hip(hipStreamSynchronize(s), "hipStreamSynchronize")) return (uint32_t)-1;
    return c;
}

// Reduce_By_Key (reduce.cpp:36-62) in one pass: segmented scan of the values
// and compaction of segment tails (see trtc_select.hpp).
uint32_t general_reduce_by_key(const DVVectorLike& kin, const DVVectorLike& vin, DVVectorLike& kout,
                               DVVectorLike& vout, const Functor& pred, const Functor& op)
{
    if (!ensure_init()) return (uint32_t)-1;
    size_t n = kin.size();
    if (n == 0) return 0;
    if (vin.size() < n) {
        TRTC_ERR("Reduce_By_Key: values shorter than keys");
        return (uint32_t)-1;
    }
    if (n >= 0xFFFFFFFFull) {
        TRTC_ERR("Reduce_By_Key: %zu keys (the count is a 32-bit value)", n);
        return (uint32_t)-1;
    }
    std::string V = canon_type(vout.name_elem_cls());
    size_t vsz = vout.elem_size();
    std::string name = "trtc_reduce_by_key_k<" + vcls(kin) + ", " + vcls(vin) + ", " + vcls(kout) + ", " +
                       vcls(vout) + ", " + vcls(pred) + ", " + vcls(op) + ", " + V + ">";
    Args args;
    args.view(kin);
    args.view(vin);
    args.view(kout);
    args.view(vout);
    args.view(pred);
    args.view(op);
    // state {V v; unsigned flag; unsigned tails;} fits align_up(V, 8) + 8 bytes
    const size_t esz = std::max(kin.elem_size(), vsz);
    return run_selection(name, args, n, (size_t)kScanBlock * tile_items_for(esz), kScanBlock, align_up(vsz, 8) + 8,
                         false);
}

// ---------------------------------------------------------------- scan_by_key

bool general_scan_by_key(const DVVectorLike& keys, const DVVectorLike& vals, DVVectorLike& out, const Functor& pred,
                         const Functor& op, const DeviceViewable* init, bool exclusive)
{
    if (!ensure_init()) return false;
    size_t n = keys.size();
    if (n == 0) return true;
    if (vals.size() < n || out.size() < n) {
        TRTC_ERR("Scan_By_Key: values/output shorter than keys");
        return false;
    }
    std::string V = canon_type(out.name_elem_cls());
    size_t vsz = out.elem_size();
    ViewBuf initv;
    if (!init_value(init, vals.name_elem_cls(), out.name_elem_cls(), initv)) return false;
    std::string name = "trtc_scan_by_key_k<" + vcls(keys) + ", " + vcls(vals) + ", " + vcls(out) + ", " + vcls(pred) +
                       ", " + vcls(op) + ", " + V + ">";
    KernelRef k;
    if (!get_template_kernel(name, name, k)) return false;
    hipStream_t s = cur_stream();
    const unsigned block = wide_block(std::max(keys.elem_size(), vsz));
    size_t tile = (size_t)block * tile_items_for(std::max(keys.elem_size(), vsz));
    size_t tiles = (n + tile - 1) / tile;
    size_t seg_bytes = align_up(vsz + 4, std::min<size_t>(std::max<size_t>(vsz, 4), 8));
    LookbackScratch lb;
    if (!lb.make(tiles, seg_bytes, s)) return false;
    trtc::ScanByKeyArgs a{n, {lb.agg, lb.incl, lb.counter}, exclusive ? 1 : 0};
    Args args;
    args.view(keys);
    args.view(vals);
    args.view(out);
    args.view(pred);
    args.view(op);
    args.raw(initv.data(), initv.size());
    args.val(a);
    return launch_kernel(k, dim3((unsigned)tiles), dim3(block), args.data(), 0, s);
}

}  // namespace

namespace trtc_host {
bool reduce_view(const DVVectorLike& vec, const DeviceViewable* init, const Functor& op, ViewBuf& ret)
{
    return general_reduce(vec, init, op, ret);
}

bool extreme_view(const DVVectorLike& vec, const Functor& comp, bool max, size_t& id, bool& handled)
{
    return general_extreme(vec, comp, max, id, handled);
}

bool inclusive_scan_fixed_scratch(const DVVectorLike& vin, DVVectorLike& vout, const Functor& op, void* scratch,
                                  size_t scratch_bytes)
{
    if (!scratch) return false;
    return general_scan(vin, vout, op, nullptr, false, scratch, scratch_bytes);
}

// Grid of the long-tile pass (trtc_radix_pass_long_k, trtc_seg_radix_pass_long_k):
// its workgroups take tiles until none is left, tile runs going to virtual
// XCD blockIdx.x mod 8, so the grid is a multiple of 8 (every virtual XCD has
// workgroups) and at most one workgroup per CU (the pass holds 145 KB of LDS):
// a workgroup moves to its next tile without a launch in between.
size_t long_pass_grid(size_t tiles)
{
    const size_t cus = (size_t)std::max(8, context().num_cus);
    const size_t want = std::min((tiles + 7) / 8, cus / 8);
    return 8 * std::max<size_t>(1, want);
}

bool radix_sort_from(const void* keys_src, const void* vals_src, void* keys, void* vals, size_t n, int key_bytes,
                     int kind, bool descending, int val_bytes, hipStream_t s)
{
    return radix_sort(keys, vals, n, key_bytes, kind, descending, val_bytes, s, keys_src, vals_src);
}
}  // namespace trtc_host

// ================================================================ public API
bool TRTC_Sort(DVVectorLike& vec, const Functor& comp) { return sort_dispatch(vec, nullptr, comp); }
bool TRTC_Sort(DVVectorLike& vec) { return TRTC_Sort(vec, Functor("Less")); }
bool TRTC_Sort_By_Key(DVVectorLike& keys, DVVectorLike& values, const Functor& comp)
{
    return sort_dispatch(keys, &values, comp);
}
bool TRTC_Sort_By_Key(DVVectorLike& keys, DVVectorLike& values)
{
    return TRTC_Sort_By_Key(keys, values, Functor("Less"));
}

bool TRTC_Is_Sorted(const DVVectorLike& vec, const Functor& comp, bool& result)
{
    // reference sort.cpp:67-83
    result = true;
    if (vec.size() <= 1) return true;
    static TRTC_For s_for({"vec", "comp", "res"}, "idx", "    if (comp(vec[idx+1], vec[idx])) res[0] = false;\n");
    DVVector dvres("bool", 1, &result);
    const DeviceViewable* args[] = {&vec, &comp, &dvres};
    if (!s_for.launch_n(vec.size() - 1, args)) return false;
    dvres.to_host(&result);
    return true;
}
bool TRTC_Is_Sorted(const DVVectorLike& vec, bool& result) { return TRTC_Is_Sorted(vec, Functor("Less"), result); }

bool TRTC_Inclusive_Scan(const DVVectorLike& vec_in, DVVectorLike& vec_out, const Functor& binary_op)
{
    return general_scan(vec_in, vec_out, binary_op, nullptr, false);
}
bool TRTC_Inclusive_Scan(const DVVectorLike& vec_in, DVVectorLike& vec_out)
{
    return TRTC_Inclusive_Scan(vec_in, vec_out, Functor("Plus"));
}
bool TRTC_Exclusive_Scan(const DVVectorLike& vec_in, DVVectorLike& vec_out)
{
    return general_scan(vec_in, vec_out, Functor("Plus"), nullptr, true);
}
bool TRTC_Exclusive_Scan(const DVVectorLike& vec_in, DVVectorLike& vec_out, const DeviceViewable& init,
                         const Functor& binary_op)
{
    return general_scan(vec_in, vec_out, binary_op, &init, true);
}
bool TRTC_Exclusive_Scan(const DVVectorLike& vec_in, DVVectorLike& vec_out, const DeviceViewable& init)
{
    return TRTC_Exclusive_Scan(vec_in, vec_out, init, Functor("Plus"));
}

bool TRTC_Inclusive_Scan_By_Key(const DVVectorLike& vec_key, const DVVectorLike& vec_value, DVVectorLike& vec_out,
                                const Functor& binary_pred, const Functor& binary_op)
{
    return general_scan_by_key(vec_key, vec_value, vec_out, binary_pred, binary_op, nullptr, false);
}
bool TRTC_Inclusive_Scan_By_Key(const DVVectorLike& vec_key, const DVVectorLike& vec_value, DVVectorLike& vec_out,
                                const Functor& binary_pred)
{
    return TRTC_Inclusive_Scan_By_Key(vec_key, vec_value, vec_out, binary_pred, Functor("Plus"));
}
bool TRTC_Inclusive_Scan_By_Key(const DVVectorLike& vec_key, const DVVectorLike& vec_value, DVVectorLike& vec_out)
{
    return TRTC_Inclusive_Scan_By_Key(vec_key, vec_value, vec_out, Functor("EqualTo"));
}
bool TRTC_Exclusive_Scan_By_Key(const DVVectorLike& vec_key, const DVVectorLike& vec_value, DVVectorLike& vec_out)
{
    return general_scan_by_key(vec_key, vec_value, vec_out, Functor("EqualTo"), Functor("Plus"), nullptr, true);
}
bool TRTC_Exclusive_Scan_By_Key(const DVVectorLike& vec_key, const DVVectorLike& vec_value, DVVectorLike& vec_out,
                                const DeviceViewable& init, const Functor& binary_pred, const Functor& binary_op)
{
    return general_scan_by_key(vec_key, vec_value, vec_out, binary_pred, binary_op, &init, true);
}
bool TRTC_Exclusive_Scan_By_Key(const DVVectorLike& vec_key, const DVVectorLike& vec_value, DVVectorLike& vec_out,
                                const DeviceViewable& init, const Functor& binary_pred)
{
    return TRTC_Exclusive_Scan_By_Key(vec_key, vec_value, vec_out, init, binary_pred, Functor("Plus"));
}
bool TRTC_Exclusive_Scan_By_Key(const DVVectorLike& vec_key, const DVVectorLike& vec_value, DVVectorLike& vec_out,
                                const DeviceViewable& init)
{
    return TRTC_Exclusive_Scan_By_Key(vec_key, vec_value, vec_out, init, Functor("EqualTo"));
}

bool TRTC_Reduce(const DVVectorLike& vec, ViewBuf& ret) { return general_reduce(vec, nullptr, Functor("Plus"), ret); }
bool TRTC_Reduce(const DVVectorLike& vec, const DeviceViewable& init, const Functor& binary_op, ViewBuf& ret)
{
    return general_reduce(vec, &init, binary_op, ret);
}
bool TRTC_Reduce(const DVVectorLike& vec, const DeviceViewable& init, ViewBuf& ret)
{
    return general_reduce(vec, &init, Functor("Plus"), ret);
}

uint32_t TRTC_Copy_If(const DVVectorLike& vec_in, DVVectorLike& vec_out, const Functor& pred)
{
    return general_copy_if(vec_in, nullptr, vec_out, pred);
}
uint32_t TRTC_Copy_If_Stencil(const DVVectorLike& vec_in, const DVVectorLike& vec_stencil, DVVectorLike& vec_out,
                              const Functor& pred)
{
    return general_copy_if(vec_in, &vec_stencil, vec_out, pred);
}

// Element-wise converting copy (reference copy.cpp:3-12).
bool TRTC_Copy(const DVVectorLike& vec_in, DVVectorLike& vec_out)
{
    if (!ensure_init()) return false;
    size_t n = vec_in.size();
    if (n == 0) return true;
    if (vec_out.size() < n) {
        TRTC_ERR("Copy: output has %zu items, %zu needed", vec_out.size(), n);
        return false;
    }
    void* src = contiguous_data(vec_in);
    void* dst = contiguous_data(vec_out);
    // Same-type contiguous copies: 16-byte aligned ones run trtc_copy_bytes16
    // (one 16-byte vector per thread), others of 1- or 2-byte items go to
    // hipMemcpyAsync; everything else runs the element-wise map (8 reads in
    // flight per thread: 2^28 int32 in 0.370 ms, int32 -> float32 in
    // 0.382 ms; hipMemcpyAsync takes ~0.43 ms for the same 1 GiB).
    if (src && dst && canon_type(vec_in.name_elem_cls()) == canon_type(vec_out.name_elem_cls())) {
        const size_t nbytes = n * vec_in.elem_size();
        if (((((size_t)src | (size_t)dst) & 15) == 0) || vec_in.elem_size() < 4)
            return copy_bytes(src, dst, nbytes, cur_stream());
    }
    return run_map({"view_vec_in", "view_vec_out"}, "(decltype(view_vec_out)::value_t)view_vec_in[idx]",
                   "view_vec_out[idx] = _x;", n, {&vec_in, &vec_out});
}

// ---------------------------------------------------------------- scan / compaction consumers

// Reduce_By_Key (reduce.cpp:36-62)
uint32_t TRTC_Reduce_By_Key(const DVVectorLike& key_in, const DVVectorLike& value_in, DVVectorLike& key_out,
                            DVVectorLike& value_out)
{
    return general_reduce_by_key(key_in, value_in, key_out, value_out, Functor("EqualTo"), Functor("Plus"));
}
uint32_t TRTC_Reduce_By_Key(const DVVectorLike& key_in, const DVVectorLike& value_in, DVVectorLike& key_out,
                            DVVectorLike& value_out, const Functor& binary_pred)
{
    return general_reduce_by_key(key_in, value_in, key_out, value_out, binary_pred, Functor("Plus"));
}
uint32_t TRTC_Reduce_By_Key(const DVVectorLike& key_in, const DVVectorLike& value_in, DVVectorLike& key_out,
                            DVVectorLike& value_out, const Functor& binary_pred, const Functor& binary_op)
{
    return general_reduce_by_key(key_in, value_in, key_out, value_out, binary_pred, binary_op);
}

// Unique family (unique.cpp:5-95).  The default comparison keeps i when
// key[i] != key[i-1], which is !EqualTo for every arithmetic type (NaN
// included).  The in-place forms compact in place.
uint32_t TRTC_Unique(DVVectorLike& vec) { return general_unique(vec, nullptr, vec, nullptr, Functor("EqualTo")); }
uint32_t TRTC_Unique(DVVectorLike& vec, const Functor& binary_pred)
{
    return general_unique(vec, nullptr, vec, nullptr, binary_pred);
}
uint32_t TRTC_Unique_Copy(const DVVectorLike& vec_in, DVVectorLike& vec_out)
{
    return general_unique(vec_in, nullptr, vec_out, nullptr, Functor("EqualTo"));
}
uint32_t TRTC_Unique_Copy(const DVVectorLike& vec_in, DVVectorLike& vec_out, const Functor& binary_pred)
{
    return general_unique(vec_in, nullptr, vec_out, nullptr, binary_pred);
}
uint32_t TRTC_Unique_By_Key(DVVectorLike& keys, DVVectorLike& values)
{
    return general_unique(keys, &values, keys, &values, Functor("EqualTo"));
}
uint32_t TRTC_Unique_By_Key(DVVectorLike& keys, DVVectorLike& values, const Functor& binary_pred)
{
    return general_unique(keys, &values, keys, &values, binary_pred);
}
uint32_t TRTC_Unique_By_Key_Copy(const DVVectorLike& keys_in, const DVVectorLike& values_in, DVVectorLike& keys_out,
                                 DVVectorLike& values_out)
{
    return general_unique(keys_in, &values_in, keys_out, &values_out, Functor("EqualTo"));
}
uint32_t TRTC_Unique_By_Key_Copy(const DVVectorLike& keys_in, const DVVectorLike& values_in, DVVectorLike& keys_out,
                                 DVVectorLike& values_out, const Functor& binary_pred)
{
    return general_unique(keys_in, &values_in, keys_out, &values_out, binary_pred);
}

// Partition family (partition.cpp:28-130)
uint32_t TRTC_Partition(DVVectorLike& vec, const Functor& pred) { return general_partition(vec, nullptr, pred); }
uint32_t TRTC_Partition_Stencil(DVVectorLike& vec, const DVVectorLike& stencil, const Functor& pred)
{
    return general_partition(vec, &stencil, pred);
}
uint32_t TRTC_Partition_Copy(const DVVectorLike& vec_in, DVVectorLike& vec_true, DVVectorLike& vec_false,
                             const Functor& pred)
{
    return general_partition_copy(vec_in, nullptr, vec_true, vec_false, pred);
}
uint32_t TRTC_Partition_Copy_Stencil(const DVVectorLike& vec_in, const DVVectorLike& stencil, DVVectorLike& vec_true,
                                     DVVectorLike& vec_false, const Functor& pred)
{
    return general_partition_copy(vec_in, &stencil, vec_true, vec_false, pred);
}

// Remove family (remove.cpp:5-60): compaction of the items to keep (x !=
// value, or !pred), in place for the in-place forms.
namespace {
Functor keep_unequal(const DeviceViewable& value) { return Functor({{"value", &value}}, {"x"}, "        return x != value;\n"); }
Functor keep_not(const Functor& pred) { return Functor({{"pred", &pred}}, {"x"}, "        return !pred(x);\n"); }
}  // namespace

uint32_t TRTC_Remove(DVVectorLike& vec, const DeviceViewable& value)
{
    return general_copy_if(vec, nullptr, vec, keep_unequal(value));
}
uint32_t TRTC_Remove_Copy(const DVVectorLike& vec_in, DVVectorLike& vec_out, const DeviceViewable& value)
{
    return general_copy_if(vec_in, nullptr, vec_out, keep_unequal(value));
}
uint32_t TRTC_Remove_If(DVVectorLike& vec, const Functor& pred)
{
    return general_copy_if(vec, nullptr, vec, keep_not(pred));
}
uint32_t TRTC_Remove_Copy_If(const DVVectorLike& vec_in, DVVectorLike& vec_out, const Functor& pred)
{
    return general_copy_if(vec_in, nullptr, vec_out, keep_not(pred));
}
uint32_t TRTC_Remove_If_Stencil(DVVectorLike& vec, const DVVectorLike& stencil, const Functor& pred)
{
    return general_copy_if(vec, &stencil, vec, keep_not(pred));
}
uint32_t TRTC_Remove_Copy_If_Stencil(const DVVectorLike& vec_in, const DVVectorLike& stencil, DVVectorLike& vec_out,
                                     const Functor& pred)
{
    return general_copy_if(vec_in, &stencil, vec_out, keep_not(pred));
}

// Transform scans (transform_scan.cpp:4-16): the scans of a TransformView whose
// elements are unary_op(vec_in[i]) converted to the output element type.
bool TRTC_Transform_Inclusive_Scan(const DVVectorLike& vec_in, DVVectorLike& vec_out, const Functor& unary_op,
                                   const Functor& binary_op)
{
    DVTransform src(vec_in, vec_out.name_elem_cls().c_str(), unary_op);
    return general_scan(src, vec_out, binary_op, nullptr, false);
}
bool TRTC_Transform_Exclusive_Scan(const DVVectorLike& vec_in, DVVectorLike& vec_out, const Functor& unary_op,
                                   const DeviceViewable& init, const Functor& binary_op)
{
    DVTransform src(vec_in, vec_out.name_elem_cls().c_str(), unary_op);
    return general_scan(src, vec_out, binary_op, &init, true);
}

// Count / Count_If (count.cpp:4-26): sums of 0/1 indicators in 64 bits.
bool TRTC_Count(const DVVectorLike& vec, const DeviceViewable& value, size_t& ret)
{
    ret = 0;
    if (vec.size() == 0) return true;
    Functor ind({{"eq_value", &value}}, {"x"},
                ("        return (x == (" + vec.name_elem_cls() + ")eq_value) ? (uint64_t)1 : (uint64_t)0;\n").c_str());
    DVTransform src(vec, "uint64_t", ind);
    ViewBuf buf;
    if (!general_reduce(src, nullptr, Functor("Plus"), buf)) return false;
    memcpy(&ret, buf.data(), sizeof(size_t));
    return true;
}
bool TRTC_Count_If(const DVVectorLike& vec, const Functor& pred, size_t& ret)
{
    ret = 0;
    if (vec.size() == 0) return true;
    Functor ind({{"pred", &pred}}, {"x"}, "        return pred(x) ? (uint64_t)1 : (uint64_t)0;\n");
    DVTransform src(vec, "uint64_t", ind);
    ViewBuf buf;
    if (!general_reduce(src, nullptr, Functor("Plus"), buf)) return false;
    memcpy(&ret, buf.data(), sizeof(size_t));
    return true;
}

// Transform_Reduce (transform_reduce.cpp:4-15): init op u(v0) op u(v1) ... in
// the type of init.
bool TRTC_Transform_Reduce(const DVVectorLike& vec, const Functor& unary_op, const DeviceViewable& init,
                           const Functor& binary_op, ViewBuf& ret)
{
    DVTransform src(vec, init.name_view_cls().c_str(), unary_op);
    return general_reduce(src, &init, binary_op, ret);
}
