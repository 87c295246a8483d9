// Batched-segment engine: plans, descriptor tables and launches for many
// independent sorts, scans and reductions per kernel stage, and the schedule
// autotuner of the batched sort.
#include "trtc_batch.h"

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <map>
#include <mutex>

namespace trtc_host {

namespace {
size_t align256(size_t x) { return (x + 255) / 256 * 256; }
size_t ceil_div(size_t a, size_t b) { return (a + b - 1) / b; }
const size_t kNoLimit = ~(size_t)0;
}  // namespace

// Host table -> device, asynchronously on stream s: staged in pinned memory
// and copied by the 16-byte copy kernel reading the pinned buffer directly.
// (hipMemcpyAsync from the same pinned buffer returned only once the copy
// ran for tables of a few hundred KB, ~6 ms while another stream's sort held
// every CU, with the GPU idle behind this batch meanwhile.)
bool upload_async(void* dst, const void* src, size_t bytes, hipStream_t s)
{
    if (!bytes) return true;
    void* st = pinned_alloc(bytes);
    if (!st) return false;
    memcpy(st, src, bytes);
    void* dev_src = nullptr;
    bool ok;
    if (hipHostGetDevicePointer(&dev_src, st, 0) == hipSuccess && dev_src)
        ok = copy_bytes(dev_src, dst, bytes, s);
    else {
        (void)hipGetLastError();
        ok = check_hip(hipMemcpyAsync(dst, st, bytes, hipMemcpyHostToDevice, s), "hipMemcpyAsync");
    }
    pinned_release(st, s);
    return ok;
}

// ================================================================ sort
bool radix_key_info(const std::string& elem, int& key_bytes, int& kind)
{
    const ScalarInfo* si = scalar_info(elem);
    if (!si) return false;
    switch (si->code) {
    case TC_I8: case TC_I16: case TC_I32: case TC_I64: kind = 1; break;
    case TC_U8: case TC_U16: case TC_U32: case TC_U64: case TC_BOOL: kind = 0; break;
    case TC_F32: case TC_F64: kind = 2; break;
    default: return false;
    }
    key_bytes = (int)si->size;
    return true;
}

std::string describe(const SortSchedule& s)
{
    std::string r;
    for (size_t i = 0; i < s.size(); i++) {
        char b[96];
        if (s[i].max_n == kNoLimit)
            snprintf(b, sizeof(b), "%srest:%dx%d", i ? " " : "", s[i].block, s[i].items);
        else
            snprintf(b, sizeof(b), "%s<=%zu:%dx%d", i ? " " : "", s[i].max_n, s[i].block, s[i].items);
        r += b;
    }
    return r;
}

SortSchedule default_sort_schedule(int key_bytes, int val_bytes)
{
    if (key_bytes > 4 || val_bytes > 4) return {{kNoLimit, 512, 8}};
    return {{kNoLimit, 512, 16}};
}

// The long-tile pass (1024 x 32 keys, trtc_seg_radix_pass_long_k) exists for
// 4-byte keys with 4-byte or no values.
static bool long_tiles_ok(int key_bytes, int val_bytes) { return key_bytes == 4 && (val_bytes == 0 || val_bytes == 4); }
static bool is_long_class(int key_bytes, int val_bytes, int block, int items)
{
    return long_tiles_ok(key_bytes, val_bytes) && block == 1024 && items == 32;
}

// Size classes for batches whose shape changes from call to call (the request
// server): small segments in small tiles, segments above 2^18 keys in the
// tiles of large single sorts (the autotuner's choice on log-uniform sizes,
// e.g. the 4096-segment batch).
SortSchedule served_sort_schedule(int key_bytes, int val_bytes)
{
    if (key_bytes > 4 || val_bytes > 4) return {{2048, 256, 8}, {1u << 18, 512, 8}, {kNoLimit, 1024, 8}};
    if (long_tiles_ok(key_bytes, val_bytes)) return {{4096, 256, 16}, {1u << 18, 512, 16}, {kNoLimit, 1024, 32}};
    return {{4096, 256, 16}, {1u << 18, 512, 16}, {kNoLimit, 1024, 16}};
}

// The segment's offset from a 128-byte line in keys.  In every pass, tile 0
// of a multi-tile segment gives up the lead of that pass's input buffer
// (computed by the pass from the pointer), so that its later tiles start on
// line boundaries: a 256-byte load instruction of the pass then covers two
// lines, not three.  Per 1024x16 pass over 128 segments of u32/u32 pairs
// 2^20 - 4 keys apart (16-byte but not line aligned): 0.648 ms without the
// lead, 0.596 ms with it (0.587 ms for segments 2^20 keys apart).  Pass 0
// reads the request's source, which a server's requests place anywhere.
constexpr size_t kLineBytes = 128;
unsigned seg_lead(const SortSeg& g, int KB)
{
    return (unsigned)(((uintptr_t)g.keys & (kLineBytes - 1)) / (unsigned)KB);
}

// Tiles of a segment: a multi-tile segment has room for the largest lead of
// any pass (a pass with a smaller one may leave its last tile empty).
size_t seg_tiles(const SortSeg& g, size_t tile, int KB)
{
    const size_t max_lead = kLineBytes / (size_t)KB - 1;
    return g.n > tile ? ceil_div(g.n + max_lead, tile) : ceil_div(g.n, tile);
}

bool plan_sort(const std::vector<SortSeg>& segs, int KB, int VB, int kind, bool descending,
               const SortSchedule& sched, SortPlan& plan)
{
    if (!(KB == 1 || KB == 2 || KB == 4 || KB == 8) || !(VB == 0 || VB == 1 || VB == 2 || VB == 4 || VB == 8)) {
        TRTC_ERR("batched sort: unsupported key/value widths %d/%d", KB, VB);
        return false;
    }
    if (sched.empty()) {
        TRTC_ERR("batched sort: empty schedule");
        return false;
    }
    plan = SortPlan();
    plan.KB = KB;
    plan.VB = VB;
    plan.kind = kind;
    plan.desc = descending;
    plan.cls.resize(sched.size());
    for (size_t c = 0; c < sched.size(); c++) {
        if (sched[c].items == 32 && !is_long_class(KB, VB, sched[c].block, sched[c].items)) {
            TRTC_ERR("batched sort: %dx%d tiles are for 4-byte keys with 4-byte or no values (1024 x 32)",
                     sched[c].block, sched[c].items);
            return false;
        }
        plan.cls[c].block = sched[c].block;
        plan.cls[c].items = sched[c].items;
        plan.cls[c].tile = (size_t)sched[c].block * sched[c].items;
    }
    for (const SortSeg& g : segs) {
        if (g.n == 0 || (g.n == 1 && !g.src_keys)) continue;  // nothing to move
        if (g.n >= (1ull << 30)) {
            TRTC_ERR("batched sort: segment of %zu keys (limit 2^30 - 1); sort it with TRTC_Sort_By_Key", g.n);
            return false;
        }
        size_t c = 0;
        while (c + 1 < sched.size() && g.n > sched[c].max_n) c++;
        plan.cls[c].segs.push_back(g);
        plan.keys_total += g.n;
    }
    // histogram chunk: 16K..1M keys (16 to 1024 per thread of the 1024-thread
    // workgroups), small enough that the batch's multi-tile keys give every
    // CU about four workgroups; each chunk costs ~12 us of its own (counter
    // clearing, column sums, global adds), so large batches take large chunks
    // (a served batch of 1e9 keys: 8227 chunks of 128K keys took 1.2 ms)
    {
        size_t multi_keys = 0;
        for (auto& C : plan.cls)
            for (auto& g : C.segs)
                if (g.n > C.tile) multi_keys += g.n;
        size_t want = multi_keys / ((size_t)std::max(1, context().num_cus) * 4);
        size_t ck = 16384;
        while (ck < want && ck < ((size_t)1 << 20)) ck <<= 1;
        plan.hist_chunk = ck;
    }
    size_t off = 0;
    for (auto& C : plan.cls) {
        // multi-tile segments first, largest first (they start earliest and
        // the one-tile segments fill in behind them); stable otherwise
        const size_t tile = C.tile;
        std::stable_sort(C.segs.begin(), C.segs.end(), [tile](const SortSeg& a, const SortSeg& b) {
            bool ma = a.n > tile, mb = b.n > tile;
            if (ma != mb) return ma;
            return ma && a.n > b.n;
        });
        C.tmp_off.assign(C.segs.size(), 0);
        for (size_t i = 0; i < C.segs.size(); i++) {
            size_t nt = seg_tiles(C.segs[i], tile, KB);
            C.num_tiles += (unsigned)nt;
            if (nt > 1) {
                C.multi_tiles += (unsigned)nt;
                C.multi_segs++;
                C.hist_chunks += (unsigned)ceil_div(C.segs[i].n, plan.hist_chunk);
                // the ping-pong copy is as far past a line boundary as the
                // segment's own buffers (the ping-pong buffers themselves are
                // 256-byte aligned), so passes 1-3 give up the same lead
                const size_t per = kLineBytes / (size_t)KB;
                const size_t at = plan.tmp_elems;
                C.tmp_off[i] = at + (seg_lead(C.segs[i], KB) + per - at % per) % per;
                plan.tmp_elems = C.tmp_off[i] + C.segs[i].n;
            }
        }
        C.off_descs = off;
        off += align256(C.segs.size() * sizeof(trtc::SegSortDesc));
        C.off_tile_seg = off;
        off += align256((size_t)C.num_tiles * 4);
        C.off_hist_chunk = off;
        off += align256((size_t)C.hist_chunks * 8);
    }
    plan.table_bytes = off;
    for (auto& C : plan.cls) {
        C.off_zero = off;
        C.off_cnt = off;
        off += 256;
        C.off_done = off;
        off += align256((size_t)C.multi_segs * 4);
        C.off_hist = off;
        off += align256((size_t)C.multi_segs * KB * 256 * 4);
        C.off_status = off;
        C.status_pass_words = (size_t)C.multi_tiles * 256;
        off += align256((size_t)KB * C.status_pass_words * 4);
        C.zero_bytes = off - C.off_zero;
        C.off_dstart = off;
        off += align256((size_t)C.multi_segs * KB * 256 * 4);
        C.off_rank = off;  // written by the histogram pre-pass for every multi-tile segment
        off += align256((size_t)C.multi_segs * 4);
    }
    plan.scratch_bytes = off;
    return true;
}

void fill_sort_tables(const SortPlan& plan, char* scratch, char* tmp_keys, char* tmp_vals, char* table)
{
    (void)scratch;
    for (const auto& C : plan.cls) {
        trtc::SegSortDesc* D = (trtc::SegSortDesc*)(table + C.off_descs);
        unsigned* TS = (unsigned*)(table + C.off_tile_seg);
        unsigned* HC = (unsigned*)(table + C.off_hist_chunk);
        unsigned t = 0, m = 0, hc = 0;
        unsigned long long row = 0;
        for (size_t i = 0; i < C.segs.size(); i++) {
            const SortSeg& g = C.segs[i];
            size_t nt = seg_tiles(g, C.tile, plan.KB);
            trtc::SegSortDesc& d = D[i];
            d.keys = (unsigned long long)g.keys;
            d.vals = (unsigned long long)g.vals;
            d.src_keys = (unsigned long long)(g.src_keys ? g.src_keys : g.keys);
            d.src_vals = (unsigned long long)(g.src_keys ? g.src_vals : g.vals);
            d.n = g.n;
            d.first_tile = t;
            if (nt > 1) {
                d.tmp_keys = (unsigned long long)(tmp_keys + C.tmp_off[i] * plan.KB);
                d.tmp_vals = plan.VB ? (unsigned long long)(tmp_vals + C.tmp_off[i] * plan.VB) : 0ull;
                d.multi = m++;
                d.status_row = row;
                row += nt;
                d.hist_chunks = (unsigned)ceil_div(g.n, plan.hist_chunk);
                for (unsigned c = 0; c < d.hist_chunks; c++) {
                    HC[2 * hc] = (unsigned)i;
                    HC[2 * hc + 1] = c;
                    hc++;
                }
            } else {
                d.tmp_keys = d.keys;
                d.tmp_vals = d.vals;
                d.multi = 0xFFFFFFFFu;
                d.status_row = 0;
                d.hist_chunks = 0;
            }
            for (size_t k = 0; k < nt; k++) TS[t++] = (unsigned)i;
        }
    }
}

namespace {
std::string seg_hist_name(const SortPlan& plan) { return "trtc_seg_radix_hist_k<" + std::to_string(plan.KB) + ">"; }
std::string seg_pass_name(const SortPlan& plan, const SortPlan::Cls& C)
{
    if (is_long_class(plan.KB, plan.VB, C.block, C.items))
        return "trtc_seg_radix_pass_long_k<" + std::to_string(plan.VB) + ", " + std::to_string(plan.kind) + ">";
    return "trtc_seg_radix_pass_k<" + std::to_string(plan.KB) + ", " + std::to_string(plan.VB) + ", " +
           std::to_string(C.block) + ", " + std::to_string(C.items) + ", " + std::to_string(plan.kind) + ">";
}
std::string seg_scan_name(const ScanPlan& plan)
{
    return "trtc_seg_scan_k<" + plan.op_cls + ", " + plan.T + ", " + std::to_string(plan.items) + ">";
}
std::string seg_reduce_name(const ReducePlan& plan) { return "trtc_seg_reduce_k<" + plan.op_cls + ", " + plan.T + ">"; }
}  // namespace

bool prepare_sort(const SortPlan& plan)
{
    KernelRef k;
    for (const auto& C : plan.cls) {
        if (!C.num_tiles) continue;
        if (C.multi_tiles && !get_template_kernel(seg_hist_name(plan), "", k)) return false;
        if (!get_template_kernel(seg_pass_name(plan, C), "", k)) return false;
    }
    return true;
}

bool prepare_scan(const ScanPlan& plan)
{
    KernelRef k;
    return !plan.num_tiles || get_template_kernel(seg_scan_name(plan), seg_scan_name(plan), k);
}

bool prepare_reduce(const ReducePlan& plan)
{
    KernelRef k;
    return !plan.num_tiles || get_template_kernel(seg_reduce_name(plan), seg_reduce_name(plan), k);
}

bool enqueue_sort(const SortPlan& plan, char* scratch, hipStream_t s)
{
    for (const auto& C : plan.cls) {
        if (!C.num_tiles) continue;
        if (!fill_u32(scratch + C.off_zero, C.zero_bytes / 4, 0, s)) return false;
        trtc::SegRadixArgs a{};
        a.descs = (const trtc::SegSortDesc*)(scratch + C.off_descs);
        a.tile_seg = (const unsigned*)(scratch + C.off_tile_seg);
        a.num_tiles = C.num_tiles;
        a.num_multi_tiles = C.multi_tiles;
        a.tile_counter = (unsigned*)(scratch + C.off_cnt);
        a.status = (unsigned*)(scratch + C.off_status);
        a.seg_hist = (unsigned*)(scratch + C.off_hist);
        a.seg_dstart = (unsigned*)(scratch + C.off_dstart);
        a.seg_done = (unsigned*)(scratch + C.off_done);
        a.seg_rank_lds = (unsigned*)(scratch + C.off_rank);
        a.hist_chunk = (const unsigned*)(scratch + C.off_hist_chunk);
        a.num_hist_chunks = C.hist_chunks;
        a.hist_chunk_keys = (unsigned)plan.hist_chunk;
        a.pass = 0;
        a.shift = 0;
        a.kind = plan.kind;
        a.descending = plan.desc ? 1 : 0;
        if (C.multi_tiles) {
            KernelRef hk;
            if (!get_template_kernel(seg_hist_name(plan), "", hk)) return false;
            void* args[] = {&a};
            if (!launch_kernel(hk, dim3(C.hist_chunks), dim3(1024), args, 0, s)) return false;  // TRTC_HIST_BLOCK
        }
        KernelRef pk;
        if (!get_template_kernel(seg_pass_name(plan, C), "", pk)) return false;
        // the long-tile pass hands tiles to the 8 XCDs in runs of 8 (8 tile
        // counters per pass) to workgroups that loop until none is left
        const bool lt = is_long_class(plan.KB, plan.VB, C.block, C.items);
        const unsigned grid = lt ? (unsigned)long_pass_grid(C.num_tiles) : C.num_tiles;
        for (int p = 0; p < plan.KB; p++) {
            a.pass = p;
            a.shift = 8 * p;
            a.tile_counter = (unsigned*)(scratch + C.off_cnt) + (lt ? 8 * p : p);
            a.status = (unsigned*)(scratch + C.off_status) + p * C.status_pass_words;
            void* args[] = {&a};
            if (!launch_kernel(pk, dim3(grid), dim3(C.block), args, 0, s)) return false;
        }
    }
    return true;
}

namespace {
// Odd digit counts end in the ping-pong buffers for multi-tile segments.
bool copy_back_odd(const SortPlan& plan, char* tmp_keys, char* tmp_vals, hipStream_t s)
{
    if (!(plan.KB & 1)) return true;
    for (const auto& C : plan.cls)
        for (size_t i = 0; i < C.segs.size(); i++) {
            const SortSeg& g = C.segs[i];
            if (g.n <= C.tile) continue;
            if (!check_hip(hipMemcpyAsync(g.keys, tmp_keys + C.tmp_off[i] * plan.KB, g.n * plan.KB,
                                          hipMemcpyDeviceToDevice, s),
                           "hipMemcpyAsync"))
                return false;
            if (plan.VB && !check_hip(hipMemcpyAsync(g.vals, tmp_vals + C.tmp_off[i] * plan.VB, g.n * plan.VB,
                                                     hipMemcpyDeviceToDevice, s),
                                      "hipMemcpyAsync"))
                return false;
        }
    return true;
}
}  // namespace

bool launch_sort_plan(const SortPlan& plan, char* scratch, char* tmp_keys, char* tmp_vals, hipStream_t s)
{
    if (!plan.keys_total) return true;
    std::vector<char> table(plan.table_bytes);
    fill_sort_tables(plan, scratch, tmp_keys, tmp_vals, table.data());
    if (!upload_async(scratch, table.data(), table.size(), s)) return false;
    if (!enqueue_sort(plan, scratch, s)) return false;
    return copy_back_odd(plan, tmp_keys, tmp_vals, s);
}

bool run_sort_batch(const std::vector<SortSeg>& segs, int KB, int VB, int kind, bool descending,
                    const SortSchedule& sched, hipStream_t s)
{
    SortPlan plan;
    if (!plan_sort(segs, KB, VB, kind, descending, sched, plan)) return false;
    if (!plan.keys_total) return true;
    PoolBuf scratch(plan.scratch_bytes, s);
    PoolBuf tk(plan.tmp_elems * KB, s), tv(plan.tmp_elems * VB, s);
    if (!scratch.p || (plan.tmp_elems && !tk.p) || (plan.tmp_elems && VB && !tv.p)) return false;
    return launch_sort_plan(plan, (char*)scratch.p, (char*)tk.p, (char*)tv.p, s);
}

// ---------------------------------------------------------------- autotuner
namespace {
struct Tuner
{
    std::mutex mtx;
    bool on = true;
    bool forced = false;  // a fixed schedule set by the caller
    SortSchedule force;
    std::map<std::string, SortSchedule> best;
    std::vector<std::string> log;
};
Tuner& tuner()
{
    static Tuner* t = new Tuner();
    return *t;
}

// Batch shape: widths plus the number of segments per power-of-two size,
// each count rounded to a power of two (shapes that differ less share a
// schedule).
std::string signature(const std::vector<SortSeg>& segs, int KB, int VB)
{
    size_t cnt[64] = {0};
    for (auto& g : segs)
        if (g.n >= 2) {
            int b = 63 - __builtin_clzll((unsigned long long)g.n);
            cnt[b]++;
        }
    std::string sig = "k" + std::to_string(KB) + "v" + std::to_string(VB);
    for (int b = 0; b < 64; b++)
        if (cnt[b]) {
            int r = 63 - __builtin_clzll((unsigned long long)cnt[b]);
            sig += " " + std::to_string(b) + ":" + std::to_string(r);
        }
    return sig;
}

// Segments above 2^18 keys may take 1024-thread tiles (16384 keys for 4-byte
// pairs, or the 32768-key long tiles of the largest single sorts for 4-byte
// keys with 4-byte or no values); the tuner keeps them only where they win.
std::vector<SortSchedule> candidates(int KB, int VB)
{
    if (KB > 4 || VB > 4)
        return {{{kNoLimit, 512, 8}},
                {{kNoLimit, 256, 8}},
                {{2048, 256, 8}, {kNoLimit, 512, 8}},
                {{2048, 256, 8}, {1u << 18, 512, 8}, {kNoLimit, 1024, 8}}};
    std::vector<SortSchedule> c = {{{kNoLimit, 512, 16}},
                                   {{kNoLimit, 512, 8}},
                                   {{kNoLimit, 256, 16}},
                                   {{2048, 256, 8}, {kNoLimit, 512, 16}},
                                   {{4096, 256, 16}, {kNoLimit, 512, 16}},
                                   {{2048, 256, 8}, {1u << 18, 512, 16}, {kNoLimit, 1024, 16}},
                                   {{4096, 256, 16}, {1u << 18, 512, 16}, {kNoLimit, 1024, 16}}};
    if (long_tiles_ok(KB, VB)) {
        c.push_back({{2048, 256, 8}, {1u << 18, 512, 16}, {kNoLimit, 1024, 32}});
        c.push_back({{4096, 256, 16}, {1u << 18, 512, 16}, {kNoLimit, 1024, 32}});
    }
    return c;
}

// Copies the segments' data into a shadow buffer (adjacent segments in one
// copy) so that tuning runs never touch the caller's data.
bool fill_shadow(const std::vector<SortSeg>& segs, const std::vector<SortSeg>& shadow, int KB, int VB, hipStream_t s)
{
    size_t i = 0;
    while (i < segs.size()) {
        size_t j = i + 1, n = segs[i].n;
        while (j < segs.size() && !segs[j].src_keys && !segs[i].src_keys &&
               segs[j].keys == segs[j - 1].keys + segs[j - 1].n * KB &&
               (!VB || segs[j].vals == segs[j - 1].vals + segs[j - 1].n * VB)) {
            n += segs[j].n;
            j++;
        }
        if (n) {
            const char* sk = segs[i].src_keys ? segs[i].src_keys : segs[i].keys;
            const char* sv = segs[i].src_keys ? segs[i].src_vals : segs[i].vals;
            if (!check_hip(hipMemcpyAsync(shadow[i].keys, sk, n * KB, hipMemcpyDeviceToDevice, s), "hipMemcpyAsync"))
                return false;
            if (VB && !check_hip(hipMemcpyAsync(shadow[i].vals, sv, n * VB, hipMemcpyDeviceToDevice, s),
                                 "hipMemcpyAsync"))
                return false;
        }
        i = j;
    }
    return true;
}
}  // namespace

bool parse_schedule(const std::string& spec, SortSchedule& out)
{
    out.clear();
    size_t i = 0;
    while (i < spec.size()) {
        while (i < spec.size() && spec[i] == ' ') i++;
        if (i >= spec.size()) break;
        size_t j = spec.find(' ', i);
        std::string tok = spec.substr(i, j == std::string::npos ? std::string::npos : j - i);
        i = j == std::string::npos ? spec.size() : j;
        SortClass c;
        int b = 0, it = 0;
        unsigned long long mx = 0;
        if (sscanf(tok.c_str(), "rest:%dx%d", &b, &it) == 2) {
            c.max_n = kNoLimit;
        } else if (sscanf(tok.c_str(), "<=%llu:%dx%d", &mx, &b, &it) == 3) {
            c.max_n = (size_t)mx;
        } else {
            return false;
        }
        if (!(b == 256 || b == 512 || b == 1024) || !(it == 4 || it == 8 || it == 16 || (it == 32 && b == 1024)))
            return false;
        c.block = b;
        c.items = it;
        out.push_back(c);
    }
    if (out.empty() || out.back().max_n != kNoLimit) return false;
    for (size_t k = 1; k < out.size(); k++)
        if (out[k - 1].max_n >= out[k].max_n) return false;
    return true;
}

bool set_forced_schedule(const char* spec)
{
    Tuner& T = tuner();
    std::lock_guard<std::mutex> lk(T.mtx);
    if (!spec || !*spec) {
        T.forced = false;
        return true;
    }
    SortSchedule s;
    if (!parse_schedule(spec, s)) {
        TRTC_ERR("batch schedule '%s' not understood (e.g. \"<=2048:256x8 rest:512x16\")", spec);
        return false;
    }
    T.forced = true;
    T.force = s;
    return true;
}

void set_autotune(bool on)
{
    std::lock_guard<std::mutex> lk(tuner().mtx);
    tuner().on = on;
}
bool autotune_on()
{
    std::lock_guard<std::mutex> lk(tuner().mtx);
    return tuner().on;
}

std::string tuning_report()
{
    Tuner& T = tuner();
    std::lock_guard<std::mutex> lk(T.mtx);
    std::string js = "[";
    for (size_t i = 0; i < T.log.size(); i++) js += (i ? ", " : "") + T.log[i];
    return js + "]";
}

SortSchedule schedule_for(const std::vector<SortSeg>& segs, int KB, int VB, int kind, bool descending, hipStream_t s)
{
    Tuner& T = tuner();
    std::string sig = signature(segs, KB, VB);
    {
        std::lock_guard<std::mutex> lk(T.mtx);
        if (T.forced) return T.force;
        auto it = T.best.find(sig);
        if (it != T.best.end()) return it->second;
        if (!T.on) return default_sort_schedule(KB, VB);
    }
    // shadow copy of the data, laid out like the segments
    std::vector<SortSeg> live;
    size_t total = 0;
    for (auto& g : segs)
        if (g.n >= 2) {
            live.push_back(g);
            total += g.n;
        }
    if (!total) return default_sort_schedule(KB, VB);
    PoolBuf sk(total * KB, s), sv(total * VB, s);
    if (!sk.p || (VB && !sv.p)) return default_sort_schedule(KB, VB);
    std::vector<SortSeg> shadow(live.size());
    size_t off = 0;
    for (size_t i = 0; i < live.size(); i++) {
        shadow[i] = SortSeg();
        shadow[i].keys = (char*)sk.p + off * KB;
        shadow[i].vals = VB ? (char*)sv.p + off * VB : nullptr;
        shadow[i].n = live[i].n;
        off += live[i].n;
    }
    hipEvent_t e0, e1;
    hipEventCreate(&e0);
    hipEventCreate(&e1);
    SortSchedule best = default_sort_schedule(KB, VB);
    float best_ms = 1e30f;
    std::string entry = "{\"signature\": \"" + sig + "\", \"segments\": " + std::to_string(live.size()) +
                        ", \"keys\": " + std::to_string(total) + ", \"candidates\": [";
    bool first = true;
    for (const SortSchedule& cand : candidates(KB, VB)) {
        float ms_min = 1e30f;
        bool ok = true;
        for (int rep = 0; rep < 3 && ok; rep++) {  // rep 0 also compiles / warms up
            ok = fill_shadow(live, shadow, KB, VB, s);
            ok = ok && hipEventRecord(e0, s) == hipSuccess;
            ok = ok && run_sort_batch(shadow, KB, VB, kind, descending, cand, s);
            ok = ok && hipEventRecord(e1, s) == hipSuccess && hipEventSynchronize(e1) == hipSuccess;
            float ms = 0;
            if (ok && hipEventElapsedTime(&ms, e0, e1) == hipSuccess && rep > 0) ms_min = std::min(ms_min, ms);
        }
        if (!ok) continue;
        char b[160];
        snprintf(b, sizeof(b), "%s{\"schedule\": \"%s\", \"ms\": %.4f}", first ? "" : ", ", describe(cand).c_str(),
                 ms_min);
        entry += b;
        first = false;
        if (ms_min < best_ms) {
            best_ms = ms_min;
            best = cand;
        }
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    entry += "], \"chosen\": \"" + describe(best) + "\"}";
    std::lock_guard<std::mutex> lk(T.mtx);
    T.best[sig] = best;
    T.log.push_back(entry);
    return best;
}

// ================================================================ scan
// Tiles of the single-input scan (ScanBlock x ScanItems in the device
// header): 512 x 32 for elements of up to 4 bytes, 1024 x 16 for 8 bytes.
// On the 4096-segment batch (3.9e8 int32) 256 x 16 tiles took 1.41 ms.
int default_scan_items(size_t tsz) { return tsz <= 4 ? 32 : 16; }
unsigned seg_scan_block(size_t tsz) { return tsz == 8 ? 1024u : 512u; }

// Elements that tile 0 of a multi-tile segment gives up, so that its later
// tiles start on 128-byte lines (as seg_lead for sorts).  Tile 0 still loads
// its whole extent through 16-byte pieces: the rest of the segment waits in
// its lookback for tile 0, and element loads there took the 4096-segment
// scan from 0.81 to 1.33 ms (0.78 ms as done now).
unsigned scan_seg_lead(const ScanSeg& g, size_t tile, size_t tsz)
{
    return g.n > tile ? (unsigned)(((uintptr_t)g.in & (kLineBytes - 1)) / tsz) : 0u;
}

size_t scan_seg_tiles(const ScanSeg& g, size_t tile, size_t tsz)
{
    return ceil_div(g.n + scan_seg_lead(g, tile, tsz), tile);
}

bool plan_scan(const std::vector<ScanSeg>& segs, const std::string& T, const Functor& op, const ViewBuf& init,
               bool exclusive, int items, ScanPlan& plan)
{
    plan = ScanPlan();
    const ScalarInfo* si = scalar_info(T);
    if (!si) {
        TRTC_ERR("batched scan: element type %s is not a scalar", T.c_str());
        return false;
    }
    plan.T = si->canon;
    plan.tsz = si->size;
    plan.op_cls = canon_type(op.name_view_cls());
    plan.op_view = op.view();
    plan.init = init;
    plan.init.resize(plan.tsz, 0);
    plan.exclusive = exclusive;
    plan.items = items;
    plan.tile = (size_t)seg_scan_block(plan.tsz) * items;
    plan.words = (plan.tsz + 3) / 4;
    for (auto& g : segs)
        if (g.n) {
            plan.segs.push_back(g);
            plan.num_tiles += (unsigned)scan_seg_tiles(g, plan.tile, plan.tsz);
        }
    size_t off = 0;
    plan.off_descs = off;
    off += align256(plan.segs.size() * sizeof(trtc::SegScanDesc));
    plan.off_tile_seg = off;
    off += align256((size_t)plan.num_tiles * 4);
    plan.table_bytes = off;
    plan.off_zero = off;
    off += 256 + align256(2 * (size_t)plan.num_tiles * plan.words * 8);
    plan.zero_bytes = off - plan.off_zero;
    plan.scratch_bytes = off;
    return true;
}

void fill_scan_tables(const ScanPlan& plan, char* scratch, char* table)
{
    (void)scratch;
    trtc::SegScanDesc* D = (trtc::SegScanDesc*)(table + plan.off_descs);
    unsigned* TS = (unsigned*)(table + plan.off_tile_seg);
    unsigned t = 0;
    for (size_t i = 0; i < plan.segs.size(); i++) {
        const ScanSeg& g = plan.segs[i];
        D[i].in = (unsigned long long)g.in;
        D[i].out = (unsigned long long)g.out;
        D[i].n = g.n;
        D[i].first_tile = t;
        D[i].lead = scan_seg_lead(g, plan.tile, plan.tsz);
        size_t nt = scan_seg_tiles(g, plan.tile, plan.tsz);
        for (size_t k = 0; k < nt; k++) TS[t++] = (unsigned)i;
    }
}

bool enqueue_scan(const ScanPlan& plan, char* scratch, hipStream_t s)
{
    if (!plan.num_tiles) return true;
    if (!fill_u32(scratch + plan.off_zero, plan.zero_bytes / 4, 0, s)) return false;
    std::string name = seg_scan_name(plan);
    KernelRef k;
    if (!get_template_kernel(name, name, k)) return false;
    trtc::SegScanArgs a;
    a.descs = (const trtc::SegScanDesc*)(scratch + plan.off_descs);
    a.tile_seg = (const unsigned*)(scratch + plan.off_tile_seg);
    a.num_tiles = plan.num_tiles;
    a.lb.tile_counter = (unsigned*)(scratch + plan.off_zero);
    a.lb.agg = (unsigned long long*)(scratch + plan.off_zero + 256);
    a.lb.incl = a.lb.agg + (size_t)plan.num_tiles * plan.words;
    a.exclusive = plan.exclusive ? 1 : 0;
    ViewBuf opv = plan.op_view, initv = plan.init;
    void* args[] = {opv.data(), initv.data(), &a};
    return launch_kernel(k, dim3(plan.num_tiles), dim3(seg_scan_block(plan.tsz)), args, 0, s);
}

bool run_scan_batch(const std::vector<ScanSeg>& segs, const std::string& T, const Functor& op, const ViewBuf& init,
                    bool exclusive, int items, hipStream_t s)
{
    ScanPlan plan;
    if (!plan_scan(segs, T, op, init, exclusive, items, plan)) return false;
    if (!plan.num_tiles) return true;
    PoolBuf scratch(plan.scratch_bytes, s);
    if (!scratch.p) return false;
    return launch_scan_plan(plan, (char*)scratch.p, s);
}

bool launch_scan_plan(const ScanPlan& plan, char* scratch, hipStream_t s)
{
    if (!plan.num_tiles) return true;
    std::vector<char> table(plan.table_bytes);
    fill_scan_tables(plan, scratch, table.data());
    if (!upload_async(scratch, table.data(), table.size(), s)) return false;
    return enqueue_scan(plan, scratch, s);
}

// ================================================================ reduce
namespace {
const size_t kReduceTile = 16384;  // TRTC_SEG_REDUCE_TILE in the device header
}

bool plan_reduce(const std::vector<ReduceSeg>& segs, const std::string& T, const Functor& op, ReducePlan& plan)
{
    plan = ReducePlan();
    const ScalarInfo* si = scalar_info(T);
    if (!si) {
        TRTC_ERR("batched reduce: element type %s is not a scalar", T.c_str());
        return false;
    }
    plan.T = si->canon;
    plan.tsz = si->size;
    plan.op_cls = canon_type(op.name_view_cls());
    plan.op_view = op.view();
    // partials are Maybe<T> = {T v; unsigned has}
    size_t al = std::min<size_t>(std::max<size_t>(plan.tsz, 4), 8);
    size_t maybe = (plan.tsz + 4 + al - 1) / al * al;
    plan.words = (maybe + 3) / 4;
    for (auto& g : segs) {
        if (!g.n) {
            TRTC_ERR("batched reduce: empty segment");
            return false;
        }
        plan.segs.push_back(g);
        plan.num_tiles += (unsigned)ceil_div(g.n, kReduceTile);
    }
    size_t off = 0;
    plan.off_descs = off;
    off += align256(plan.segs.size() * sizeof(trtc::SegReduceDesc));
    plan.off_tile_seg = off;
    off += align256((size_t)plan.num_tiles * 4);
    plan.table_bytes = off;
    plan.off_zero = off;
    off += align256(plan.segs.size() * 4) + align256((size_t)plan.num_tiles * plan.words * 8);
    plan.zero_bytes = off - plan.off_zero;
    plan.scratch_bytes = off;
    return true;
}

void fill_reduce_tables(const ReducePlan& plan, char* scratch, char* table)
{
    (void)scratch;
    trtc::SegReduceDesc* D = (trtc::SegReduceDesc*)(table + plan.off_descs);
    unsigned* TS = (unsigned*)(table + plan.off_tile_seg);
    unsigned t = 0;
    for (size_t i = 0; i < plan.segs.size(); i++) {
        const ReduceSeg& g = plan.segs[i];
        size_t nt = ceil_div(g.n, kReduceTile);
        D[i].in = (unsigned long long)g.in;
        D[i].n = g.n;
        D[i].first_tile = t;
        D[i].num_tiles = (unsigned)nt;
        D[i].result = (unsigned long long)g.result;
        for (size_t k = 0; k < nt; k++) TS[t++] = (unsigned)i;
    }
}

bool enqueue_reduce(const ReducePlan& plan, char* scratch, hipStream_t s)
{
    if (!plan.num_tiles) return true;
    if (!fill_u32(scratch + plan.off_zero, plan.zero_bytes / 4, 0, s)) return false;
    std::string name = seg_reduce_name(plan);
    KernelRef k;
    if (!get_template_kernel(name, name, k)) return false;
    trtc::SegReduceArgs a;
    a.descs = (const trtc::SegReduceDesc*)(scratch + plan.off_descs);
    a.tile_seg = (const unsigned*)(scratch + plan.off_tile_seg);
    a.num_tiles = plan.num_tiles;
    a.seg_done = (unsigned*)(scratch + plan.off_zero);
    a.partials = (unsigned long long*)(scratch + plan.off_zero + align256(plan.segs.size() * 4));
    ViewBuf opv = plan.op_view;
    void* args[] = {opv.data(), &a};
    return launch_kernel(k, dim3(plan.num_tiles), dim3(256), args, 0, s);
}

bool run_reduce_batch(const std::vector<ReduceSeg>& segs, const std::string& T, const Functor& op, hipStream_t s)
{
    ReducePlan plan;
    if (!plan_reduce(segs, T, op, plan)) return false;
    if (!plan.num_tiles) return true;
    PoolBuf scratch(plan.scratch_bytes, s);
    if (!scratch.p) return false;
    return launch_reduce_plan(plan, (char*)scratch.p, s);
}

bool launch_reduce_plan(const ReducePlan& plan, char* scratch, hipStream_t s)
{
    if (!plan.num_tiles) return true;
    std::vector<char> table(plan.table_bytes);
    fill_reduce_tables(plan, scratch, table.data());
    if (!upload_async(scratch, table.data(), table.size(), s)) return false;
    return enqueue_reduce(plan, scratch, s);
}

}  // namespace trtc_host
