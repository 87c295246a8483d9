// Batched-segment engine (internal): many independent sorts, scans and
// reductions in one launch per kernel stage.  Used by the segmented API, the
// graph-captured Batch object and the request server.
#pragma once
#include <string>
#include <vector>

#include "../device/trtc_args.h"
#include "trtc_internal.h"

namespace trtc_host {

// ---------------------------------------------------------------- sort
struct SortSeg
{
    char* keys;
    char* vals;  // null for keys-only sorts
    size_t n;
    const char* src_keys = nullptr;  // read by the first pass instead of keys (out-of-place sort)
    const char* src_vals = nullptr;
};

// Segments of up to max_n keys run in a launch with tiles of block*items keys.
struct SortClass
{
    size_t max_n;
    int block;
    int items;
};
typedef std::vector<SortClass> SortSchedule;
std::string describe(const SortSchedule& s);
SortSchedule default_sort_schedule(int key_bytes, int val_bytes);
SortSchedule served_sort_schedule(int key_bytes, int val_bytes);

struct SortPlan
{
    int KB = 4, VB = 0, kind = 0;
    bool desc = false;
    struct Cls
    {
        int block = 512, items = 16;
        size_t tile = 8192;
        std::vector<SortSeg> segs;  // multi-tile segments first (largest first), then one-tile ones
        std::vector<size_t> tmp_off;  // element offset into the ping-pong buffers (multi-tile only)
        unsigned num_tiles = 0, multi_tiles = 0, multi_segs = 0, hist_chunks = 0;
        size_t off_descs = 0, off_tile_seg = 0, off_hist_chunk = 0;  // in the uploaded table block
        size_t off_zero = 0, zero_bytes = 0;         // in scratch: cleared before the launches
        size_t off_cnt = 0, off_done = 0, off_hist = 0, off_status = 0, status_pass_words = 0;
        size_t off_dstart = 0, off_rank = 0;
    };
    std::vector<Cls> cls;
    size_t hist_chunk = 131072;  // keys per histogram workgroup
    size_t tmp_elems = 0;     // keys in multi-tile segments (ping-pong buffer length)
    size_t table_bytes = 0;   // descriptor tables, uploaded at the start of scratch
    size_t scratch_bytes = 0; // tables + lookback state + histograms
    size_t keys_total = 0;
};

// Sizes and layouts only (no device work).  Segments with nothing to move
// (empty, or one key sorted in place) are dropped.
// Every segment must hold fewer than 2^30 keys.
bool plan_sort(const std::vector<SortSeg>& segs, int key_bytes, int val_bytes, int kind, bool descending,
               const SortSchedule& sched, SortPlan& plan);
// Writes the descriptor tables (device addresses inside scratch / tmp) into
// host memory `table` (plan.table_bytes).
void fill_sort_tables(const SortPlan& plan, char* scratch, char* tmp_keys, char* tmp_vals, char* table);
// Issues the clears, histogram pre-passes and digit passes on stream s; the
// tables must already be in scratch.
bool enqueue_sort(const SortPlan& plan, char* scratch, hipStream_t s);
// Resolves (compiling if needed) every kernel the plan launches, so that a
// later enqueue under graph capture does no compilation or module loading.
bool prepare_sort(const SortPlan& plan);
// Tables + upload + enqueue (+ copy-back for odd digit counts) with caller
// memory: scratch of plan.scratch_bytes, ping-pong buffers of
// plan.tmp_elems keys / values.
bool launch_sort_plan(const SortPlan& plan, char* scratch, char* tmp_keys, char* tmp_vals, hipStream_t s);
// One-shot: plan, allocate from the pool, upload, enqueue.
bool run_sort_batch(const std::vector<SortSeg>& segs, int key_bytes, int val_bytes, int kind, bool descending,
                    const SortSchedule& sched, hipStream_t s);

// Schedule for this batch shape: tuned (timed on a copy of the data) when
// autotuning is on and the shape's signature has not been tuned yet,
// otherwise the cached or default one.
SortSchedule schedule_for(const std::vector<SortSeg>& segs, int key_bytes, int val_bytes, int kind, bool descending,
                          hipStream_t s);
void set_autotune(bool on);
// Parses describe()'s format ("<=2048:256x8 rest:512x16").
bool parse_schedule(const std::string& spec, SortSchedule& out);
// A schedule every batched sort uses instead of the tuned one (NULL or "":
// back to automatic).
bool set_forced_schedule(const char* spec);
bool autotune_on();
std::string tuning_report();

// ---------------------------------------------------------------- scan
struct ScanSeg
{
    const char* in;
    char* out;
    size_t n;
};

struct ScanPlan
{
    std::string T;        // element type (input and output)
    std::string op_cls;   // functor type name
    ViewBuf op_view, init;
    size_t tsz = 4;
    int items = 16;
    size_t tile = 4096;
    bool exclusive = false;
    std::vector<ScanSeg> segs;
    unsigned num_tiles = 0;
    size_t words = 1;
    size_t off_descs = 0, off_tile_seg = 0, table_bytes = 0;
    size_t off_zero = 0, zero_bytes = 0, scratch_bytes = 0;
};
bool plan_scan(const std::vector<ScanSeg>& segs, const std::string& T, const Functor& op, const ViewBuf& init,
               bool exclusive, int items, ScanPlan& plan);
void fill_scan_tables(const ScanPlan& plan, char* scratch, char* table);
bool enqueue_scan(const ScanPlan& plan, char* scratch, hipStream_t s);
bool prepare_scan(const ScanPlan& plan);
bool launch_scan_plan(const ScanPlan& plan, char* scratch, hipStream_t s);
bool run_scan_batch(const std::vector<ScanSeg>& segs, const std::string& T, const Functor& op, const ViewBuf& init,
                    bool exclusive, int items, hipStream_t s);
int default_scan_items(size_t tsz);

// ---------------------------------------------------------------- reduce
struct ReduceSeg
{
    const char* in;
    size_t n;
    char* result;  // device address of one element
};

struct ReducePlan
{
    std::string T, op_cls;
    ViewBuf op_view;
    size_t tsz = 4;
    std::vector<ReduceSeg> segs;
    unsigned num_tiles = 0;
    size_t words = 2;
    size_t off_descs = 0, off_tile_seg = 0, table_bytes = 0;
    size_t off_zero = 0, zero_bytes = 0, scratch_bytes = 0;
};
bool plan_reduce(const std::vector<ReduceSeg>& segs, const std::string& T, const Functor& op, ReducePlan& plan);
void fill_reduce_tables(const ReducePlan& plan, char* scratch, char* table);
bool enqueue_reduce(const ReducePlan& plan, char* scratch, hipStream_t s);
bool prepare_reduce(const ReducePlan& plan);
bool launch_reduce_plan(const ReducePlan& plan, char* scratch, hipStream_t s);
bool run_reduce_batch(const std::vector<ReduceSeg>& segs, const std::string& T, const Functor& op, hipStream_t s);

// Upload `bytes` of host data to device `dst` on stream s through a pinned
// staging block (asynchronous; the host data may be freed on return).
bool upload_async(void* dst, const void* src, size_t bytes, hipStream_t s);

}  // namespace trtc_host
