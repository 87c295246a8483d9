// Kernel argument blocks shared verbatim by the host drivers and the device
// kernels (plain structs, identical layout on both sides).
#pragma once
#if !defined(__HIPCC_RTC__)
#include <stddef.h>
#endif

namespace trtc {

struct LookbackArgs
{
    unsigned long long* agg;   // [tiles * words] status words, zero on entry
    unsigned long long* incl;  // [tiles * words] status words, zero on entry
    unsigned* tile_counter;    // dynamic tile ids, zero on entry
};

struct ScanArgs
{
    size_t n;
    LookbackArgs lb;
    int exclusive;
};

struct ReduceArgs
{
    size_t n;
    size_t per_wave;               // elements per wave range (multiple of the chunk)
    unsigned long long* partials;  // [gridDim.x * words] status words, zero on entry
    unsigned* done;                // zero on entry
    void* result;                  // one element of T
    int has_init;
};

struct CopyIfArgs
{
    size_t n;
    LookbackArgs lb;
    unsigned long long* count;  // out: total kept (written by the last tile)
    unsigned num_tiles;
    int has_stencil;            // informational: kernels take the stencil type (_NoStencil: none)
};

struct ScanByKeyArgs
{
    size_t n;
    LookbackArgs lb;
    int exclusive;
};

// Device-side plan of one single-sequence radix sort, written by the
// histogram pre-pass once it knows which digits are constant (one bin holds
// every key): per pass the key / value buffers to read and write (kin 0: the
// pass is skipped), and the copy that brings the result into the caller's
// buffers when an odd number of passes ran in place (ksrc 0: none).  The
// passes and the copy read it, so skipping needs no host synchronisation.
// rank_lds: bit p set when pass p's digits are spread thinly enough over the
// 256 bins for the long-tile pass to rank them with LDS count|mask words
// rather than ballots (few lanes of a slot share a digit).
struct RadixPlan
{
    unsigned long long kin[8], kout[8], vin[8], vout[8];
    unsigned long long copy_ksrc, copy_kdst, copy_vsrc, copy_vdst;
    unsigned rank_lds;
    unsigned copy_mask;  // bit p: pass p's launch copies kin -> kout (its digit is constant)
};

struct RadixHistArgs
{
    const void* keys;
    size_t n;
    int kind;
    int descending;
    int passes;
    unsigned long long* hist;         // [KB][256] global counts, zero on entry
    unsigned long long* digit_start;  // [KB][256] out: exclusive digit offsets
    unsigned* done;                   // workgroup completion counter, zero on entry
    unsigned* zero_ptr;               // words to clear for the passes (status + tile counters)
    size_t zero_words;
    // buffers of the sort for the plan (plan null: no plan is written)
    RadixPlan* plan;
    const void* vals_src;             // first pass input (keys: `keys`)
    void* keys_dst;                   // caller's buffers (result)
    void* vals_dst;
    void* keys_tmp;                   // ping-pong buffers
    void* vals_tmp;
    int val_bytes;
    int self_clear;                   // the last workgroup zeroes hist and done again after use
};

struct RadixPassArgs
{
    const void* keys_in;
    void* keys_out;
    const void* vals_in;
    void* vals_out;
    size_t n;
    const unsigned long long* digit_start;  // [256] for this pass
    void* status;                           // [tiles][256] status words, zero on entry
    unsigned* tile_counter;                 // zero on entry
    int shift;
    int kind;
    int descending;
    const RadixPlan* plan;                  // non-null: buffers from plan (pass `pass`), skip when kin is 0
    int pass;
};

struct RadixCopyArgs
{
    const RadixPlan* plan;
    size_t key_bytes;                       // n * key width
    size_t val_bytes;                       // n * value width (0: keys only)
};

struct MergeArgs
{
    size_t na;  // length of the first sorted input
    size_t nb;  // length of the second
    const unsigned long long* splits;  // [tiles + 1] merge-path splits at the tile starts, or null
};

struct MergeSortArgs
{
    size_t n;
    size_t width;                // run length of the input of a merge pass
    int with_idx;                // carry the input index of every key
    unsigned long long* splits;  // per merge tile: its first diagonal's split (null: search in the merge)
};

// ---- batched segments: many independent sequences in one launch --------
// A launch covers the tiles of a list of segments; tile ids 0..num_tiles-1
// are handed out by an atomic counter in order, and tile_seg maps each one to
// its segment.  Segments that span several tiles come first (their tiles use
// lookback state); single-tile segments follow.

struct SegSortDesc
{
    unsigned long long src_keys, src_vals;  // input of the first pass (== keys/vals for in-place sorts)
    unsigned long long keys, vals;          // device addresses of the segment (the sorted result)
    unsigned long long tmp_keys, tmp_vals;  // ping-pong buffers (== keys/vals for single-tile segments)
    unsigned long long n;                   // keys in the segment (< 2^30)
    unsigned first_tile;                    // launch tile id of the segment's tile 0
    unsigned multi;                         // slot among multi-tile segments, 0xFFFFFFFF if single-tile
    unsigned long long status_row;          // multi-tile: lookback row of tile 0 (rows of 256 words)
    unsigned hist_chunks;                   // multi-tile: histogram workgroups of the segment
};

struct SegRadixArgs
{
    const SegSortDesc* descs;
    const unsigned* tile_seg;  // [num_tiles] segment index of each launch tile
    unsigned num_tiles;
    unsigned num_multi_tiles;  // tiles 0..num_multi_tiles-1 belong to multi-tile segments
    unsigned* tile_counter;    // zero on entry (one per pass; 8 per pass for the long-tile pass)
    unsigned* status;          // [num_multi_tiles][256] lookback words of this pass, zero on entry
    unsigned* seg_hist;        // [multi segments][KB][256] digit counts, zero on entry
    unsigned* seg_dstart;      // [multi segments][KB][256] exclusive digit offsets
    unsigned* seg_done;        // [multi segments] arrival counters, zero on entry
    unsigned* seg_rank_lds;    // [multi segments] bit p: the long-tile pass ranks digit p with LDS words
    const unsigned* hist_chunk;  // [num_hist_chunks][2]: segment index, chunk index inside it
    unsigned num_hist_chunks;
    unsigned hist_chunk_keys;    // keys per histogram chunk
    int pass;                  // digit index (reads keys when even, tmp when odd)
    int shift;
    int kind;
    int descending;
};

struct SegScanDesc
{
    unsigned long long in, out;  // device addresses
    unsigned long long n;
    unsigned first_tile;
    unsigned lead;  // multi-tile: (in % 128) / element size; tile 0 holds tile - lead elements
                    // so that later tiles start on 128-byte lines
};

struct SegScanArgs
{
    const SegScanDesc* descs;
    const unsigned* tile_seg;
    unsigned num_tiles;
    LookbackArgs lb;
    int exclusive;
};

struct SegReduceDesc
{
    unsigned long long in;  // device address
    unsigned long long n;
    unsigned first_tile;
    unsigned num_tiles;
    unsigned long long result;  // device address of one element of T
};

struct SegReduceArgs
{
    const SegReduceDesc* descs;
    const unsigned* tile_seg;
    unsigned num_tiles;
    unsigned long long* partials;  // [num_tiles * words] status words, zero on entry
    unsigned* seg_done;            // [segments] arrival counters, zero on entry
};

}  // namespace trtc
