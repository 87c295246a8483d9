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
    int has_stencil;            // predicate reads the stencil, not the input
};

struct ScanByKeyArgs
{
    size_t n;
    LookbackArgs lb;
    int exclusive;
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
};

}  // namespace trtc
