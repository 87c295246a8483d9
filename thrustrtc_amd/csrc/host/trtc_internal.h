// Internal interfaces of the host runtime (not installed).
#pragma once
#include <hip/hip_runtime_api.h>

#include <atomic>
#include <cstdint>
#include <functional>
#include <cstdio>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "DVVector.h"
#include "TRTCContext.h"
#include "functor.h"

namespace trtc_host {

// ---------------------------------------------------------------- types
enum TypeCode {
    TC_I8 = 0, TC_U8, TC_I16, TC_U16, TC_I32, TC_U32, TC_I64, TC_U64, TC_F32, TC_F64, TC_BOOL, TC_OTHER
};

struct ScalarInfo
{
    const char* canon;  // canonical device type name (e.g. "int32_t")
    TypeCode code;
    size_t size;
};

// Scalar type info for any spelling the reference accepts ("int", "int32_t",
// "unsigned long long", "size_t", ...); nullptr for non-scalars.
const ScalarInfo* scalar_info(const std::string& name);
// Canonical spelling of a type name: scalars are mapped to their fixed-width
// name, VectorView<...> / RangeView<...> arguments are canonicalised too.
std::string canon_type(const std::string& name);
// Numeric conversion between scalar types with C++ cast semantics.
bool convert_scalar(TypeCode from, const void* src, TypeCode to, void* dst);
const char* type_name(TypeCode tc);
size_t type_size(TypeCode tc);

// ---------------------------------------------------------------- layout
// Alignment of the device type named view_cls whose view has view_size bytes.
size_t view_align(const std::string& view_cls, size_t view_size);
// Size/alignment of registered structs (functors).
void register_struct_layout(const std::string& name, size_t size, size_t align);

// ---------------------------------------------------------------- context
struct Context
{
    bool ok = false;
    int device = 0;
    std::string arch;      // full target id, e.g. gfx950:sramecc+:xnack-
    std::string arch_base; // gfx950
    int arch_num = 0;      // 950
    int num_cus = 0;
    size_t lds_per_block = 0;
    bool verbose = false;
    bool kernel_debug = false;
    // bumped whenever code blocks, headers, include directories or constants
    // change (they are part of every generated kernel's source)
    std::atomic<unsigned long long> generation{0};

    std::mutex mtx;
    std::vector<std::string> include_dirs;
    std::vector<std::string> header_names;
    std::vector<std::string> header_contents;
    std::vector<std::string> code_blocks;
    std::vector<std::pair<std::string, ViewBuf>> constants;
};

Context& context();
// Initialise the process-wide context (once) and make its device current on
// this thread.  False if no usable GPU.
bool ensure_init();
hipStream_t cur_stream();
void set_stream(hipStream_t s);

bool check_hip(hipError_t e, const char* what);

// ---------------------------------------------------------------- structs
// Register a struct body; returns its name _S_<hash>.  Thread-safe.
std::string add_struct(const std::string& body);
// Definitions (in registration order) of every registered struct named in
// `text`, including the structs those definitions name in turn.
std::string struct_defs_for(const std::string& text);

// ---------------------------------------------------------------- memory
// Stream-ordered caching pool for temporaries: blocks are reused only on the
// stream that freed them.
void* pool_alloc(size_t bytes, hipStream_t s);
void pool_free(void* p, hipStream_t s);

struct PoolBuf
{
    void* p = nullptr;
    hipStream_t s = nullptr;
    PoolBuf() {}
    PoolBuf(size_t bytes, hipStream_t st) : s(st) { p = bytes ? pool_alloc(bytes, st) : nullptr; }
    ~PoolBuf() { if (p) pool_free(p, s); }
    PoolBuf(const PoolBuf&) = delete;
    PoolBuf& operator=(const PoolBuf&) = delete;
    template <class T> T* as() const { return (T*)p; }
};

// ---------------------------------------------------------------- kernels
struct KernelRef
{
    const void* aot = nullptr;       // host stub of an ahead-of-time kernel
    hipFunction_t jit = nullptr;     // function of a run-time compiled module
    const char* name = "";           // name expression (stable storage)
    bool valid() const { return aot || jit; }
};

// Ahead-of-time kernels register themselves under their name expression.
void register_aot(const char* name_expr, const void* fn);
const void* find_aot(const std::string& name_expr);

// A kernel template instantiation by name expression: the ahead-of-time one if
// it exists, otherwise compiled by hipRTC (cached in memory and on disk).
// `type_text` names every type involved (used to pull in struct definitions).
bool get_template_kernel(const std::string& name_expr, const std::string& type_text, KernelRef& out);
// The same lookup without compiling or loading anything: an ahead-of-time
// kernel, or one an earlier call already compiled and loaded in this process.
bool find_template_kernel(const std::string& name_expr, const std::string& type_text, KernelRef& out);
// An extern "C" kernel named `fn_name` defined in `kernel_src` (user kernels).
bool get_source_kernel(const std::string& kernel_src, const std::string& fn_name, KernelRef& out);
// A kernel resolved once per `key` (plus the context generation) by `build`;
// later calls with the same key return it without running `build`.
bool memo_kernel(const std::string& key, const std::function<bool(KernelRef&)>& build, KernelRef& out);
// Grid of a grid-stride launch over `items` units of `per_block` each, at most `cap` blocks.
unsigned capped_grid(size_t items, size_t per_block, size_t cap);
// Compile only (no device needed): for offline cache warm-up and checks.
bool compile_only(const std::string& arch, const std::string& name_expr, const std::string& type_text,
                  std::string* log);

// Disk cache entries of run-time compiled code objects: the header is checked
// (magic, code length, FNV-1a checksum) before anything is handed to the
// loader.  True and the parts for a well-formed entry.
bool parse_cache_entry(const std::string& blob, std::string& lowered, std::string& code);
// JIT lookups so far: in-memory hits, disk-cache hits, hipRTC compilations,
// disk entries rejected by parse_cache_entry.
void jit_cache_stats(unsigned long long out[4]);

bool launch_kernel(const KernelRef& k, dim3 grid, dim3 block, void** args, size_t shmem, hipStream_t s);
bool sync_if_debug(hipStream_t s, const char* what);
// Kernel timing: while enabled, every launch is bracketed by HIP events on its
// stream; profile_end() synchronises and returns per-kernel counts and times.
void profile_begin();
std::string profile_end();

// While a thread captures a HIP graph, launches skip the timing events and
// the debug synchronisation (neither may be recorded into a graph).
void set_capturing(bool on);
bool capturing();
// hipGraphLaunch with the same timing hook as kernel launches.
bool launch_graph(hipGraphExec_t g, const char* name, hipStream_t s);

// Pinned host staging blocks for asynchronous uploads: a released block is
// reused only after the stream has passed the point of release.
void* pinned_alloc(size_t bytes);
void pinned_release(void* p, hipStream_t s);

// Built-in utility kernels.
bool fill_u32(void* p, size_t words, unsigned v, hipStream_t s);
// The stream's radix histogram block (8 * 256 64-bit counters, then the
// completion counter; zero between sorts), or null if it cannot be allocated.
constexpr size_t kRadixHistBlockBytes = 8 * 256 * 8 + 256;
void* radix_hist_block(hipStream_t s);
// Zeroes words with agent-scope atomic stores (lookback state in static
// device memory, read by the next kernel's agent-scope atomics).
bool clear_u32_agent(void* p, size_t words, hipStream_t s);
// Device-to-device copy of nbytes: a 16-byte-per-thread kernel when both
// pointers are 16-byte aligned, hipMemcpyAsync otherwise.
bool copy_bytes(const void* src, void* dst, size_t nbytes, hipStream_t s);
bool iota_u32(unsigned* p, size_t n, hipStream_t s);
bool gather_bytes(const void* src, void* dst, const unsigned* idx, size_t n, size_t elem, hipStream_t s);

// Radix key class of an element type: width in bytes and kind (0 unsigned /
// bool, 1 signed, 2 floating point); false for other types.
bool radix_key_info(const std::string& elem, int& key_bytes, int& kind);

// Inclusive_Scan with its lookback state in caller memory of scratch_bytes
// (allocates and compiles nothing; fails if the kernel does not exist yet or
// the input needs more scratch).
bool inclusive_scan_fixed_scratch(const DVVectorLike& vin, DVVectorLike& vout, const Functor& op, void* scratch,
                                  size_t scratch_bytes);

// Single-call onesweep radix sort (the TRTC_Sort_By_Key path) reading its
// input from keys_src/vals_src and leaving the result in keys/vals.
// Grid of the long-tile radix passes for `tiles` tiles (one workgroup per CU,
// a multiple of 8).
size_t long_pass_grid(size_t tiles);
bool radix_sort_from(const void* keys_src, const void* vals_src, void* keys, void* vals, size_t n, int key_bytes,
                     int kind, bool descending, int val_bytes, hipStream_t s);

// ---------------------------------------------------------------- vectors
// Raw device pointer of a contiguous vector-like (DVVector, DVVectorAdaptor, a DVRange of one),
// nullptr otherwise.
void* contiguous_data(const DVVectorLike& v);

// The single-launch reduction of the primitive core over any view (init
// first when given; an empty vector without init is an error).
bool reduce_view(const DVVectorLike& vec, const DeviceViewable* init, const Functor& op, ViewBuf& ret);
// Min/Max_Element of a plain arithmetic vector in one streaming launch;
// `handled` is false (and nothing runs) for any other vector.
bool extreme_view(const DVVectorLike& vec, const Functor& comp, bool max, size_t& id, bool& handled);

uint64_t fnv1a(const void* data, size_t n, uint64_t h = 1469598103934665603ull);

// Element-wise map over [0, n) for bodies that read at index idx, then write
// at idx: `value` is an expression in `idx` and the named views (args, same
// order), `store` a statement that writes `_x` at idx.  8 indices per thread,
// all read before any is written (in-place calls are safe).
bool run_map(const std::vector<const char*>& names, const char* value, const char* store, size_t n,
             std::initializer_list<const DeviceViewable*> args);

}  // namespace trtc_host

#define TRTC_ERR(...) (fprintf(stderr, "[ThrustRTC] " __VA_ARGS__), fprintf(stderr, "\n"))
