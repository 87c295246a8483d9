// Process-wide context: device selection, type tables, struct registry, the
// stream-ordered memory pool, and the kernel registry (ahead-of-time kernels
// plus hipRTC instantiation with an in-memory and on-disk code-object cache).
//
// Replaces the reference's TRTCContext (TRTCContext.cpp, internal/
// impl_context.inl:47-953), its dlopen'd CUDA/NVRTC shims (internal/
// cuda_wrapper.*, nvtrc_wrapper.*) and its unqlite PTX cache.  Differences
// by design: kernels are looked up by a short name expression instead of
// re-assembling and hashing the whole source on every launch, the disk cache
// is one file per code object keyed by (source hash, target id, hipRTC
// version), and temporaries come from a pool instead of cuMemAlloc+memset.
#include <hip/hiprtc.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <atomic>
#include <functional>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <tuple>
#include <memory>
#include <sstream>

#include "trtc_internal.h"

#include "embedded_headers.inc"  // k_embedded_count, k_embedded_names[], k_embedded_sources[]

namespace trtc_host {

uint64_t fnv1a(const void* data, size_t n, uint64_t h)
{
    const unsigned char* p = (const unsigned char*)data;
    for (size_t i = 0; i < n; i++) {
        h ^= p[i];
        h *= 1099511628211ull;
    }
    return h;
}

bool check_hip(hipError_t e, const char* what)
{
    if (e == hipSuccess) return true;
    TRTC_ERR("%s failed: %s (%d)", what, hipGetErrorString(e), (int)e);
    (void)hipGetLastError();  // reported here; later checks must not see it again
    return false;
}

// ================================================================ types
static const ScalarInfo k_scalars[] = {
    {"int8_t", TC_I8, 1},   {"uint8_t", TC_U8, 1},   {"int16_t", TC_I16, 2}, {"uint16_t", TC_U16, 2},
    {"int32_t", TC_I32, 4}, {"uint32_t", TC_U32, 4}, {"int64_t", TC_I64, 8}, {"uint64_t", TC_U64, 8},
    {"float", TC_F32, 4},   {"double", TC_F64, 8},   {"bool", TC_BOOL, 1},
};

const ScalarInfo* scalar_info(const std::string& name)
{
    static const std::unordered_map<std::string, int> alias = {
        {"int8_t", 0}, {"char", 0}, {"signed char", 0},
        {"uint8_t", 1}, {"unsigned char", 1},
        {"int16_t", 2}, {"short", 2}, {"signed short", 2},
        {"uint16_t", 3}, {"unsigned short", 3},
        {"int32_t", 4}, {"int", 4}, {"signed int", 4},
        {"uint32_t", 5}, {"unsigned int", 5}, {"unsigned", 5},
        {"int64_t", 6}, {"long", 6}, {"long long", 6}, {"signed long long", 6}, {"ptrdiff_t", 6},
        {"uint64_t", 7}, {"unsigned long", 7}, {"unsigned long long", 7}, {"size_t", 7},
        {"float", 8}, {"double", 9}, {"bool", 10},
    };
    auto it = alias.find(name);
    if (it == alias.end()) return nullptr;
    return &k_scalars[it->second];
}

const char* type_name(TypeCode tc) { return tc < TC_OTHER ? k_scalars[tc].canon : "?"; }
size_t type_size(TypeCode tc) { return tc < TC_OTHER ? k_scalars[tc].size : 0; }

std::string canon_type(const std::string& name)
{
    const ScalarInfo* si = scalar_info(name);
    if (si) return si->canon;
    for (const char* tmpl : {"VectorView<", "RangeView<"}) {
        size_t L = strlen(tmpl);
        if (name.size() > L + 1 && name.compare(0, L, tmpl) == 0 && name.back() == '>')
            return std::string(tmpl) + canon_type(name.substr(L, name.size() - L - 1)) + ">";
    }
    return name;
}

template <class T> static bool load_as(TypeCode tc, const void* p, T& v)
{
    switch (tc) {
    case TC_I8: v = (T) * (const int8_t*)p; return true;
    case TC_U8: v = (T) * (const uint8_t*)p; return true;
    case TC_I16: v = (T) * (const int16_t*)p; return true;
    case TC_U16: v = (T) * (const uint16_t*)p; return true;
    case TC_I32: v = (T) * (const int32_t*)p; return true;
    case TC_U32: v = (T) * (const uint32_t*)p; return true;
    case TC_I64: v = (T) * (const int64_t*)p; return true;
    case TC_U64: v = (T) * (const uint64_t*)p; return true;
    case TC_F32: v = (T) * (const float*)p; return true;
    case TC_F64: v = (T) * (const double*)p; return true;
    case TC_BOOL: v = (T) * (const bool*)p; return true;
    default: return false;
    }
}

bool convert_scalar(TypeCode from, const void* src, TypeCode to, void* dst)
{
#define CV(T) { T v; if (!load_as<T>(from, src, v)) return false; memcpy(dst, &v, sizeof(T)); return true; }
    switch (to) {
    case TC_I8: CV(int8_t)
    case TC_U8: CV(uint8_t)
    case TC_I16: CV(int16_t)
    case TC_U16: CV(uint16_t)
    case TC_I32: CV(int32_t)
    case TC_U32: CV(uint32_t)
    case TC_I64: CV(int64_t)
    case TC_U64: CV(uint64_t)
    case TC_F32: CV(float)
    case TC_F64: CV(double)
    case TC_BOOL: CV(bool)
    default: return false;
    }
#undef CV
}

// ================================================================ layout
static std::mutex g_layout_mtx;
static std::unordered_map<std::string, std::pair<size_t, size_t>> g_struct_layout;  // size, align

void register_struct_layout(const std::string& name, size_t size, size_t align)
{
    std::lock_guard<std::mutex> lk(g_layout_mtx);
    g_struct_layout[name] = {size, align};
}

size_t view_align(const std::string& view_cls, size_t view_size)
{
    const ScalarInfo* si = scalar_info(view_cls);
    if (si) return si->size;
    if (view_cls.compare(0, 11, "VectorView<") == 0 || view_cls.compare(0, 10, "RangeView<") == 0) return 8;
    if (view_cls == "cuFloatComplex") return 8;
    if (view_cls == "cuDoubleComplex") return 16;
    {
        std::lock_guard<std::mutex> lk(g_layout_mtx);
        auto it = g_struct_layout.find(view_cls);
        if (it != g_struct_layout.end()) return it->second.second;
    }
    static const char* builtin[] = {"Identity", "Maximum", "Minimum", "EqualTo", "NotEqualTo", "Greater", "Less",
                                    "GreaterEqual", "LessEqual", "Plus", "Minus", "Multiplies", "Divides",
                                    "Modulus", "Negate"};
    for (const char* b : builtin)
        if (view_cls == b) return 1;
    size_t a = 8;
    while (a > 1 && (view_size % a) != 0) a >>= 1;
    return a;
}

// ================================================================ context
static Context* g_ctx = nullptr;
static std::once_flag g_ctx_once;
static thread_local bool t_device_set = false;
static thread_local hipStream_t t_stream = nullptr;

Context& context()
{
    static Context c;
    return c;
}

static void init_context()
{
    Context& c = context();
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count < 1) return;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) dev = 0;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev) != hipSuccess) return;
    c.device = dev;
    c.arch = prop.gcnArchName;
    c.arch_base = c.arch.substr(0, c.arch.find(':'));
    c.arch_num = 0;
    if (c.arch_base.compare(0, 3, "gfx") == 0) c.arch_num = atoi(c.arch_base.c_str() + 3);
    c.num_cus = prop.multiProcessorCount;
    c.lds_per_block = prop.sharedMemPerBlock;
    c.ok = true;
    g_ctx = &c;
}

bool ensure_init()
{
    std::call_once(g_ctx_once, init_context);
    Context& c = context();
    if (!c.ok) return false;
    if (!t_device_set) {
        if (hipSetDevice(c.device) != hipSuccess) return false;
        t_device_set = true;
    }
    return true;
}

hipStream_t cur_stream() { return t_stream; }
void set_stream(hipStream_t s) { t_stream = s; }

// ================================================================ structs
static std::mutex g_struct_mtx;
static std::vector<std::pair<std::string, std::string>> g_structs;  // name, full definition
static std::unordered_map<std::string, size_t> g_struct_index;

std::string add_struct(const std::string& body)
{
    uint64_t h = fnv1a(body.data(), body.size());
    char name[32];
    snprintf(name, sizeof(name), "_S_%016llx", (unsigned long long)h);
    std::lock_guard<std::mutex> lk(g_struct_mtx);
    if (g_struct_index.count(name)) return name;
    std::string def = std::string("struct ") + name + "\n{\n    typedef " + name + " CurType;\n" + body + "};\n";
    g_struct_index[name] = g_structs.size();
    g_structs.push_back({name, def});
    return name;
}

static void collect_struct_refs(const std::string& text, std::vector<size_t>& out)
{
    size_t pos = 0;
    while ((pos = text.find("_S_", pos)) != std::string::npos) {
        if (pos + 19 <= text.size()) {
            std::string nm = text.substr(pos, 19);
            auto it = g_struct_index.find(nm);
            if (it != g_struct_index.end() &&
                std::find(out.begin(), out.end(), it->second) == out.end()) {
                out.push_back(it->second);
                collect_struct_refs(g_structs[it->second].second, out);
            }
        }
        pos += 3;
    }
}

std::string struct_defs_for(const std::string& text)
{
    std::lock_guard<std::mutex> lk(g_struct_mtx);
    std::vector<size_t> idx;
    collect_struct_refs(text, idx);
    std::sort(idx.begin(), idx.end());
    std::string s;
    for (size_t i : idx) s += g_structs[i].second;
    return s;
}

// ================================================================ memory pool
namespace {
struct Pool
{
    std::mutex mtx;
    // (stream, size class) -> free blocks
    std::map<std::pair<hipStream_t, size_t>, std::vector<void*>> free_blocks;
    std::unordered_map<void*, size_t> sizes;
};
Pool& pool()
{
    static Pool* p = new Pool();  // never destroyed: blocks may be freed during exit
    return *p;
}
size_t size_class(size_t b)
{
    if (b <= 256) return 256;
    if (b >= (1u << 20)) return (b + (2u << 20) - 1) & ~(size_t)((2u << 20) - 1);
    size_t c = 256;
    while (c < b) c <<= 1;
    return c;
}
}  // namespace

void* pool_alloc(size_t bytes, hipStream_t s)
{
    Pool& P = pool();
    size_t cls = size_class(bytes);
    {
        std::lock_guard<std::mutex> lk(P.mtx);
        auto it = P.free_blocks.find({s, cls});
        if (it != P.free_blocks.end() && !it->second.empty()) {
            void* p = it->second.back();
            it->second.pop_back();
            return p;
        }
    }
    void* p = nullptr;
    if (hipMalloc(&p, cls) != hipSuccess) {
        // release cached blocks and retry once
        std::vector<void*> to_free;
        {
            std::lock_guard<std::mutex> lk(P.mtx);
            for (auto& kv : P.free_blocks) {
                for (void* q : kv.second) to_free.push_back(q);
                kv.second.clear();
            }
            for (void* q : to_free) P.sizes.erase(q);
        }
        hipDeviceSynchronize();
        for (void* q : to_free) hipFree(q);
        (void)hipGetLastError();
        if (hipMalloc(&p, cls) != hipSuccess) {
            TRTC_ERR("device allocation of %zu bytes failed", cls);
            return nullptr;
        }
    }
    std::lock_guard<std::mutex> lk(P.mtx);
    P.sizes[p] = cls;
    return p;
}

void pool_free(void* p, hipStream_t s)
{
    if (!p) return;
    Pool& P = pool();
    std::lock_guard<std::mutex> lk(P.mtx);
    auto it = P.sizes.find(p);
    if (it == P.sizes.end()) return;
    P.free_blocks[{s, it->second}].push_back(p);
}

// Per-stream histogram block of the radix sort: KB*256 64-bit counters and
// the pre-pass's completion counter, all zero between sorts (the pre-pass's
// last workgroup clears what it used), so a sort needs no clearing launch
// before its pre-pass.  Sorts on one stream run one after another; another
// stream has its own block.  Keyed by (current device, stream): the null
// stream is one handle on every device.  Allocated and cleared on a stream's
// first sort, kept until the process ends.
void* radix_hist_block(hipStream_t s)
{
    static std::mutex mtx;
    static std::map<std::pair<int, hipStream_t>, void*>* blocks = new std::map<std::pair<int, hipStream_t>, void*>();
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) {
        (void)hipGetLastError();
        return nullptr;
    }
    std::lock_guard<std::mutex> lk(mtx);
    const std::pair<int, hipStream_t> key(dev, s);
    auto it = blocks->find(key);
    if (it != blocks->end()) return it->second;
    void* p = nullptr;
    if (hipMalloc(&p, kRadixHistBlockBytes) != hipSuccess) {
        (void)hipGetLastError();
        return nullptr;
    }
    if (!fill_u32(p, kRadixHistBlockBytes / 4, 0, s)) {
        hipFree(p);
        return nullptr;
    }
    (*blocks)[key] = p;
    return p;
}

// ================================================================ kernels
namespace {
struct Registry
{
    std::mutex mtx;
    std::unordered_map<std::string, const void*> aot;
    std::unordered_map<uint64_t, KernelRef> jit;   // source hash -> function
    std::vector<hipModule_t> modules;
    std::unordered_map<std::string, std::unique_ptr<std::string>> names;
};
Registry& registry()
{
    static Registry* r = new Registry();
    return *r;
}

std::string cache_dir()
{
    const char* d = getenv("TRTC_CACHE_DIR");
    if (d && *d) return d;
    const char* x = getenv("XDG_CACHE_HOME");
    if (x && *x) return std::string(x) + "/thrustrtc_amd";
    const char* h = getenv("HOME");
    if (h && *h) return std::string(h) + "/.cache/thrustrtc_amd";
    return "";
}

void mkdirs(const std::string& path)
{
    std::string cur;
    for (size_t i = 0; i < path.size(); i++) {
        cur += path[i];
        if (path[i] == '/' && cur.size() > 1) mkdir(cur.c_str(), 0755);
    }
    mkdir(path.c_str(), 0755);
}

bool read_file(const std::string& fn, std::string& out)
{
    std::ifstream f(fn, std::ios::binary);
    if (!f) return false;
    std::stringstream ss;
    ss << f.rdbuf();
    out = ss.str();
    return true;
}

void write_file_atomic(const std::string& fn, const std::string& data)
{
    std::string tmp = fn + ".tmp" + std::to_string(getpid());
    {
        std::ofstream f(tmp, std::ios::binary);
        if (!f) return;
        f.write(data.data(), (std::streamsize)data.size());
        if (!f) return;
    }
    rename(tmp.c_str(), fn.c_str());
}

std::string hiprtc_version()
{
    int a = 0, b = 0;
    hiprtcVersion(&a, &b);
    return std::to_string(a) + "." + std::to_string(b);
}

// Full source of a JIT translation unit.
std::string assemble_source(const std::string& body_text, const std::string& type_text, bool templates)
{
    Context& c = context();
    std::string src;
    {
        std::lock_guard<std::mutex> lk(c.mtx);
        for (auto& b : c.code_blocks) src += b;
    }
    src += templates ? "#include \"trtc_kernels.hpp\"\n" : "#include \"trtc_builtin.hpp\"\n";
    src += struct_defs_for(type_text);
    src += body_text;
    return src;
}

// Compile with hipRTC.  name_expr empty: no template instantiation request.
bool rtc_compile(const std::string& src, const std::string& name_expr, const std::string& arch,
                 std::string& code, std::string& lowered, std::string* log_out)
{
    Context& c = context();
    std::vector<std::string> hnames, hcont;
    for (int i = 0; i < k_embedded_count; i++) {
        hnames.push_back(k_embedded_names[i]);
        hcont.push_back(k_embedded_sources[i]);
    }
    std::vector<std::string> opts_s = {"--offload-arch=" + arch, "-O3", "-std=c++17", "-Xclang", "-target-feature",
                                       "-Xclang", "-packed-fp32-ops", "-DDEVICE_ONLY"};
    {
        std::lock_guard<std::mutex> lk(c.mtx);
        for (size_t i = 0; i < c.header_names.size(); i++) {
            hnames.push_back(c.header_names[i]);
            hcont.push_back(c.header_contents[i]);
        }
        for (auto& d : c.include_dirs) opts_s.push_back("-I" + d);
    }
    std::vector<const char*> hn, hc, opts;
    for (auto& s : hnames) hn.push_back(s.c_str());
    for (auto& s : hcont) hc.push_back(s.c_str());
    for (auto& s : opts_s) opts.push_back(s.c_str());

    hiprtcProgram prog;
    if (hiprtcCreateProgram(&prog, src.c_str(), "trtc_jit.hip", (int)hn.size(), hc.data(), hn.data()) !=
        HIPRTC_SUCCESS) {
        TRTC_ERR("hiprtcCreateProgram failed");
        return false;
    }
    if (!name_expr.empty()) hiprtcAddNameExpression(prog, name_expr.c_str());
    hiprtcResult r = hiprtcCompileProgram(prog, (int)opts.size(), opts.data());
    size_t ls = 0;
    hiprtcGetProgramLogSize(prog, &ls);
    std::string log(ls, '\0');
    if (ls) hiprtcGetProgramLog(prog, &log[0]);
    if (log_out) *log_out = log;
    if (r != HIPRTC_SUCCESS) {
        TRTC_ERR("hipRTC compilation failed:\n%s\n---- source ----\n%s", log.c_str(), src.c_str());
        hiprtcDestroyProgram(&prog);
        return false;
    }
    if (c.verbose) fprintf(stderr, "---- trtc_jit.hip (%s) ----\n%s\n", name_expr.c_str(), src.c_str());
    if (!name_expr.empty()) {
        const char* low = nullptr;
        if (hiprtcGetLoweredName(prog, name_expr.c_str(), &low) != HIPRTC_SUCCESS || !low) {
            TRTC_ERR("hipRTC: no lowered name for %s", name_expr.c_str());
            hiprtcDestroyProgram(&prog);
            return false;
        }
        lowered = low;
    }
    size_t cs = 0;
    hiprtcGetCodeSize(prog, &cs);
    code.resize(cs);
    hiprtcGetCode(prog, &code[0]);
    hiprtcDestroyProgram(&prog);
    return true;
}

// Hash of the embedded device headers, computed once: every JIT key covers
// them, and hashing their ~136 KB on each lookup cost ~0.2 ms per call.
uint64_t embedded_sources_hash()
{
    static const uint64_t h = [] {
        uint64_t x = fnv1a("trtc-embedded", 13);
        for (int i = 0; i < k_embedded_count; i++) {
            x = fnv1a(k_embedded_names[i], strlen(k_embedded_names[i]), x);
            x = fnv1a(k_embedded_sources[i], strlen(k_embedded_sources[i]), x);
        }
        return x;
    }();
    return h;
}

struct JitStats
{
    std::atomic<unsigned long long> mem_hits{0}, disk_hits{0}, compiles{0}, rejected{0};
};
JitStats& jit_stats()
{
    static JitStats* s = new JitStats();
    return *s;
}

// Disk cache entry: "TRTCCO2 <code bytes> <fnv1a-64 of the code, hex> <lowered name>\n" followed by the
// code object.  compile + write = make_cache_entry; read = parse_cache_entry, which checks the length and
// the checksum on the host so that a truncated, damaged or foreign file never reaches hipModuleLoadData.
std::string make_cache_entry(const std::string& lowered, const std::string& code)
{
    char head[96];
    snprintf(head, sizeof(head), "TRTCCO2 %zu %016llx ", code.size(),
             (unsigned long long)fnv1a(code.data(), code.size()));
    return head + lowered + "\n" + code;
}
}  // namespace

bool parse_cache_entry(const std::string& blob, std::string& lowered, std::string& code)
{
    static const char magic[] = "TRTCCO2 ";
    if (blob.size() < sizeof(magic) - 1 || blob.compare(0, sizeof(magic) - 1, magic) != 0) return false;
    size_t nl = blob.find('\n');
    if (nl == std::string::npos || nl > 4096) return false;
    const std::string head = blob.substr(sizeof(magic) - 1, nl - (sizeof(magic) - 1));
    unsigned long long len = 0, sum = 0;
    int used = 0;
    if (sscanf(head.c_str(), "%llu %16llx %n", &len, &sum, &used) != 2 || used <= 0) return false;
    const std::string name = head.substr((size_t)used);
    if (name.empty() || name.find(' ') != std::string::npos) return false;
    if (len == 0 || blob.size() - (nl + 1) != len) return false;
    if (fnv1a(blob.data() + nl + 1, (size_t)len) != sum) return false;
    lowered = name;
    code = blob.substr(nl + 1);
    return true;
}

void jit_cache_stats(unsigned long long out[4])
{
    JitStats& st = jit_stats();
    out[0] = st.mem_hits;
    out[1] = st.disk_hits;
    out[2] = st.compiles;
    out[3] = st.rejected;
}

namespace {
// Kernel for a JIT translation unit: the in-memory table, then the disk cache
// (keyed by source, arch and hipRTC version), then hipRTC.  With compile ==
// false only the in-memory table is consulted (no compilation, no module load).
bool jit_get(const std::string& src, const std::string& name_expr, const std::string& fn_name, KernelRef& out,
             bool compile = true)
{
    Context& c = context();
    std::string key_text = src + '\x01' + name_expr + '\x01' + c.arch + '\x01' + hiprtc_version();
    uint64_t key = fnv1a(key_text.data(), key_text.size(), embedded_sources_hash());
    Registry& R = registry();
    JitStats& stats = jit_stats();
    {
        std::lock_guard<std::mutex> lk(R.mtx);
        auto it = R.jit.find(key);
        if (it != R.jit.end()) {
            out = it->second;
            stats.mem_hits++;
            return true;
        }
    }
    if (!compile) return false;
    std::string code, lowered = fn_name;
    std::string dir = cache_dir();
    char hex[32];
    snprintf(hex, sizeof(hex), "%016llx", (unsigned long long)key);
    std::string safe_arch = c.arch;
    for (auto& ch : safe_arch)
        if (ch == ':' || ch == '+' || ch == '/') ch = '_';
    std::string fn = dir.empty() ? "" : dir + "/" + safe_arch + "/" + hex + ".co";
    bool have = false;
    if (!fn.empty()) {
        std::string blob;
        if (read_file(fn, blob)) {
            have = parse_cache_entry(blob, lowered, code);
            if (!have) {
                lowered = fn_name;
                stats.rejected++;
            }
        }
    }
    if (have) {
        stats.disk_hits++;
    } else {
        if (!rtc_compile(src, name_expr, c.arch, code, lowered, nullptr)) return false;
        stats.compiles++;
        if (name_expr.empty()) lowered = fn_name;
        if (!fn.empty()) {
            mkdirs(dir + "/" + safe_arch);
            write_file_atomic(fn, make_cache_entry(lowered, code));
        }
    }
    hipModule_t mod;
    if (hipModuleLoadData(&mod, code.data()) != hipSuccess) {
        (void)hipGetLastError();
        if (have) {  // a valid entry the loader still refuses (e.g. another driver): recompile once
            if (!rtc_compile(src, name_expr, c.arch, code, lowered, nullptr)) return false;
            stats.compiles++;
            if (name_expr.empty()) lowered = fn_name;
            if (!fn.empty()) write_file_atomic(fn, make_cache_entry(lowered, code));
            if (!check_hip(hipModuleLoadData(&mod, code.data()), "hipModuleLoadData")) return false;
        } else {
            TRTC_ERR("hipModuleLoadData failed");
            return false;
        }
    }
    hipFunction_t f;
    if (!check_hip(hipModuleGetFunction(&f, mod, lowered.c_str()), "hipModuleGetFunction")) return false;
    // upload constant objects declared by code blocks
    {
        std::lock_guard<std::mutex> lk(c.mtx);
        for (auto& co : c.constants) {
            hipDeviceptr_t dp;
            size_t bytes = 0;
            if (hipModuleGetGlobal(&dp, &bytes, mod, co.first.c_str()) == hipSuccess)
                hipMemcpyHtoD(dp, (void*)co.second.data(), std::min(bytes, co.second.size()));
            else
                (void)hipGetLastError();
        }
    }
    KernelRef k;
    k.jit = f;
    std::lock_guard<std::mutex> lk(R.mtx);
    R.modules.push_back(mod);
    R.jit[key] = k;
    out = k;
    return true;
}
}  // namespace

const char* stable_name(const std::string& n)
{
    Registry& R = registry();
    std::lock_guard<std::mutex> lk(R.mtx);
    auto& p = R.names[n];
    if (!p) p.reset(new std::string(n));
    return p->c_str();
}

void register_aot(const char* name_expr, const void* fn)
{
    Registry& R = registry();
    std::lock_guard<std::mutex> lk(R.mtx);
    R.aot[name_expr] = fn;
}

const void* find_aot(const std::string& name_expr)
{
    Registry& R = registry();
    std::lock_guard<std::mutex> lk(R.mtx);
    auto it = R.aot.find(name_expr);
    return it == R.aot.end() ? nullptr : it->second;
}

bool get_template_kernel(const std::string& name_expr, const std::string& type_text, KernelRef& out)
{
    const void* a = find_aot(name_expr);
    if (a) {
        out.aot = a;
        out.jit = nullptr;
        out.name = stable_name(name_expr);
        return true;
    }
    std::string src = assemble_source("", type_text, true);
    if (!jit_get(src, name_expr, "", out)) return false;
    out.name = stable_name(name_expr);
    return true;
}

bool find_template_kernel(const std::string& name_expr, const std::string& type_text, KernelRef& out)
{
    const void* a = find_aot(name_expr);
    if (a) {
        out.aot = a;
        out.jit = nullptr;
        out.name = stable_name(name_expr);
        return true;
    }
    std::string src = assemble_source("", type_text, true);
    if (!jit_get(src, name_expr, "", out, false)) return false;
    out.name = stable_name(name_expr);
    return true;
}

// Resolved kernels of generated sources, keyed by what the source is built
// from plus the context generation (code blocks, headers, include
// directories and constants change the source of every later kernel): a
// repeated launch skips building and hashing the whole source.
bool memo_kernel(const std::string& key, const std::function<bool(KernelRef&)>& build, KernelRef& out)
{
    static std::mutex mtx;
    static std::unordered_map<std::string, KernelRef>* memo = new std::unordered_map<std::string, KernelRef>();
    const std::string full = key + "\x02" + std::to_string(context().generation.load());
    {
        std::lock_guard<std::mutex> lk(mtx);
        auto it = memo->find(full);
        if (it != memo->end()) {
            out = it->second;
            return true;
        }
    }
    if (!build(out)) return false;
    std::lock_guard<std::mutex> lk(mtx);
    (*memo)[full] = out;
    return true;
}

bool get_source_kernel(const std::string& kernel_src, const std::string& fn_name, KernelRef& out)
{
    std::string src = assemble_source(kernel_src, kernel_src, false);
    if (!jit_get(src, "", fn_name, out)) return false;
    out.name = stable_name(fn_name);
    return true;
}

bool compile_only(const std::string& arch, const std::string& name_expr, const std::string& type_text, std::string* log)
{
    std::string src = assemble_source("", type_text, true);
    std::string code, lowered;
    return rtc_compile(src, name_expr, arch, code, lowered, log);
}

bool sync_if_debug(hipStream_t s, const char* what)
{
    if (!context().kernel_debug) return true;
    return check_hip(hipStreamSynchronize(s), what);
}

namespace {
struct ProfRec
{
    const char* name;
    hipEvent_t a, b;
};
struct Prof
{
    std::mutex mtx;
    bool on = false;
    std::vector<ProfRec> recs;
    std::vector<hipEvent_t> spare;
    hipEvent_t get()
    {
        if (!spare.empty()) {
            hipEvent_t e = spare.back();
            spare.pop_back();
            return e;
        }
        hipEvent_t e;
        hipEventCreate(&e);
        return e;
    }
};
Prof& prof()
{
    static Prof* p = new Prof();
    return *p;
}
}  // namespace

void profile_begin()
{
    Prof& P = prof();
    std::lock_guard<std::mutex> lk(P.mtx);
    for (auto& r : P.recs) {
        P.spare.push_back(r.a);
        P.spare.push_back(r.b);
    }
    P.recs.clear();
    P.on = true;
}

std::string profile_end()
{
    Prof& P = prof();
    std::lock_guard<std::mutex> lk(P.mtx);
    P.on = false;
    std::map<std::string, std::pair<size_t, double>> agg;
    for (auto& r : P.recs) {
        hipEventSynchronize(r.b);
        float ms = 0;
        hipEventElapsedTime(&ms, r.a, r.b);
        auto& x = agg[r.name];
        x.first += 1;
        x.second += ms;
    }
    std::string js = "{";
    bool first = true;
    for (auto& kv : agg) {
        char buf[128];
        snprintf(buf, sizeof(buf), "\": {\"count\": %zu, \"total_ms\": %.6f}", kv.second.first, kv.second.second);
        js += std::string(first ? "" : ", ") + "\"" + kv.first + buf;
        first = false;
    }
    return js + "}";
}

namespace {
thread_local bool t_capturing = false;
}
void set_capturing(bool on) { t_capturing = on; }
bool capturing() { return t_capturing; }

bool launch_graph(hipGraphExec_t g, const char* name, hipStream_t s)
{
    Prof& P = prof();
    ProfRec rec{nullptr, nullptr, nullptr};
    if (P.on) {
        std::lock_guard<std::mutex> lk(P.mtx);
        rec.name = name;
        rec.a = P.get();
        rec.b = P.get();
        hipEventRecord(rec.a, s);
    }
    hipError_t e = hipGraphLaunch(g, s);
    if (rec.name) {
        std::lock_guard<std::mutex> lk(P.mtx);
        hipEventRecord(rec.b, s);
        P.recs.push_back(rec);
    }
    if (!check_hip(e, "hipGraphLaunch")) return false;
    return sync_if_debug(s, "graph execution");
}

// ---------------------------------------------------------------- pinned staging
namespace {
struct Pinned
{
    std::mutex mtx;
    std::multimap<size_t, void*> free_blocks;                    // size -> block
    std::vector<std::tuple<void*, size_t, hipEvent_t>> pending;  // released, awaiting the stream
    std::unordered_map<void*, size_t> sizes;
};
Pinned& pinned()
{
    static Pinned* p = new Pinned();
    return *p;
}
}  // namespace

void* pinned_alloc(size_t bytes)
{
    Pinned& P = pinned();
    std::lock_guard<std::mutex> lk(P.mtx);
    for (size_t i = 0; i < P.pending.size();) {
        auto& t = P.pending[i];
        if (hipEventQuery(std::get<2>(t)) == hipSuccess) {
            hipEventDestroy(std::get<2>(t));
            P.free_blocks.emplace(std::get<1>(t), std::get<0>(t));
            P.pending[i] = P.pending.back();
            P.pending.pop_back();
        } else {
            i++;
        }
    }
    size_t cls = 4096;
    while (cls < bytes) cls <<= 1;
    auto it = P.free_blocks.find(cls);
    if (it != P.free_blocks.end()) {
        void* p = it->second;
        P.free_blocks.erase(it);
        return p;
    }
    void* p = nullptr;
    if (!check_hip(hipHostMalloc(&p, cls, hipHostMallocDefault), "hipHostMalloc")) return nullptr;
    P.sizes[p] = cls;
    return p;
}

void pinned_release(void* p, hipStream_t s)
{
    if (!p) return;
    Pinned& P = pinned();
    std::lock_guard<std::mutex> lk(P.mtx);
    hipEvent_t e;
    if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) return;  // leaks p; never reused
    hipEventRecord(e, s);
    P.pending.emplace_back(p, P.sizes[p], e);
}

bool launch_kernel(const KernelRef& k, dim3 grid, dim3 block, void** args, size_t shmem, hipStream_t s)
{
    if (grid.x == 0) return true;
    Prof& P = prof();
    ProfRec rec{nullptr, nullptr, nullptr};
    if (P.on && !t_capturing) {
        std::lock_guard<std::mutex> lk(P.mtx);
        rec.name = k.name;
        rec.a = P.get();
        rec.b = P.get();
        hipEventRecord(rec.a, s);
    }
    hipError_t e;
    if (k.aot) e = hipLaunchKernel(k.aot, grid, block, args, shmem, s);
    else if (k.jit)
        e = hipModuleLaunchKernel(k.jit, grid.x, grid.y, grid.z, block.x, block.y, block.z, (unsigned)shmem, s, args,
                                  nullptr);
    else return false;
    if (rec.name) {
        std::lock_guard<std::mutex> lk(P.mtx);
        hipEventRecord(rec.b, s);
        P.recs.push_back(rec);
    }
    if (!check_hip(e, "kernel launch")) return false;
    if (t_capturing) return true;
    return sync_if_debug(s, "kernel execution");
}

static bool get_builtin(const char* name, KernelRef& k)
{
    const void* a = find_aot(name);
    if (!a) {
        TRTC_ERR("built-in kernel %s missing", name);
        return false;
    }
    k.aot = a;
    k.name = name;
    return true;
}

static unsigned grid_for(size_t n, unsigned block)
{
    size_t g = (n + block - 1) / block;
    size_t cap = (size_t)std::max(1, context().num_cus) * 16;
    if (g > cap) g = cap;
    return (unsigned)std::max<size_t>(g, 1);
}

bool fill_u32(void* p, size_t words, unsigned v, hipStream_t s)
{
    if (!words) return true;
    KernelRef k;
    if (!get_builtin("trtc_fill_u32", k)) return false;
    void* args[] = {&p, &words, &v};
    return launch_kernel(k, dim3(grid_for(words, 256)), dim3(256), args, 0, s);
}

bool clear_u32_agent(void* p, size_t words, hipStream_t s)
{
    if (!words) return true;
    KernelRef k;
    if (!get_builtin("trtc_clear_u32_agent", k)) return false;
    void* args[] = {&p, &words};
    return launch_kernel(k, dim3(grid_for(words, 256)), dim3(256), args, 0, s);
}

bool copy_bytes(const void* src, void* dst, size_t nbytes, hipStream_t s)
{
    if (!nbytes) return true;
    if ((((size_t)src | (size_t)dst) & 15) != 0)
        return check_hip(hipMemcpyAsync(dst, src, nbytes, hipMemcpyDeviceToDevice, s), "hipMemcpyAsync");
    KernelRef k;
    if (!get_builtin("trtc_copy_bytes16", k)) return false;
    void* args[] = {&src, &dst, &nbytes};
    // one vector per thread up to 16 GiB (a grid-strided loop keeps 16-byte
    // copies ~10% below this), grid-stride beyond: the grid stays under HIP's
    // 2^32 work-item limit
    return launch_kernel(k, dim3(capped_grid(nbytes / 16, 256, (size_t)1 << 22)), dim3(256), args, 0, s);
}

bool iota_u32(unsigned* p, size_t n, hipStream_t s)
{
    if (!n) return true;
    KernelRef k;
    if (!get_builtin("trtc_iota_u32", k)) return false;
    void* args[] = {&p, &n};
    return launch_kernel(k, dim3(grid_for(n, 256)), dim3(256), args, 0, s);
}

bool gather_bytes(const void* src, void* dst, const unsigned* idx, size_t n, size_t elem, hipStream_t s)
{
    if (!n) return true;
    KernelRef k;
    if (!get_builtin("trtc_gather_bytes", k)) return false;
    void* args[] = {&src, &dst, &idx, &n, &elem};
    return launch_kernel(k, dim3(grid_for(n, 256)), dim3(256), args, 0, s);
}

void* contiguous_data(const DVVectorLike& v)
{
    if (auto p = dynamic_cast<const DVVector*>(&v)) return p->data();
    if (auto p = dynamic_cast<const DVVectorAdaptor*>(&v)) return p->data();
    if (auto p = dynamic_cast<const DVRange*>(&v)) return p->contiguous();
    return nullptr;
}

}  // namespace trtc_host

using namespace trtc_host;

// ================================================================ public context API
static std::string g_hiprtc_path;

void set_libnvrtc_path(const char* path) { g_hiprtc_path = path ? path : ""; }

bool TRTC_Try_Init() { return ensure_init(); }

int TRTC_Get_PTX_Arch()
{
    if (!ensure_init()) return -1;
    return context().arch_num;
}

void TRTC_Set_Verbose(bool verbose) { context().verbose = verbose; }
void TRTC_Set_Kernel_Debug(bool kernel_debug) { context().kernel_debug = kernel_debug; }

void TRTC_Add_Include_Dir(const char* path)
{
    Context& c = context();
    std::lock_guard<std::mutex> lk(c.mtx);
    c.include_dirs.push_back(path);
    c.generation++;
}

void TRTC_Add_Built_In_Header(const char* name, const char* content)
{
    Context& c = context();
    std::lock_guard<std::mutex> lk(c.mtx);
    c.header_names.push_back(name);
    c.header_contents.push_back(content);
    c.generation++;
}

void TRTC_Add_Code_Block(const char* code)
{
    Context& c = context();
    std::lock_guard<std::mutex> lk(c.mtx);
    c.code_blocks.push_back(code);
    c.generation++;
}

void TRTC_Add_Inlcude_Filename(const char* fn)
{
    std::string line = std::string("#include \"") + fn + "\"\n";
    TRTC_Add_Code_Block(line.c_str());
}

void TRTC_Add_Constant_Object(const char* name, const DeviceViewable& obj)
{
    std::string line = "__constant__ " + obj.name_view_cls() + " " + name + ";\n";
    TRTC_Add_Code_Block(line.c_str());
    Context& c = context();
    std::lock_guard<std::mutex> lk(c.mtx);
    c.constants.push_back({name, obj.view()});
    c.generation++;
}

std::string TRTC_Add_Struct(const char* struct_body) { return add_struct(struct_body); }

void TRTC_Wait()
{
    if (!ensure_init()) return;
    check_hip(hipDeviceSynchronize(), "hipDeviceSynchronize");
}

void TRTC_Set_Stream(void* stream) { set_stream((hipStream_t)stream); }
void* TRTC_Get_Stream() { return (void*)cur_stream(); }

// Size / member offsets of a device type, measured on the device by a probe
// kernel (the reference compiles a similar probe, impl_context.inl:372-552).
size_t TRTC_Size_Of(const char* cls)
{
    const ScalarInfo* si = scalar_info(cls);
    if (si) return si->size;
    std::string c = cls;
    if (c.compare(0, 11, "VectorView<") == 0) return 16;
    {
        // structs whose layout the host computed (functors, views, zipped elements)
        std::lock_guard<std::mutex> lk(g_layout_mtx);
        auto it = g_struct_layout.find(c);
        if (it != g_struct_layout.end()) return it->second.first;
    }
    static std::mutex m;
    static std::unordered_map<std::string, size_t> cache;
    {
        std::lock_guard<std::mutex> lk(m);
        auto it = cache.find(c);
        if (it != cache.end()) return it->second;
    }
    if (!ensure_init()) return 0;
    std::string src = "extern \"C\" __global__ void trtc_sizeof_probe(size_t* out) { out[0] = sizeof(" + c + "); }\n";
    KernelRef k;
    if (!get_source_kernel(src, "trtc_sizeof_probe", k)) return 0;
    hipStream_t s = cur_stream();
    PoolBuf buf(sizeof(size_t), s);
    size_t* p = buf.as<size_t>();
    void* args[] = {&p};
    if (!launch_kernel(k, dim3(1), dim3(1), args, 0, s)) return 0;
    size_t r = 0;
    if (!check_hip(hipMemcpyAsync(&r, p, sizeof(size_t), hipMemcpyDeviceToHost, s), "hipMemcpyAsync")) return 0;
    hipStreamSynchronize(s);
    std::lock_guard<std::mutex> lk(m);
    cache[c] = r;
    return r;
}

bool TRTC_Query_Struct(const char* name_struct, const std::vector<const char*>& name_members, size_t* offsets)
{
    if (!ensure_init()) return false;
    std::string src = "extern \"C\" __global__ void trtc_offsetof_probe(size_t* out) {\n";
    size_t n = name_members.size();
    for (size_t i = 0; i < n; i++)
        src += "    out[" + std::to_string(i) + "] = __builtin_offsetof(" + name_struct + ", " + name_members[i] + ");\n";
    src += "    out[" + std::to_string(n) + "] = sizeof(" + name_struct + ");\n}\n";
    KernelRef k;
    if (!get_source_kernel(src, "trtc_offsetof_probe", k)) return false;
    hipStream_t s = cur_stream();
    PoolBuf buf(sizeof(size_t) * (n + 1), s);
    size_t* p = buf.as<size_t>();
    void* args[] = {&p};
    if (!launch_kernel(k, dim3(1), dim3(1), args, 0, s)) return false;
    if (!check_hip(hipMemcpyAsync(offsets, p, sizeof(size_t) * (n + 1), hipMemcpyDeviceToHost, s), "hipMemcpyAsync"))
        return false;
    return check_hip(hipStreamSynchronize(s), "hipStreamSynchronize");
}

DeviceViewable* dv_from_viewbuf(const ViewBuf& buf, const char* type)
{
    const ScalarInfo* si = scalar_info(type);
    if (si) {
        switch (si->code) {
        case TC_I8: return new DVInt8(*(int8_t*)buf.data());
        case TC_U8: return new DVUInt8(*(uint8_t*)buf.data());
        case TC_I16: return new DVInt16(*(int16_t*)buf.data());
        case TC_U16: return new DVUInt16(*(uint16_t*)buf.data());
        case TC_I32: return new DVInt32(*(int32_t*)buf.data());
        case TC_U32: return new DVUInt32(*(uint32_t*)buf.data());
        case TC_I64: return new DVInt64(*(int64_t*)buf.data());
        case TC_U64: return new DVUInt64(*(uint64_t*)buf.data());
        case TC_F32: return new DVFloat(*(float*)buf.data());
        case TC_F64: return new DVDouble(*(double*)buf.data());
        case TC_BOOL: return new DVBool(*(bool*)buf.data());
        default: break;
        }
    }
    return new SomeDeviceViewable(type, buf.data(), buf.size());
}

// ================================================================ user kernels
// TRTC_Kernel: `extern "C" __global__ void saxpy(<view types> <names>) { body }`
// instantiated per combination of argument view types (reference
// impl_context.inl:554-600, TRTCContext.cpp:173-207).
static std::string kernel_signature(const std::vector<std::string>& names, const DeviceViewable** args)
{
    std::string s;
    for (size_t i = 0; i < names.size(); i++) {
        if (i) s += ", ";
        s += args[i]->name_view_cls() + " " + names[i];
    }
    return s;
}

TRTC_Kernel::TRTC_Kernel(const std::vector<const char*>& param_names, const char* code_body)
{
    for (auto p : param_names) m_param_names.push_back(p);
    m_code_body = code_body;
}

static bool build_user_kernel(const std::vector<std::string>& names, const std::string& body,
                              const DeviceViewable** args, KernelRef& k)
{
    std::string src = "extern \"C\" __global__ void saxpy(" + kernel_signature(names, args) + ")\n{\n" + body + "\n}\n";
    return get_source_kernel(src, "saxpy", k);
}

bool TRTC_Kernel::calc_optimal_block_size(const DeviceViewable** args, int& sizeBlock, unsigned sharedMemBytes)
{
    if (!ensure_init()) return false;
    KernelRef k;
    if (!build_user_kernel(m_param_names, m_code_body, args, k)) return false;
    int grid = 0;
    return check_hip(hipModuleOccupancyMaxPotentialBlockSize(&grid, &sizeBlock, k.jit, sharedMemBytes, 0),
                     "hipModuleOccupancyMaxPotentialBlockSize");
}

bool TRTC_Kernel::calc_number_blocks(const DeviceViewable** args, int sizeBlock, int& numBlocks, unsigned sharedMemBytes)
{
    if (!ensure_init()) return false;
    KernelRef k;
    if (!build_user_kernel(m_param_names, m_code_body, args, k)) return false;
    int per_cu = 0;
    if (!check_hip(hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k.jit, sizeBlock, sharedMemBytes),
                   "hipModuleOccupancyMaxActiveBlocksPerMultiprocessor"))
        return false;
    numBlocks = per_cu * context().num_cus;
    return true;
}

bool TRTC_Kernel::launch(dim_type gridDim, dim_type blockDim, const DeviceViewable** args, unsigned sharedMemBytes)
{
    if (!ensure_init()) return false;
    KernelRef k;
    if (!build_user_kernel(m_param_names, m_code_body, args, k)) return false;
    size_t n = m_param_names.size();
    std::vector<ViewBuf> views(n);
    std::vector<void*> ptrs(n);
    for (size_t i = 0; i < n; i++) {
        views[i] = args[i]->view();
        ptrs[i] = views[i].data();
    }
    return launch_kernel(k, dim3(gridDim.x, gridDim.y, gridDim.z), dim3(blockDim.x, blockDim.y, blockDim.z),
                         ptrs.data(), sharedMemBytes, cur_stream());
}

// TRTC_For: the body runs once per index `name_iter` in [begin, end)
// (reference impl_context.inl:801-898).
TRTC_For::TRTC_For(const std::vector<const char*>& param_names, const char* name_iter, const char* code_body)
{
    for (auto p : param_names) m_param_names.push_back(p);
    m_name_iter = name_iter;
    m_code_body = code_body;
}

// The generated kernel: the body is a device function called once per index
// with its own copies of the parameters (as in a one-index-per-thread kernel:
// a `return` or a write to a parameter affects that index only); each thread
// takes kForItems indices kForBlock apart (their memory operations overlap
// across indices: a 2 GiB int32 copy 3.7 -> 4.9 TB/s), and an outer
// grid-stride loop covers any n with a capped grid.
constexpr unsigned kForBlock = 256, kForItems = 8;

static std::string for_source(const std::string& sig, const std::vector<std::string>& names,
                              const std::string& name_iter, const std::string& body)
{
    std::string call;
    for (const auto& nm : names) call += ", " + nm;
    const std::string B = std::to_string(kForBlock), I = std::to_string(kForItems);
    return "__device__ __forceinline__ void _trtc_for_body(size_t " + name_iter + std::string(sig.empty() ? "" : ", ") +
           sig + ")\n{\n" + body + "\n}\n"
           "extern \"C\" __global__ void __launch_bounds__(" + B + ") trtc_for(size_t _trtc_begin, size_t _trtc_end" +
           std::string(sig.empty() ? "" : ", ") + sig + ")\n{\n"
           "    const size_t _trtc_stride = (size_t)gridDim.x * " + B + " * " + I + ";\n"
           "    for (size_t _trtc_b = _trtc_begin + (size_t)blockIdx.x * " + B + " * " + I + " + threadIdx.x;\n"
           "         _trtc_b < _trtc_end; _trtc_b += _trtc_stride) {\n"
           "#pragma unroll\n"
           "        for (int _trtc_k = 0; _trtc_k < " + I + "; _trtc_k++) {\n"
           "            const size_t _trtc_i = _trtc_b + (size_t)_trtc_k * " + B + ";\n"
           "            if (_trtc_i < _trtc_end) _trtc_for_body(_trtc_i" + call + ");\n"
           "        }\n"
           "    }\n"
           "}\n";
}

// Blocks of a grid-stride launch over `items` work units of `per_block` each,
// capped so that the launch stays far below HIP's 2^32 work-item limit.
unsigned trtc_host::capped_grid(size_t items, size_t per_block, size_t cap)
{
    size_t g = (items + per_block - 1) / per_block;
    return (unsigned)std::max<size_t>(1, std::min(g, cap));
}

bool TRTC_For::launch(size_t begin, size_t end, const DeviceViewable** args)
{
    if (!ensure_init()) return false;
    if (end <= begin) return true;
    std::string sig = kernel_signature(m_param_names, args);
    KernelRef k;
    const std::string memo = "for\x01" + m_name_iter + "\x01" + sig + "\x01" + m_code_body;
    if (!memo_kernel(memo, [&](KernelRef& out) {
            return get_source_kernel(for_source(sig, m_param_names, m_name_iter, m_code_body), "trtc_for", out);
        }, k))
        return false;
    size_t n = m_param_names.size();
    std::vector<ViewBuf> views(n);
    std::vector<void*> ptrs(n + 2);
    ptrs[0] = &begin;
    ptrs[1] = &end;
    for (size_t i = 0; i < n; i++) {
        views[i] = args[i]->view();
        ptrs[i + 2] = views[i].data();
    }
    const unsigned blocks = capped_grid(end - begin, (size_t)kForBlock * kForItems, (size_t)1 << 20);
    return launch_kernel(k, dim3(blocks), dim3(kForBlock), ptrs.data(), 0, cur_stream());
}

bool TRTC_For::launch_n(size_t n, const DeviceViewable** args) { return launch(0, n, args); }
