// ORACLE — test infrastructure only.
//
// A CPU restatement of the reference's primitive-core semantics (fynv/ThrustRTC),
// used as the checker by tests/, by __graft_entry__.smoke() and by the
// cpu_baseline leg of bench.py.  Nothing in the product (thrustrtc_amd/) may
// import, link or call it.
//
// Every function cites the reference file:line it restates.  Type codes:
//   0 int8  1 uint8  2 int16  3 uint16  4 int32  5 uint32
//   6 int64 7 uint64 8 float  9 double
// Binary ops: 0 Plus, 1 Maximum, 2 Minimum, 3 Multiplies   (built_in.h:14-150)
//
// Parity status of each function (pinned by tests/test_oracle.py):
//   or_sort / or_sort_by_key ..... pinned: docs known answers, shuffled 1..10000
//                                  (test_sort.cpp:82-168), brute force over all
//                                  permutations n<=7, -0/+0 and Greater-tie cases
//   or_inclusive/exclusive_scan .. pinned: docs/test_scan known answers, closed
//                                  form (i+1)(i+2)/2, numpy cumsum cross-check
//   or_scan_by_key ............... pinned: test_scan_by_key known answers,
//                                  triangular numbers per segment, per-segment scan
//   or_reduce .................... pinned: test_reduce known answers, closed form
//   or_reduce_ref_tree_f32 ....... pinned: equals exact integer sums on integer-
//                                  valued floats; brute-force tree on tiny n
//   or_copy_if ................... pinned: test_copy known answers, numpy mask
//   or_unique / or_reduce_by_key . pinned: the reference tests' inputs
//                                  (test_unique.py) with their results, groupby
//                                  restatements (itertools.groupby), per-run
//                                  numpy reductions
//   or_select / or_partition ..... pinned: test_remove / test_partition inputs
//                                  with their results, numpy boolean masks
//   or_transform_scan ............ pinned: test_transform_scan inputs with their
//                                  results, numpy cumsum of the transform
//   or_count ..................... pinned: test_count (20 of 0..1999 have
//                                  x % 100 == 47), numpy count_nonzero
//
// The parallel functions at the bottom (bl_*) are the host baseline that
// bench.py times: __gnu_parallel / OpenMP on all host cores.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <type_traits>
#include <vector>
#include <parallel/algorithm>
#include <parallel/numeric>
#include <omp.h>

namespace {

// ---------------------------------------------------------------------------
// Key map of the radix path: built_in.h:430-458 (d_u32 / d_u64).
// Integers map by a plain (sign-extending) cast; a negative float maps to
// -(bits & 0x7FFFFFFF), which makes -0.0 and +0.0 the same key.
// ---------------------------------------------------------------------------
template <class T> inline uint64_t key_map(T v)
{
    if (sizeof(T) <= 4) return (uint64_t)(uint32_t)v;
    return (uint64_t)v;
}
template <> inline uint64_t key_map<float>(float v)
{
    uint32_t u;
    memcpy(&u, &v, 4);
    if (u & 0x80000000u) u = (uint32_t)(0u - (u & 0x7FFFFFFFu));
    return u;
}
template <> inline uint64_t key_map<double>(double v)
{
    uint64_t u;
    memcpy(&u, &v, 8);
    if (u & 0x8000000000000000ull) u = 0ull - (u & 0x7FFFFFFFFFFFFFFFull);
    return u;
}

// Argmin with the reference's comparison `vec[i2] < vec[i1] ? i2 : i1`
// (extrema.cpp:7) evaluated left to right: the first minimal element.
template <class T> size_t arg_min(const T* v, size_t n)
{
    size_t best = 0;
    for (size_t i = 1; i < n; i++)
        if (v[i] < v[best]) best = i;
    return best;
}

// Sort order of radix_sort_32/64 (radix_sort.cpp:165-433): the key is
// diff = d_u(x) - d_u(min) modulo 2^32 (2^64 for 64-bit types), and every
// one-bit pass is a stable split, so the result is the stable sort by diff.
// "Greater" (radix_sort_reverse_*) puts the 1-bit side first in each pass:
// the stable sort by diff in descending order (ties keep input order).
template <class T>
void sort_perm(const T* keys, size_t n, int descending, std::vector<size_t>& perm)
{
    perm.resize(n);
    std::iota(perm.begin(), perm.end(), (size_t)0);
    if (n < 2) return;
    size_t im = arg_min(keys, n);
    uint64_t kmin = key_map(keys[im]);
    uint64_t mask = sizeof(T) <= 4 ? 0xFFFFFFFFull : ~0ull;
    std::vector<uint64_t> diff(n);
    for (size_t i = 0; i < n; i++) diff[i] = (key_map(keys[i]) - kmin) & mask;
    if (descending)
        std::stable_sort(perm.begin(), perm.end(),
                         [&](size_t a, size_t b) { return diff[a] > diff[b]; });
    else
        std::stable_sort(perm.begin(), perm.end(),
                         [&](size_t a, size_t b) { return diff[a] < diff[b]; });
}

template <class T>
void apply_perm(T* data, const std::vector<size_t>& perm)
{
    std::vector<T> tmp(data, data + perm.size());
    for (size_t i = 0; i < perm.size(); i++) data[i] = tmp[perm[i]];
}

void apply_perm_bytes(void* data, size_t elem, const std::vector<size_t>& perm)
{
    char* p = (char*)data;
    std::vector<char> tmp(p, p + elem * perm.size());
    for (size_t i = 0; i < perm.size(); i++) memcpy(p + i * elem, tmp.data() + perm[i] * elem, elem);
}

// ---------------------------------------------------------------------------
// Binary ops (built_in.h:14-150).  Integer Plus/Multiplies wrap modulo 2^w
// (computed in the unsigned type: the device code wraps, std:: on signed
// overflow is undefined).  Maximum is `x < y ? y : x`, Minimum `x > y ? y : x`.
// ---------------------------------------------------------------------------
template <class T> struct Unsigned { typedef T type; };
template <> struct Unsigned<int8_t> { typedef uint8_t type; };
template <> struct Unsigned<int16_t> { typedef uint16_t type; };
template <> struct Unsigned<int32_t> { typedef uint32_t type; };
template <> struct Unsigned<int64_t> { typedef uint64_t type; };

template <class T> inline T apply_op(int op, T x, T y)
{
    typedef typename Unsigned<T>::type U;
    switch (op) {
    case 0:
        if (std::is_integral<T>::value) return (T)(U)((U)x + (U)y);
        return (T)(x + y);
    case 1: return x < y ? y : x;
    case 2: return x > y ? y : x;
    case 3:
        if (std::is_integral<T>::value) return (T)(U)((U)x * (U)y);
        return (T)(x * y);
    }
    return x;
}
template <> inline float apply_op<float>(int op, float x, float y)
{
    switch (op) {
    case 0: return x + y;
    case 1: return x < y ? y : x;
    case 2: return x > y ? y : x;
    case 3: return x * y;
    }
    return x;
}
template <> inline double apply_op<double>(int op, double x, double y)
{
    switch (op) {
    case 0: return x + y;
    case 1: return x < y ? y : x;
    case 2: return x > y ? y : x;
    case 3: return x * y;
    }
    return x;
}

// Inclusive scan: scan.cpp:4-16 + general_scan.cpp:66-88; out[i] = in[0] op ...
// op in[i], the left operand always the earlier partial.
template <class T> void incl_scan(const T* in, T* out, size_t n, int op)
{
    if (n == 0) return;
    T acc = in[0];
    out[0] = acc;
    for (size_t i = 1; i < n; i++) {
        acc = apply_op<T>(op, acc, in[i]);
        out[i] = acc;
    }
}

// Exclusive scan: scan.cpp:18-39 — an inclusive scan of
// src(i) = i > 0 ? in[i-1] : init, so out[0] = init, out[i] = init op in[0..i-1].
template <class T> void excl_scan(const T* in, T* out, size_t n, T init, int op)
{
    if (n == 0) return;
    T acc = init;
    T prev = in[0];
    out[0] = acc;
    for (size_t i = 1; i < n; i++) {
        acc = apply_op<T>(op, acc, prev);
        prev = in[i];
        out[i] = acc;
    }
}

// Scan by key: scan_by_key.cpp:4-53 with general_scan_by_key.cpp:36-55.
// A segment starts at i == 0 or where !(key[i-1] == key[i]) (EqualTo).
// Inclusive: out[i] = head ? v[i] : out[i-1] op v[i].
// Exclusive: the inclusive segmented scan of src(i) = head ? init : v[i-1].
template <class K, class V>
void scan_by_key(const K* key, const V* val, V* out, size_t n, int exclusive, V init, int op)
{
    V acc = V();
    V prev = V();
    for (size_t i = 0; i < n; i++) {
        bool head = (i == 0) || !(key[i - 1] == key[i]);
        V src = exclusive ? (head ? init : prev) : val[i];
        prev = val[i];
        acc = head ? src : apply_op<V>(op, acc, src);
        out[i] = acc;
    }
}

// Reduce: reduce.cpp:5-31.  Without init the result starts from element 0;
// with init, init is element 0 of the reduced sequence (src(0) = init).
template <class T> T reduce_seq(const T* in, size_t n, int has_init, T init, int op)
{
    T acc;
    size_t i = 0;
    if (has_init) acc = init;
    else {
        if (n == 0) return T();
        acc = in[0];
        i = 1;
    }
    for (; i < n; i++) acc = apply_op<T>(op, acc, in[i]);
    return acc;
}

}  // namespace

#define DISPATCH_TYPE(tc, MACRO) \
    switch (tc) { \
    case 0: MACRO(int8_t); break; \
    case 1: MACRO(uint8_t); break; \
    case 2: MACRO(int16_t); break; \
    case 3: MACRO(uint16_t); break; \
    case 4: MACRO(int32_t); break; \
    case 5: MACRO(uint32_t); break; \
    case 6: MACRO(int64_t); break; \
    case 7: MACRO(uint64_t); break; \
    case 8: MACRO(float); break; \
    case 9: MACRO(double); break; \
    default: return -1; \
    }

namespace {

// ---------------------------------------------------------------------------
// Scan / compaction consumers.  Key equivalence of the restatements below:
// a == b (EqualTo, built_in.h), or a / div == b / div for integer keys with
// div > 1 (an equivalence relation, like every binary_pred the tests use).
// Predicates on one value v (pmode): 0: v % a == b (C++ truncating %),
// 1: v == a, 2: v != 0 (Identity on an integer stencil).
// ---------------------------------------------------------------------------
template <class K> inline bool key_eq(K x, K y, int64_t div)
{
    if constexpr (std::is_integral<K>::value) {
        if (div > 1) return (int64_t)x / div == (int64_t)y / div;
    }
    return x == y;
}

inline bool pred_of(int64_t v, int pmode, int64_t a, int64_t b)
{
    if (pmode == 0) return v % a == b;
    if (pmode == 1) return v == a;
    return v != 0;
}

// unique.cpp:5-95: item i is kept when i == 0 or !pred(key[i], key[i-1]);
// values (vsize bytes each) follow their keys.
template <class K>
uint64_t unique_keys(const K* keys, K* kout, const char* vals, char* vout, uint64_t vsize, uint64_t n, int64_t div)
{
    uint64_t c = 0;
    for (uint64_t i = 0; i < n; i++) {
        if (i > 0 && key_eq(keys[i], keys[i - 1], div)) continue;
        kout[c] = keys[i];
        if (vals) memcpy(vout + c * vsize, vals + i * vsize, vsize);
        c++;
    }
    return c;
}

// reduce.cpp:36-62: Inclusive_Scan_By_Key of the values (segments start where
// !pred(key[i-1], key[i])), then the key and scanned value at every segment
// tail (i == n-1 or !pred(key[i], key[i+1])).
template <class K, class V>
uint64_t reduce_by_key(const K* keys, const V* vals, K* kout, V* vout, uint64_t n, int op, int64_t div)
{
    uint64_t c = 0;
    V run = V();
    for (uint64_t i = 0; i < n; i++) {
        bool head = i == 0 || !key_eq(keys[i - 1], keys[i], div);
        run = head ? vals[i] : apply_op<V>(op, run, vals[i]);
        bool tail = i == n - 1 || !key_eq(keys[i], keys[i + 1], div);
        if (tail) {
            kout[c] = keys[i];
            vout[c] = run;
            c++;
        }
    }
    return c;
}

template <class T> inline T unary_of(int u, T x)
{
    if (u == 1) return (T)(T(0) - x);  // Negate (wraps for unsigned)
    if (u == 2) return (T)(x * x);
    return x;
}

}  // namespace

extern "C" {

// TRTC_Sort (sort.cpp:7-35, radix_sort.cpp:165-237): in-place stable sort.
int or_sort(int tc, void* keys, uint64_t n, int descending)
{
    std::vector<size_t> perm;
#define DO(T) sort_perm<T>((const T*)keys, n, descending, perm); apply_perm<T>((T*)keys, perm);
    DISPATCH_TYPE(tc, DO)
#undef DO
    return 0;
}

// TRTC_Sort_By_Key (sort.cpp:37-65, radix_sort.cpp:315-433): the first n values
// follow their keys; value_size is the value element size in bytes.
int or_sort_by_key(int tc, void* keys, void* values, uint64_t value_size, uint64_t n, int descending)
{
    std::vector<size_t> perm;
#define DO(T) sort_perm<T>((const T*)keys, n, descending, perm); apply_perm<T>((T*)keys, perm);
    DISPATCH_TYPE(tc, DO)
#undef DO
    apply_perm_bytes(values, value_size, perm);
    return 0;
}

// The permutation itself (perm[i] = input index of output element i).
int or_sort_perm(int tc, const void* keys, uint64_t n, int descending, uint64_t* perm_out)
{
    std::vector<size_t> perm;
#define DO(T) sort_perm<T>((const T*)keys, n, descending, perm);
    DISPATCH_TYPE(tc, DO)
#undef DO
    for (size_t i = 0; i < n; i++) perm_out[i] = perm[i];
    return 0;
}

int or_inclusive_scan(int tc, const void* in, void* out, uint64_t n, int op)
{
#define DO(T) incl_scan<T>((const T*)in, (T*)out, n, op);
    DISPATCH_TYPE(tc, DO)
#undef DO
    return 0;
}

// init points at one element of type tc.
int or_exclusive_scan(int tc, const void* in, void* out, uint64_t n, const void* init, int op)
{
#define DO(T) excl_scan<T>((const T*)in, (T*)out, n, *(const T*)init, op);
    DISPATCH_TYPE(tc, DO)
#undef DO
    return 0;
}

// Keys of type ktc, values of type vtc; init points at one value.
int or_scan_by_key(int ktc, int vtc, const void* keys, const void* values, void* out,
                   uint64_t n, int exclusive, const void* init, int op)
{
#define DOV(V) \
    switch (ktc) { \
    case 0: scan_by_key<int8_t, V>((const int8_t*)keys, (const V*)values, (V*)out, n, exclusive, *(const V*)init, op); break; \
    case 1: scan_by_key<uint8_t, V>((const uint8_t*)keys, (const V*)values, (V*)out, n, exclusive, *(const V*)init, op); break; \
    case 2: scan_by_key<int16_t, V>((const int16_t*)keys, (const V*)values, (V*)out, n, exclusive, *(const V*)init, op); break; \
    case 3: scan_by_key<uint16_t, V>((const uint16_t*)keys, (const V*)values, (V*)out, n, exclusive, *(const V*)init, op); break; \
    case 4: scan_by_key<int32_t, V>((const int32_t*)keys, (const V*)values, (V*)out, n, exclusive, *(const V*)init, op); break; \
    case 5: scan_by_key<uint32_t, V>((const uint32_t*)keys, (const V*)values, (V*)out, n, exclusive, *(const V*)init, op); break; \
    case 6: scan_by_key<int64_t, V>((const int64_t*)keys, (const V*)values, (V*)out, n, exclusive, *(const V*)init, op); break; \
    case 7: scan_by_key<uint64_t, V>((const uint64_t*)keys, (const V*)values, (V*)out, n, exclusive, *(const V*)init, op); break; \
    case 8: scan_by_key<float, V>((const float*)keys, (const V*)values, (V*)out, n, exclusive, *(const V*)init, op); break; \
    case 9: scan_by_key<double, V>((const double*)keys, (const V*)values, (V*)out, n, exclusive, *(const V*)init, op); break; \
    default: return -1; \
    }
    DISPATCH_TYPE(vtc, DOV)
#undef DOV
    return 0;
}

// Reduce in the element type, sequentially (exact for integer types).
int or_reduce(int tc, const void* in, uint64_t n, int has_init, const void* init, int op, void* result)
{
#define DO(T) { T iv = has_init ? *(const T*)init : T(); *(T*)result = reduce_seq<T>((const T*)in, n, has_init, iv, op); }
    DISPATCH_TYPE(tc, DO)
#undef DO
    return 0;
}

// FP64 sum of a float array: the tolerance reference for float Reduce.
double or_sum_f32_fp64(const float* in, uint64_t n)
{
    double s = 0.0;
    for (uint64_t i = 0; i < n; i++) s += (double)in[i];
    return s;
}

// The reference's exact float reduction order (general_reduce.cpp:30-95): blocks
// of 256 consecutive elements, each reduced by the halving tree
//   for s = 128..1: if (tid < s && i + s < n) buf[tid] = op(buf[tid], buf[tid+s])
// (i = block*256 + tid), the block results reduced again the same way until one
// value remains.  With init, the sequence is (init, v0, v1, ...) (reduce.cpp:18-22).
float or_reduce_ref_tree_f32(const float* in, uint64_t n, int has_init, float init)
{
    std::vector<float> cur;
    if (has_init) {
        cur.resize(n + 1);
        cur[0] = init;
        for (uint64_t i = 0; i < n; i++) cur[i + 1] = in[i];
    } else {
        cur.assign(in, in + n);
    }
    if (cur.empty()) return 0.0f;
    while (true) {
        size_t m = cur.size();
        size_t blocks = (m + 255) / 256;
        std::vector<float> next(blocks);
        for (size_t b = 0; b < blocks; b++) {
            float buf[256];
            for (size_t t = 0; t < 256; t++) {
                size_t i = b * 256 + t;
                buf[t] = i < m ? cur[i] : 0.0f;
            }
            for (size_t s = 128; s > 0; s >>= 1)
                for (size_t t = 0; t < s; t++) {
                    size_t i = b * 256 + t;
                    if (i + s < m) buf[t] = buf[t] + buf[t + s];
                }
            next[b] = buf[0];
        }
        if (blocks == 1) return next[0];
        cur.swap(next);
    }
}

// Copy_If (copy.cpp:16-21, general_copy_if.cpp:26-42): stable compaction with
// the predicate x % modulus == remainder (C++ truncating %, so negative
// multiples are kept too).  Returns the count.
uint64_t or_copy_if_mod_i32(const int32_t* in, int32_t* out, uint64_t n, int32_t modulus, int32_t remainder)
{
    uint64_t k = 0;
    for (uint64_t i = 0; i < n; i++)
        if (in[i] % modulus == remainder) out[k++] = in[i];
    return k;
}

// Copy_If_Stencil (copy.cpp:23-28): keeps in[i] where stencil[i] % m == r.
uint64_t or_copy_if_stencil_mod_i32(const int32_t* in, const int32_t* stencil, int32_t* out,
                                    uint64_t n, int32_t modulus, int32_t remainder)
{
    uint64_t k = 0;
    for (uint64_t i = 0; i < n; i++)
        if (stencil[i] % modulus == remainder) out[k++] = in[i];
    return k;
}

// Unique / Unique_By_Key (unique.cpp:5-95).  vals may be null.
uint64_t or_unique(int ktc, const void* keys, void* kout, const void* vals, void* vout, uint64_t vsize, uint64_t n,
                   int64_t div)
{
#define DO(K) return unique_keys<K>((const K*)keys, (K*)kout, (const char*)vals, (char*)vout, vsize, n, div);
    DISPATCH_TYPE(ktc, DO)
#undef DO
}

// Reduce_By_Key (reduce.cpp:36-62), keys ktc, values vtc (value_out type).
uint64_t or_reduce_by_key(int ktc, int vtc, const void* keys, const void* vals, void* kout, void* vout, uint64_t n,
                          int op, int64_t div)
{
#define DOV(V) \
    switch (ktc) { \
    case 4: return reduce_by_key<int32_t, V>((const int32_t*)keys, (const V*)vals, (int32_t*)kout, (V*)vout, n, op, div); \
    case 5: return reduce_by_key<uint32_t, V>((const uint32_t*)keys, (const V*)vals, (uint32_t*)kout, (V*)vout, n, op, div); \
    case 6: return reduce_by_key<int64_t, V>((const int64_t*)keys, (const V*)vals, (int64_t*)kout, (V*)vout, n, op, div); \
    case 8: return reduce_by_key<float, V>((const float*)keys, (const V*)vals, (float*)kout, (V*)vout, n, op, div); \
    default: return ~0ull; \
    }
    DISPATCH_TYPE(vtc, DOV)
#undef DOV
}

// Copy_If / Remove family (copy.cpp:16-28, remove.cpp:5-60): stable
// compaction of in (type tc) keeping items whose predicate (on the int32
// stencil when given, else on the item) holds, or fails when negate is set
// (Remove keeps x != value: pmode 1 with negate).
uint64_t or_select(int tc, const void* in, const int32_t* stencil, void* out, uint64_t n, int pmode, int64_t a,
                   int64_t b, int negate)
{
#define DO(T) { \
        uint64_t c = 0; \
        for (uint64_t i = 0; i < n; i++) { \
            int64_t v = stencil ? (int64_t)stencil[i] : (int64_t)((const T*)in)[i]; \
            if (pred_of(v, pmode, a, b) != (negate != 0)) ((T*)out)[c++] = ((const T*)in)[i]; \
        } \
        return c; }
    DISPATCH_TYPE(tc, DO)
#undef DO
}

// Partition family (partition.cpp:28-130): items whose predicate holds, in
// order, to tout; the others, in order, to fout.  Returns the first count.
// The in-place forms are tout = the vector, fout appended after the count.
uint64_t or_partition(int tc, const void* in, const int32_t* stencil, void* tout, void* fout, uint64_t n, int pmode,
                      int64_t a, int64_t b)
{
#define DO(T) { \
        uint64_t t = 0, f = 0; \
        for (uint64_t i = 0; i < n; i++) { \
            int64_t v = stencil ? (int64_t)stencil[i] : (int64_t)((const T*)in)[i]; \
            if (pred_of(v, pmode, a, b)) ((T*)tout)[t++] = ((const T*)in)[i]; \
            else ((T*)fout)[f++] = ((const T*)in)[i]; \
        } \
        return t; }
    DISPATCH_TYPE(tc, DO)
#undef DO
}

// Transform_Inclusive/Exclusive_Scan (transform_scan.cpp:4-16): the scan of
// unary(in[i]) (0 Identity, 1 Negate, 2 square) in the element type;
// exclusive: init first.
int or_transform_scan(int tc, const void* in, void* out, uint64_t n, int unary, int op, int exclusive,
                      const void* init)
{
#define DO(T) { \
        const T* x = (const T*)in; \
        T* y = (T*)out; \
        std::vector<T> u(n); \
        for (uint64_t i = 0; i < n; i++) u[i] = unary_of<T>(unary, x[i]); \
        if (exclusive) excl_scan<T>(u.data(), y, n, *(const T*)init, op); \
        else incl_scan<T>(u.data(), y, n, op); \
        return 0; }
    DISPATCH_TYPE(tc, DO)
#undef DO
}

// Count / Count_If (count.cpp:4-26): number of items whose predicate holds.
uint64_t or_count(int tc, const void* in, uint64_t n, int pmode, int64_t a, int64_t b)
{
#define DO(T) { \
        uint64_t c = 0; \
        for (uint64_t i = 0; i < n; i++) c += pred_of((int64_t)((const T*)in)[i], pmode, a, b); \
        return c; }
    DISPATCH_TYPE(tc, DO)
#undef DO
}

// ---------------------------------------------------------------------------
// Host baseline (reported beside the GPU numbers, never the product): the same
// calls with __gnu_parallel / OpenMP on every host core.
// ---------------------------------------------------------------------------
int bl_threads(void) { return omp_get_max_threads(); }

struct KV { uint32_t k, v; };

// Stable sort of (key, value) pairs by uint32 key, in place on AoS pairs.
void bl_sort_pairs_u32(uint32_t* keys, uint32_t* values, uint64_t n)
{
    std::vector<KV> kv(n);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)n; i++) kv[i] = KV{keys[i], values[i]};
    __gnu_parallel::stable_sort(kv.begin(), kv.end(),
                                [](const KV& a, const KV& b) { return a.k < b.k; });
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)n; i++) { keys[i] = kv[i].k; values[i] = kv[i].v; }
}

void bl_sort_u32(uint32_t* keys, uint64_t n)
{
    __gnu_parallel::stable_sort(keys, keys + n);
}

void bl_inclusive_scan_i32(const int32_t* in, int32_t* out, uint64_t n)
{
    __gnu_parallel::partial_sum((const uint32_t*)in, (const uint32_t*)in + n, (uint32_t*)out);
}

float bl_reduce_f32(const float* in, uint64_t n)
{
    return __gnu_parallel::accumulate(in, in + n, 0.0f);
}

uint64_t bl_copy_if_mod3_i32(const int32_t* in, int32_t* out, uint64_t n)
{
    int nt = omp_get_max_threads();
    std::vector<uint64_t> cnt(nt + 1, 0);
#pragma omp parallel num_threads(nt)
    {
        int t = omp_get_thread_num();
        uint64_t b = n * t / nt, e = n * (t + 1) / nt, c = 0;
        for (uint64_t i = b; i < e; i++) c += (in[i] % 3 == 0);
        cnt[t + 1] = c;
#pragma omp barrier
#pragma omp single
        for (int i = 0; i < nt; i++) cnt[i + 1] += cnt[i];
        uint64_t k = cnt[t];
        for (uint64_t i = b; i < e; i++)
            if (in[i] % 3 == 0) out[k++] = in[i];
    }
    return cnt[nt];
}

// Batched segments on the host: OpenMP over segments (dynamic schedule, the
// sizes are uneven), std::stable_sort of (key, value) pairs and
// std::inclusive_scan (uint32 wrap) per segment.
void bl_segmented_sort_pairs_u32(uint32_t* keys, uint32_t* values, const uint64_t* offsets, uint64_t num_segments)
{
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t s = 0; s < (int64_t)num_segments; s++) {
        uint64_t b = offsets[s], e = offsets[s + 1];
        std::vector<KV> kv(e - b);
        for (uint64_t i = b; i < e; i++) kv[i - b] = KV{keys[i], values[i]};
        std::stable_sort(kv.begin(), kv.end(), [](const KV& x, const KV& y) { return x.k < y.k; });
        for (uint64_t i = b; i < e; i++) {
            keys[i] = kv[i - b].k;
            values[i] = kv[i - b].v;
        }
    }
}

void bl_segmented_inclusive_scan_i32(const int32_t* in, int32_t* out, const uint64_t* offsets, uint64_t num_segments)
{
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t s = 0; s < (int64_t)num_segments; s++) {
        uint64_t b = offsets[s], e = offsets[s + 1];
        std::inclusive_scan((const uint32_t*)in + b, (const uint32_t*)in + e, (uint32_t*)out + b);
    }
}

}  // extern "C"
