// The remaining Thrust-like algorithms of the reference's API (transform.cpp,
// replace.cpp, for_each.cpp, adjacent_difference.cpp, sequence.cpp,
// tabulate.cpp, gather.cpp, scatter.cpp, swap.cpp, fill.cpp, equal.cpp,
// extrema.cpp, inner_product.cpp, logical.cpp, find.cpp, mismatch.cpp,
// binary_search.cpp, merge.cpp, sort.cpp:85-101, partition.cpp:132-210,
// DVTuple.cpp).
//
// * Element-wise algorithms are one JIT `For` kernel each (one thread per
//   element, hipRTC-compiled once per combination of view types).
// * Searches for a first index (find, mismatch, equal, all/any/none_of,
//   is_sorted_until, is_partitioned) are the single-launch reduction of the
//   primitive core with Minimum over a computed index vector (i where the
//   condition holds, else the largest size_t) -- no atomics, one pass.
// * Extrema reduce the index vector 0..n-1 with an "earlier wins ties"
//   comparison functor (the reduction combines in index order); inner_product
//   reduces a computed vector op2(a[i], b[i]) with init and op1.
// * Single-value binary searches run one thread; the _V forms one thread per
//   value (d_lower_bound / d_upper_bound / d_binary_search).
// * Merge / Merge_By_Key: the merge-path kernel of trtc_merge.hpp.
#include <algorithm>
#include <deque>

#include "DVTuple.h"
#include "adjacent_difference.h"
#include "binary_search.h"
#include "equal.h"
#include "extrema.h"
#include "fake_vectors/DVCounter.h"
#include "fake_vectors/DVCustomVector.h"
#include "fake_vectors/DVZipped.h"
#include "fill.h"
#include "find.h"
#include "for_each.h"
#include "gather.h"
#include "inner_product.h"
#include "logical.h"
#include "merge.h"
#include "mismatch.h"
#include "partition.h"
#include "replace.h"
#include "scatter.h"
#include "sequence.h"
#include "sort.h"
#include "swap.h"
#include "tabulate.h"
#include "transform.h"
#include "trtc_internal.h"
#include "../device/trtc_args.h"

using namespace trtc_host;

namespace {

typedef std::initializer_list<const DeviceViewable*> DVList;

// Vectors read or written at every index < n must hold n items (the
// reference does not check; a short vector would be accessed out of bounds).
bool need(const DVVectorLike& v, size_t n, const char* fn, const char* what)
{
    if (v.size() >= n) return true;
    TRTC_ERR("%s: %s has %zu items, %zu needed", fn, what, v.size(), n);
    return false;
}

bool run_for(TRTC_For& f, size_t n, DVList args)
{
    std::vector<const DeviceViewable*> a(args);
    return f.launch_n(n, a.data());
}

// Index of the first i < n where `cond` holds (cond names the captured
// objects and `idx`), (size_t)-1 when none.
bool first_index(const std::vector<CapturedDeviceViewable>& caps, const char* cond, size_t n, size_t& result)
{
    result = (size_t)-1;
    if (n == 0) return true;
    std::string body = std::string("        return (") + cond + ") ? (uint64_t)idx : ~(uint64_t)0;\n";
    DVCustomVector iv(caps, "idx", body.c_str(), "uint64_t", n, true);
    ViewBuf buf;
    if (!reduce_view(iv, nullptr, Functor("Minimum"), buf)) return false;
    uint64_t r;
    memcpy(&r, buf.data(), 8);
    result = (size_t)r;
    return true;
}

// One-element uint64 device result: run `f` (n threads), read element 0.
template <class T> bool device_result(TRTC_For& f, size_t n, std::vector<const DeviceViewable*> args, T& out)
{
    uint64_t zero = 0;
    DVVector res("uint64_t", 1, &zero);
    args.push_back(&res);
    if (!f.launch_n(n, args.data())) return false;
    uint64_t r = 0;
    res.to_host(&r);
    out = (T)r;
    return true;
}

size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

}  // namespace

namespace trtc_host {
// Element-wise map over [0, n) for bodies that read at index idx and then
// write at idx: `value` is an expression in `idx` and the named views, `store`
// a statement that writes `_x` (the value) at idx.  Each thread takes 8
// indices 256 apart (coalesced), evaluates all 8 values, then stores them, so
// 8 reads per thread are in flight at once -- a one-index-per-thread For keeps
// one (2^29 int32 Transform: 0.58 ms as a For).  A thread reads all of its
// indices before it writes any, and no other thread touches them, so in-place
// calls (an output that is also an input) are safe.
bool run_map(const std::vector<const char*>& names, const char* value, const char* store, size_t n, DVList args)
{
    if (!ensure_init()) return false;
    if (n == 0) return true;
    std::vector<const DeviceViewable*> a(args);
    std::string sig;
    for (size_t i = 0; i < names.size(); i++) sig += std::string(", ") + a[i]->name_view_cls() + " " + names[i];
    KernelRef k;
    const std::string memo = std::string("map\x01") + sig + "\x01" + value + "\x01" + store;
    auto build = [&](KernelRef& out) {
        std::string src = std::string("extern \"C\" __global__ void __launch_bounds__(256) trtc_map(size_t _trtc_n") +
                          sig +
                          ")\n{\n"
                          "    auto _trtc_val = [&](size_t idx) { return " + value + "; };\n"
                          "    typedef decltype(_trtc_val((size_t)0)) _trtc_T;\n"
                          "    auto _trtc_put = [&](size_t idx, const _trtc_T& _x) { " + store + " };\n"
                          "    const size_t _trtc_stride = (size_t)gridDim.x * 2048;\n"
                          "    for (size_t _trtc_b = (size_t)blockIdx.x * 2048 + threadIdx.x; _trtc_b < _trtc_n;\n"
                          "         _trtc_b += _trtc_stride) {\n"
                          "        if (_trtc_b + 7 * 256 < _trtc_n) {\n"
                          "            _trtc_T _r[8];\n"
                          "#pragma unroll\n"
                          "            for (int _k = 0; _k < 8; _k++) _r[_k] = _trtc_val(_trtc_b + (size_t)_k * 256);\n"
                          "#pragma unroll\n"
                          "            for (int _k = 0; _k < 8; _k++) _trtc_put(_trtc_b + (size_t)_k * 256, _r[_k]);\n"
                          "        } else {\n"
                          "            for (int _k = 0; _k < 8; _k++) {\n"
                          "                const size_t idx = _trtc_b + (size_t)_k * 256;\n"
                          "                if (idx < _trtc_n) _trtc_put(idx, _trtc_val(idx));\n"
                          "            }\n"
                          "        }\n"
                          "    }\n"
                          "}\n";
        return get_source_kernel(src, "trtc_map", out);
    };
    if (!memo_kernel(memo, build, k)) return false;
    std::vector<ViewBuf> views(names.size());
    std::vector<void*> ptrs(names.size() + 1);
    ptrs[0] = &n;
    for (size_t i = 0; i < names.size(); i++) {
        views[i] = a[i]->view();
        ptrs[i + 1] = views[i].data();
    }
    // one 2048-index chunk per workgroup up to 2^31 indices, grid-stride beyond
    return launch_kernel(k, dim3(capped_grid(n, 2048, (size_t)1 << 20)), dim3(256), ptrs.data(), 0, cur_stream());
}
}  // namespace trtc_host

// ================================================================ element-wise
bool TRTC_Fill(DVVectorLike& vec, const DeviceViewable& value)
{
    return run_map({"view_vec", "value"}, "(decltype(view_vec)::value_t)value", "view_vec[idx] = _x;", vec.size(),
                   {&vec, &value});
}

bool TRTC_Transform(const DVVectorLike& vec_in, DVVectorLike& vec_out, const Functor& op)
{
    if (!need(vec_out, vec_in.size(), "Transform", "vec_out")) return false;
    return run_map({"vin", "vout", "op"}, "(decltype(vout)::value_t)op(vin[idx])", "vout[idx] = _x;", vec_in.size(),
                   {&vec_in, &vec_out, &op});
}

bool TRTC_Transform_Binary(const DVVectorLike& vec_in1, const DVVectorLike& vec_in2, DVVectorLike& vec_out,
                           const Functor& op)
{
    if (!need(vec_in2, vec_in1.size(), "Transform_Binary", "vec_in2") ||
        !need(vec_out, vec_in1.size(), "Transform_Binary", "vec_out"))
        return false;
    return run_map({"vin1", "vin2", "vout", "op"}, "(decltype(vout)::value_t)op(vin1[idx], vin2[idx])",
                   "vout[idx] = _x;", vec_in1.size(), {&vec_in1, &vec_in2, &vec_out, &op});
}

bool TRTC_Transform_If(const DVVectorLike& vec_in, DVVectorLike& vec_out, const Functor& op, const Functor& pred)
{
    static TRTC_For f({"vin", "vout", "op", "pred"}, "idx",
                      "    if (pred(vin[idx])) vout[idx] = (decltype(vout)::value_t)op(vin[idx]);\n");
    if (!need(vec_out, vec_in.size(), "Transform_If", "vec_out")) return false;
    return run_for(f, vec_in.size(), {&vec_in, &vec_out, &op, &pred});
}

bool TRTC_Transform_If_Stencil(const DVVectorLike& vec_in, const DVVectorLike& vec_stencil, DVVectorLike& vec_out,
                               const Functor& op, const Functor& pred)
{
    static TRTC_For f({"vin", "vst", "vout", "op", "pred"}, "idx",
                      "    if (pred(vst[idx])) vout[idx] = (decltype(vout)::value_t)op(vin[idx]);\n");
    if (!need(vec_stencil, vec_in.size(), "Transform_If_Stencil", "vec_stencil") ||
        !need(vec_out, vec_in.size(), "Transform_If_Stencil", "vec_out"))
        return false;
    return run_for(f, vec_in.size(), {&vec_in, &vec_stencil, &vec_out, &op, &pred});
}

bool TRTC_Transform_Binary_If_Stencil(const DVVectorLike& vec_in1, const DVVectorLike& vec_in2,
                                      const DVVectorLike& vec_stencil, DVVectorLike& vec_out, const Functor& op,
                                      const Functor& pred)
{
    static TRTC_For f({"vin1", "vin2", "vst", "vout", "op", "pred"}, "idx",
                      "    if (pred(vst[idx])) vout[idx] = (decltype(vout)::value_t)op(vin1[idx], vin2[idx]);\n");
    const size_t n = vec_in1.size();
    if (!need(vec_in2, n, "Transform_Binary_If_Stencil", "vec_in2") ||
        !need(vec_stencil, n, "Transform_Binary_If_Stencil", "vec_stencil") ||
        !need(vec_out, n, "Transform_Binary_If_Stencil", "vec_out"))
        return false;
    return run_for(f, vec_in1.size(), {&vec_in1, &vec_in2, &vec_stencil, &vec_out, &op, &pred});
}

bool TRTC_Replace(DVVectorLike& vec, const DeviceViewable& old_value, const DeviceViewable& new_value)
{
    static TRTC_For f({"v", "old_value", "new_value"}, "idx",
                      "    if (v[idx] == (decltype(v)::value_t)old_value) v[idx] = (decltype(v)::value_t)new_value;\n");
    return run_for(f, vec.size(), {&vec, &old_value, &new_value});
}

bool TRTC_Replace_If(DVVectorLike& vec, const Functor& pred, const DeviceViewable& new_value)
{
    static TRTC_For f({"v", "pred", "new_value"}, "idx",
                      "    if (pred(v[idx])) v[idx] = (decltype(v)::value_t)new_value;\n");
    return run_for(f, vec.size(), {&vec, &pred, &new_value});
}

bool TRTC_Replace_Copy(const DVVectorLike& vec_in, DVVectorLike& vec_out, const DeviceViewable& old_value,
                       const DeviceViewable& new_value)
{
    if (!need(vec_out, vec_in.size(), "Replace_Copy", "vec_out")) return false;
    // the element is read once (replace.cpp:27-28), as computed views
    // (transform, custom, zipped) evaluate their operator on every read
    return run_map({"vin", "vout", "old_value", "new_value"},
                   "[&] { auto value = vin[idx];\n"
                   "      return value == (decltype(vin)::value_t)old_value ? (decltype(vout)::value_t)new_value\n"
                   "                                                        : (decltype(vout)::value_t)value; }()",
                   "vout[idx] = _x;", vec_in.size(), {&vec_in, &vec_out, &old_value, &new_value});
}

bool TRTC_Replace_Copy_If(const DVVectorLike& vec_in, DVVectorLike& vec_out, const Functor& pred,
                          const DeviceViewable& new_value)
{
    if (!need(vec_out, vec_in.size(), "Replace_Copy_If", "vec_out")) return false;
    // one read of the element (replace.cpp:39-40)
    return run_map({"vin", "vout", "pred", "new_value"},
                   "[&] { auto value = vin[idx];\n"
                   "      return pred(value) ? (decltype(vout)::value_t)new_value : (decltype(vout)::value_t)value; }()",
                   "vout[idx] = _x;", vec_in.size(), {&vec_in, &vec_out, &pred, &new_value});
}

bool TRTC_For_Each(DVVectorLike& vec, const Functor& f)
{
    static TRTC_For k({"v", "f"}, "idx", "    f(v[idx]);\n");
    return run_for(k, vec.size(), {&vec, &f});
}

// out[0] = in[0]; out[i] = op(in[i], in[i-1]) in the input type.  An
// in-place call reads a copy of the input, so no thread sees an overwritten
// neighbour.
bool TRTC_Adjacent_Difference(const DVVectorLike& vec_in, DVVectorLike& vec_out, const Functor& binary_op)
{
    if (!ensure_init()) return false;
    size_t n = vec_in.size();
    if (n == 0) return true;
    if (!need(vec_out, n, "Adjacent_Difference", "vec_out")) return false;
    // the element-wise map: out[idx] from in[idx] and in[idx - 1]
    auto run = [&](const DVVectorLike& src) {
        return run_map({"vin", "vout", "op"},
                       "(decltype(vout)::value_t)(idx > 0 ? (decltype(vin)::value_t)op(vin[idx], vin[idx - 1])\n"
                       "                                    : vin[idx])",
                       "vout[idx] = _x;", n, {&src, &vec_out, &binary_op});
    };
    void* pi = contiguous_data(vec_in);
    if (pi && pi == contiguous_data(vec_out)) {
        hipStream_t s = cur_stream();
        PoolBuf tmp(n * vec_in.elem_size(), s);
        if (!tmp.p) return false;
        DVVectorAdaptor cpy(vec_in.name_elem_cls().c_str(), n, tmp.p);
        if (!check_hip(hipMemcpyAsync(tmp.p, pi, n * vec_in.elem_size(), hipMemcpyDeviceToDevice, s), "hipMemcpyAsync"))
            return false;
        if (!run(cpy)) return false;
        return check_hip(hipStreamSynchronize(s), "hipStreamSynchronize");
    }
    return run(vec_in);
}

bool TRTC_Adjacent_Difference(const DVVectorLike& vec_in, DVVectorLike& vec_out)
{
    return TRTC_Adjacent_Difference(vec_in, vec_out, Functor("Minus"));
}

bool TRTC_Sequence(DVVectorLike& vec)
{
    return run_map({"v"}, "(decltype(v)::value_t)idx", "v[idx] = _x;", vec.size(), {&vec});
}

bool TRTC_Sequence(DVVectorLike& vec, const DeviceViewable& value_init)
{
    return run_map({"v", "value_init"}, "(decltype(v)::value_t)((decltype(v)::value_t)value_init + (decltype(v)::value_t)idx)",
                   "v[idx] = _x;", vec.size(), {&vec, &value_init});
}

bool TRTC_Sequence(DVVectorLike& vec, const DeviceViewable& value_init, const DeviceViewable& value_step)
{
    return run_map({"v", "value_init", "value_step"},
                   "(decltype(v)::value_t)((decltype(v)::value_t)value_init +\n"
                   "             (decltype(v)::value_t)idx * (decltype(v)::value_t)value_step)",
                   "v[idx] = _x;", vec.size(), {&vec, &value_init, &value_step});
}

// The reference's tabulate.cpp applies op to each element (vec[i] =
// op(vec[i])); for a vector holding 0..n-1 that is Thrust's op(i).
bool TRTC_Tabulate(DVVectorLike& vec, const Functor& op)
{
    return run_map({"v", "op"}, "(decltype(v)::value_t)op(v[idx])", "v[idx] = _x;", vec.size(), {&vec, &op});
}

bool TRTC_Gather(const DVVectorLike& vec_map, const DVVectorLike& vec_in, DVVectorLike& vec_out)
{
    // one index per thread: with 8 random reads in flight per thread (the
    // element-wise map) a 2^28 random gather ran 5.64 ms against 5.38 ms
    static TRTC_For f({"vmap", "vin", "vout"}, "idx", "    vout[idx] = (decltype(vout)::value_t)vin[vmap[idx]];\n");
    if (!need(vec_out, vec_map.size(), "Gather", "vec_out")) return false;
    return run_for(f, vec_map.size(), {&vec_map, &vec_in, &vec_out});
}

bool TRTC_Gather_If(const DVVectorLike& vec_map, const DVVectorLike& vec_stencil, const DVVectorLike& vec_in,
                    DVVectorLike& vec_out)
{
    static TRTC_For f({"vmap", "vst", "vin", "vout"}, "idx",
                      "    if (vst[idx]) vout[idx] = (decltype(vout)::value_t)vin[vmap[idx]];\n");
    if (!need(vec_stencil, vec_map.size(), "Gather_If", "vec_stencil") ||
        !need(vec_out, vec_map.size(), "Gather_If", "vec_out"))
        return false;
    return run_for(f, vec_map.size(), {&vec_map, &vec_stencil, &vec_in, &vec_out});
}

bool TRTC_Gather_If(const DVVectorLike& vec_map, const DVVectorLike& vec_stencil, const DVVectorLike& vec_in,
                    DVVectorLike& vec_out, const Functor& pred)
{
    static TRTC_For f({"vmap", "vst", "vin", "vout", "pred"}, "idx",
                      "    if (pred(vst[idx])) vout[idx] = (decltype(vout)::value_t)vin[vmap[idx]];\n");
    if (!need(vec_stencil, vec_map.size(), "Gather_If", "vec_stencil") ||
        !need(vec_out, vec_map.size(), "Gather_If", "vec_out"))
        return false;
    return run_for(f, vec_map.size(), {&vec_map, &vec_stencil, &vec_in, &vec_out, &pred});
}

bool TRTC_Scatter(const DVVectorLike& vec_in, const DVVectorLike& vec_map, DVVectorLike& vec_out)
{
    static TRTC_For f({"vin", "vmap", "vout"}, "idx", "    vout[vmap[idx]] = (decltype(vout)::value_t)vin[idx];\n");
    if (!need(vec_map, vec_in.size(), "Scatter", "vec_map")) return false;
    return run_for(f, vec_in.size(), {&vec_in, &vec_map, &vec_out});
}

bool TRTC_Scatter_If(const DVVectorLike& vec_in, const DVVectorLike& vec_map, const DVVectorLike& vec_stencil,
                     DVVectorLike& vec_out)
{
    static TRTC_For f({"vin", "vmap", "vst", "vout"}, "idx",
                      "    if (vst[idx]) vout[vmap[idx]] = (decltype(vout)::value_t)vin[idx];\n");
    if (!need(vec_map, vec_in.size(), "Scatter_If", "vec_map") ||
        !need(vec_stencil, vec_in.size(), "Scatter_If", "vec_stencil"))
        return false;
    return run_for(f, vec_in.size(), {&vec_in, &vec_map, &vec_stencil, &vec_out});
}

bool TRTC_Scatter_If(const DVVectorLike& vec_in, const DVVectorLike& vec_map, const DVVectorLike& vec_stencil,
                     DVVectorLike& vec_out, const Functor& pred)
{
    static TRTC_For f({"vin", "vmap", "vst", "vout", "pred"}, "idx",
                      "    if (pred(vst[idx])) vout[vmap[idx]] = (decltype(vout)::value_t)vin[idx];\n");
    if (!need(vec_map, vec_in.size(), "Scatter_If", "vec_map") ||
        !need(vec_stencil, vec_in.size(), "Scatter_If", "vec_stencil"))
        return false;
    return run_for(f, vec_in.size(), {&vec_in, &vec_map, &vec_stencil, &vec_out, &pred});
}

bool TRTC_Swap(DVVectorLike& vec1, DVVectorLike& vec2)
{
    static TRTC_For f({"v1", "v2"}, "idx",
                      "    typename decltype(v1)::value_t t = v1[idx];\n"
                      "    v1[idx] = (decltype(v1)::value_t)v2[idx];\n"
                      "    v2[idx] = (decltype(v2)::value_t)t;\n");
    if (!need(vec2, vec1.size(), "Swap", "vec2")) return false;
    return run_for(f, vec1.size(), {&vec1, &vec2});
}

// ================================================================ first-index searches
bool TRTC_Find(const DVVectorLike& vec, const DeviceViewable& value, size_t& result)
{
    return first_index({{"vec", &vec}, {"value", &value}}, "vec[idx] == value", vec.size(), result);
}

bool TRTC_Find_If(const DVVectorLike& vec, const Functor& pred, size_t& result)
{
    return first_index({{"vec", &vec}, {"pred", &pred}}, "pred(vec[idx])", vec.size(), result);
}

bool TRTC_Find_If_Not(const DVVectorLike& vec, const Functor& pred, size_t& result)
{
    return first_index({{"vec", &vec}, {"pred", &pred}}, "!pred(vec[idx])", vec.size(), result);
}

bool TRTC_Mismatch(const DVVectorLike& vec1, const DVVectorLike& vec2, size_t& result)
{
    if (!need(vec2, vec1.size(), "Mismatch", "vec2")) return false;
    return first_index({{"vec1", &vec1}, {"vec2", &vec2}}, "!(vec1[idx] == vec2[idx])", vec1.size(), result);
}

bool TRTC_Mismatch(const DVVectorLike& vec1, const DVVectorLike& vec2, const Functor& pred, size_t& result)
{
    if (!need(vec2, vec1.size(), "Mismatch", "vec2")) return false;
    return first_index({{"vec1", &vec1}, {"vec2", &vec2}, {"pred", &pred}}, "!pred(vec1[idx], vec2[idx])",
                       vec1.size(), result);
}

bool TRTC_Equal(const DVVectorLike& vec1, const DVVectorLike& vec2, bool& ret)
{
    size_t r;
    if (!TRTC_Mismatch(vec1, vec2, r)) return false;
    ret = r == (size_t)-1;
    return true;
}

bool TRTC_Equal(const DVVectorLike& vec1, const DVVectorLike& vec2, const Functor& binary_pred, bool& ret)
{
    size_t r;
    if (!TRTC_Mismatch(vec1, vec2, binary_pred, r)) return false;
    ret = r == (size_t)-1;
    return true;
}

bool TRTC_All_Of(const DVVectorLike& vec, const Functor& pred, bool& ret)
{
    size_t r;
    if (!TRTC_Find_If_Not(vec, pred, r)) return false;
    ret = r == (size_t)-1;
    return true;
}

bool TRTC_Any_Of(const DVVectorLike& vec, const Functor& pred, bool& ret)
{
    size_t r;
    if (!TRTC_Find_If(vec, pred, r)) return false;
    ret = r != (size_t)-1;
    return true;
}

bool TRTC_None_Of(const DVVectorLike& vec, const Functor& pred, bool& ret)
{
    size_t r;
    if (!TRTC_Find_If(vec, pred, r)) return false;
    ret = r == (size_t)-1;
    return true;
}

bool TRTC_Is_Sorted_Until(const DVVectorLike& vec, const Functor& comp, size_t& result)
{
    size_t r;
    if (!first_index({{"vec", &vec}, {"comp", &comp}}, "idx > 0 && comp(vec[idx], vec[idx - 1])", vec.size(), r))
        return false;
    result = r == (size_t)-1 ? vec.size() : r;
    return true;
}

bool TRTC_Is_Sorted_Until(const DVVectorLike& vec, size_t& result)
{
    return TRTC_Is_Sorted_Until(vec, Functor("Less"), result);
}

// Partitioned: no item failing pred is followed by one satisfying it.
bool TRTC_Is_Partitioned(const DVVectorLike& vec, const Functor& pred, bool& result)
{
    size_t r;
    if (!first_index({{"vec", &vec}, {"pred", &pred}},
                     "idx + 1 < vec.size() && !pred(vec[idx]) && pred(vec[idx + 1])", vec.size(), r))
        return false;
    result = r == (size_t)-1;
    return true;
}

// ================================================================ extrema, inner product
namespace {
bool extreme_index(const DVVectorLike& vec, const Functor& comp, bool max, size_t& id)
{
    if (vec.size() == 0) {
        TRTC_ERR("Min/Max_Element: empty vector");
        return false;
    }
    bool handled = false;
    if (!extreme_view(vec, comp, max, id, handled)) return false;
    if (handled) return true;
    // Other vectors reduce (index, value) pairs: the running extreme stays in registers and
    // every element is read once (reducing the index counter with a functor
    // that looks both values up read two per combine: 2^28 int32 Min_Element
    // 1.05 ms).  The ordered reduction always combines (earlier, later), and
    // the later pair wins only when strictly better, so the earlier index
    // wins ties: the first smallest / the first largest.  The index is the
    // first member, at offset 0 of the result.
    uint64_t zero = 0;
    DVUInt64 first(zero);
    DVCounter idx(first, vec.size());
    DVZipped pairs({&idx, const_cast<DVVectorLike*>(&vec)}, {"index", "value"});
    Functor op({{"comp", &comp}}, {"a", "b"},
               max ? "        return comp(a.value, b.value) ? b : a;\n"
                   : "        return comp(b.value, a.value) ? b : a;\n");
    ViewBuf buf;
    if (!reduce_view(pairs, nullptr, op, buf)) return false;
    uint64_t r;
    memcpy(&r, buf.data(), 8);
    id = (size_t)r;
    return true;
}
}  // namespace

bool TRTC_Min_Element(const DVVectorLike& vec, const Functor& comp, size_t& id_min)
{
    return extreme_index(vec, comp, false, id_min);
}
bool TRTC_Min_Element(const DVVectorLike& vec, size_t& id_min) { return extreme_index(vec, Functor("Less"), false, id_min); }
bool TRTC_Max_Element(const DVVectorLike& vec, const Functor& comp, size_t& id_max)
{
    return extreme_index(vec, comp, true, id_max);
}
bool TRTC_Max_Element(const DVVectorLike& vec, size_t& id_max) { return extreme_index(vec, Functor("Less"), true, id_max); }
bool TRTC_MinMax_Element(const DVVectorLike& vec, const Functor& comp, size_t& id_min, size_t& id_max)
{
    return extreme_index(vec, comp, false, id_min) && extreme_index(vec, comp, true, id_max);
}
bool TRTC_MinMax_Element(const DVVectorLike& vec, size_t& id_min, size_t& id_max)
{
    return TRTC_MinMax_Element(vec, Functor("Less"), id_min, id_max);
}

bool TRTC_Inner_Product(const DVVectorLike& vec1, const DVVectorLike& vec2, const DeviceViewable& init, ViewBuf& ret,
                        const Functor& binary_op1, const Functor& binary_op2)
{
    if (!need(vec2, vec1.size(), "Inner_Product", "vec2")) return false;
    const std::string T = init.name_view_cls();
    std::string body = "        return (" + T + ")op2(vec1[idx], vec2[idx]);\n";
    DVCustomVector prod({{"vec1", &vec1}, {"vec2", &vec2}, {"op2", &binary_op2}}, "idx", body.c_str(), T.c_str(),
                        vec1.size(), true);
    return reduce_view(prod, &init, binary_op1, ret);
}

bool TRTC_Inner_Product(const DVVectorLike& vec1, const DVVectorLike& vec2, const DeviceViewable& init, ViewBuf& ret)
{
    return TRTC_Inner_Product(vec1, vec2, init, ret, Functor("Plus"), Functor("Multiplies"));
}

// ================================================================ binary searches
bool TRTC_Lower_Bound(const DVVectorLike& vec, const DeviceViewable& value, const Functor& comp, size_t& result)
{
    static TRTC_For f({"vec", "value", "comp", "res"}, "idx", "    res[0] = d_lower_bound(vec, value, comp);\n");
    return device_result(f, 1, {&vec, &value, &comp}, result);
}
bool TRTC_Lower_Bound(const DVVectorLike& vec, const DeviceViewable& value, size_t& result)
{
    return TRTC_Lower_Bound(vec, value, Functor("Less"), result);
}

bool TRTC_Upper_Bound(const DVVectorLike& vec, const DeviceViewable& value, const Functor& comp, size_t& result)
{
    static TRTC_For f({"vec", "value", "comp", "res"}, "idx", "    res[0] = d_upper_bound(vec, value, comp);\n");
    return device_result(f, 1, {&vec, &value, &comp}, result);
}
bool TRTC_Upper_Bound(const DVVectorLike& vec, const DeviceViewable& value, size_t& result)
{
    return TRTC_Upper_Bound(vec, value, Functor("Less"), result);
}

bool TRTC_Binary_Search(const DVVectorLike& vec, const DeviceViewable& value, const Functor& comp, bool& result)
{
    static TRTC_For f({"vec", "value", "comp", "res"}, "idx",
                      "    res[0] = d_binary_search(vec, value, comp) ? 1 : 0;\n");
    return device_result(f, 1, {&vec, &value, &comp}, result);
}
bool TRTC_Binary_Search(const DVVectorLike& vec, const DeviceViewable& value, bool& result)
{
    return TRTC_Binary_Search(vec, value, Functor("Less"), result);
}

bool TRTC_Lower_Bound_V(const DVVectorLike& vec, const DVVectorLike& values, DVVectorLike& result, const Functor& comp)
{
    static TRTC_For f({"vec", "values", "res", "comp"}, "idx",
                      "    res[idx] = (decltype(res)::value_t)d_lower_bound(vec, values[idx], comp);\n");
    if (!need(result, values.size(), "Lower_Bound_V", "result")) return false;
    return run_for(f, values.size(), {&vec, &values, &result, &comp});
}
bool TRTC_Lower_Bound_V(const DVVectorLike& vec, const DVVectorLike& values, DVVectorLike& result)
{
    return TRTC_Lower_Bound_V(vec, values, result, Functor("Less"));
}

bool TRTC_Upper_Bound_V(const DVVectorLike& vec, const DVVectorLike& values, DVVectorLike& result, const Functor& comp)
{
    static TRTC_For f({"vec", "values", "res", "comp"}, "idx",
                      "    res[idx] = (decltype(res)::value_t)d_upper_bound(vec, values[idx], comp);\n");
    if (!need(result, values.size(), "Upper_Bound_V", "result")) return false;
    return run_for(f, values.size(), {&vec, &values, &result, &comp});
}
bool TRTC_Upper_Bound_V(const DVVectorLike& vec, const DVVectorLike& values, DVVectorLike& result)
{
    return TRTC_Upper_Bound_V(vec, values, result, Functor("Less"));
}

bool TRTC_Binary_Search_V(const DVVectorLike& vec, const DVVectorLike& values, DVVectorLike& result,
                          const Functor& comp)
{
    static TRTC_For f({"vec", "values", "res", "comp"}, "idx",
                      "    res[idx] = (decltype(res)::value_t)d_binary_search(vec, values[idx], comp);\n");
    if (!need(result, values.size(), "Binary_Search_V", "result")) return false;
    return run_for(f, values.size(), {&vec, &values, &result, &comp});
}
bool TRTC_Binary_Search_V(const DVVectorLike& vec, const DVVectorLike& values, DVVectorLike& result)
{
    return TRTC_Binary_Search_V(vec, values, result, Functor("Less"));
}

// First item for which pred fails in a partitioned vector (std::partition_point).
bool TRTC_Partition_Point(const DVVectorLike& vec, const Functor& pred, size_t& result)
{
    static TRTC_For f({"vec", "pred", "res"}, "idx",
                      "    size_t lo = 0, hi = vec.size();\n"
                      "    while (lo < hi) {\n"
                      "        size_t mid = (lo + hi) >> 1;\n"
                      "        if (pred(vec[mid])) lo = mid + 1;\n"
                      "        else hi = mid;\n"
                      "    }\n"
                      "    res[0] = lo;\n");
    return device_result(f, 1, {&vec, &pred}, result);
}

// ================================================================ merge
namespace {
bool merge_views(const DVVectorLike& k1, const DVVectorLike& k2, const DVVectorLike* v1, const DVVectorLike* v2,
                 DVVectorLike& kout, DVVectorLike* vout, const Functor& comp)
{
    if (!ensure_init()) return false;
    const size_t n = k1.size() + k2.size();
    if (n == 0) return true;
    if (kout.size() < n || (vout && vout->size() < n)) {
        TRTC_ERR("Merge: output shorter than the two inputs");
        return false;
    }
    if (vout && (!need(*v1, k1.size(), "Merge_By_Key", "value1") || !need(*v2, k2.size(), "Merge_By_Key", "value2")))
        return false;
    // tile of trtc_merge_k / trtc_merge_by_key_k (2 x 2^27 int32, per call):
    // Merge 512 x 16 0.63 ms (256 x 8 0.70); Merge_By_Key 256 x 8 1.02 ms
    // (16 items: 1.5-2.0 ms, three LDS buffers halve the workgroups per CU).
    // Keys wider than 8 bytes keep 2048-output tiles (LDS).
    const bool big = vout != nullptr || k1.elem_size() > 8;
    const unsigned mblock = big ? 256 : 512, mitems = big ? 8 : 16;
    const std::string shape = ", " + std::to_string(mblock) + ", " + std::to_string(mitems) + ">";
    std::string name;
    std::deque<ViewBuf> bufs;
    std::vector<void*> args;
    auto add = [&](const DeviceViewable& d) {
        bufs.push_back(d.view());
        args.push_back(bufs.back().data());
    };
    if (vout) {
        name = "trtc_merge_by_key_k<" + canon_type(k1.name_view_cls()) + ", " + canon_type(k2.name_view_cls()) + ", " +
               canon_type(v1->name_view_cls()) + ", " + canon_type(v2->name_view_cls()) + ", " +
               canon_type(kout.name_view_cls()) + ", " + canon_type(vout->name_view_cls()) + ", " +
               canon_type(comp.name_view_cls()) + shape;
        add(k1);
        add(k2);
        add(*v1);
        add(*v2);
        add(kout);
        add(*vout);
    } else {
        name = "trtc_merge_k<" + canon_type(k1.name_view_cls()) + ", " + canon_type(k2.name_view_cls()) + ", " +
               canon_type(kout.name_view_cls()) + ", " + canon_type(comp.name_view_cls()) + shape;
        add(k1);
        add(k2);
        add(kout);
    }
    add(comp);
    const size_t tile = (size_t)mblock * mitems;
    const size_t tiles = (n + tile - 1) / tile;
    hipStream_t s = cur_stream();
    // merge-path splits of every tile start, all searched in parallel
    PoolBuf splits((tiles + 1) * 8, s);
    if (!splits.p) return false;
    trtc::MergeArgs a{k1.size(), k2.size(), splits.as<unsigned long long>()};
    {
        const std::string sname = "trtc_merge_split_k<" + canon_type(k1.name_view_cls()) + ", " +
                                  canon_type(k2.name_view_cls()) + ", " + canon_type(comp.name_view_cls()) + ", " +
                                  std::to_string(tile) + ">";
        KernelRef sk;
        if (!get_template_kernel(sname, sname, sk)) return false;
        ViewBuf b1 = k1.view(), b2 = k2.view(), bc = comp.view();
        unsigned long long* sp = splits.as<unsigned long long>();
        void* sargs[] = {b1.data(), b2.data(), bc.data(), &a, &sp};
        if (!launch_kernel(sk, dim3((unsigned)((tiles + 1 + 255) / 256)), dim3(256), sargs, 0, s)) return false;
    }
    args.push_back(&a);
    KernelRef k;
    if (!get_template_kernel(name, name, k)) return false;
    return launch_kernel(k, dim3((unsigned)tiles), dim3(mblock), args.data(), 0, s);
}
}  // namespace

bool TRTC_Merge(const DVVectorLike& vec1, const DVVectorLike& vec2, DVVectorLike& vec_out, const Functor& comp)
{
    return merge_views(vec1, vec2, nullptr, nullptr, vec_out, nullptr, comp);
}
bool TRTC_Merge(const DVVectorLike& vec1, const DVVectorLike& vec2, DVVectorLike& vec_out)
{
    return TRTC_Merge(vec1, vec2, vec_out, Functor("Less"));
}
bool TRTC_Merge_By_Key(const DVVectorLike& keys1, const DVVectorLike& keys2, const DVVectorLike& value1,
                       const DVVectorLike& value2, DVVectorLike& keys_out, DVVectorLike& value_out,
                       const Functor& comp)
{
    return merge_views(keys1, keys2, &value1, &value2, keys_out, &value_out, comp);
}
bool TRTC_Merge_By_Key(const DVVectorLike& keys1, const DVVectorLike& keys2, const DVVectorLike& value1,
                       const DVVectorLike& value2, DVVectorLike& keys_out, DVVectorLike& value_out)
{
    return TRTC_Merge_By_Key(keys1, keys2, value1, value2, keys_out, value_out, Functor("Less"));
}

// ================================================================ DVTuple
// The member list text is the one DVZipped generates for its element struct
// ("    <type> <name>;" per member), so equal member lists name one struct.
DVTuple::DVTuple(const std::vector<CapturedDeviceViewable>& elem_map)
{
    std::string body;
    size_t end = 0, align = 1;
    m_view_elems.resize(elem_map.size());
    m_offsets.resize(elem_map.size() + 1);
    for (size_t i = 0; i < elem_map.size(); i++) {
        const DeviceViewable* obj = elem_map[i].obj;
        m_view_elems[i] = obj->view();
        body += "    " + obj->name_view_cls() + " " + elem_map[i].obj_name + ";\n";
        size_t a = view_align(obj->name_view_cls(), m_view_elems[i].size());
        end = align_up(end, a);
        m_offsets[i] = end;
        end += m_view_elems[i].size();
        align = std::max(align, a);
    }
    m_offsets[elem_map.size()] = align_up(std::max<size_t>(end, 1), align);
    m_name_view_cls = add_struct(body);
    register_struct_layout(m_name_view_cls, m_offsets.back(), align);
}

ViewBuf DVTuple::view() const
{
    ViewBuf buf(m_offsets.back(), 0);
    for (size_t i = 0; i < m_view_elems.size(); i++)
        memcpy(buf.data() + m_offsets[i], m_view_elems[i].data(), m_view_elems[i].size());
    return buf;
}
