// Batched segmented primitives (extensions; the reference has no segmented
// sort) and the graph-captured TRTC_Batch.  Each segment [offsets[s],
// offsets[s+1]) ends up exactly as a separate call on that sub-range would
// leave it; all segments share one launch per kernel stage.
#include <memory>

#include "batch.h"
#include "scan.h"
#include "sort.h"
#include "trtc_batch.h"

using namespace trtc_host;

namespace {

bool check_offsets(const size_t* offsets, size_t num_segments, size_t limit, const char* who)
{
    if (!offsets && num_segments) {
        TRTC_ERR("%s: offsets is NULL", who);
        return false;
    }
    for (size_t s = 0; s < num_segments; s++)
        if (offsets[s + 1] < offsets[s] || offsets[s + 1] > limit) {
            TRTC_ERR("%s: bad segment %zu [%zu, %zu) (vector length %zu)", who, s, offsets[s], offsets[s + 1], limit);
            return false;
        }
    return true;
}

struct SortSpec
{
    std::vector<SortSeg> segs;
    int KB = 0, VB = 0, kind = 0;
    bool desc = false;
};

bool sort_spec(DVVectorLike& keys, DVVectorLike* values, const size_t* offsets, size_t num_segments,
               const Functor& comp, SortSpec& sp, const char* who)
{
    const std::string& cn = comp.name_view_cls();
    if (!(cn == "Less" || cn == "Greater") || !radix_key_info(keys.name_elem_cls(), sp.KB, sp.kind)) {
        TRTC_ERR("%s: comparator '%s' on '%s' is not supported (Less/Greater on integer and floating-point keys)", who,
                 cn.c_str(), keys.name_elem_cls().c_str());
        return false;
    }
    sp.desc = cn == "Greater";
    char* kp = (char*)contiguous_data(keys);
    char* vp = values ? (char*)contiguous_data(*values) : nullptr;
    if (!kp || (values && !vp)) {
        TRTC_ERR("%s: keys and values must be DVVector or DVVectorAdaptor", who);
        return false;
    }
    size_t limit = keys.size();
    if (values) {
        sp.VB = (int)values->elem_size();
        if (!(sp.VB == 1 || sp.VB == 2 || sp.VB == 4 || sp.VB == 8)) {
            TRTC_ERR("%s: value elements of %d bytes are not supported", who, sp.VB);
            return false;
        }
        limit = std::min(limit, values->size());
    }
    if (!check_offsets(offsets, num_segments, limit, who)) return false;
    for (size_t s = 0; s < num_segments; s++) {
        size_t b = offsets[s], n = offsets[s + 1] - offsets[s];
        SortSeg g;
        g.keys = kp + b * sp.KB;
        g.vals = vp ? vp + b * sp.VB : nullptr;
        g.n = n;
        sp.segs.push_back(g);
    }
    return true;
}

struct ScanSpec
{
    std::vector<ScanSeg> segs;
    std::string T;
};

bool scan_spec(const DVVectorLike& in, DVVectorLike& out, const size_t* offsets, size_t num_segments, ScanSpec& sp,
               const char* who)
{
    char* ip = (char*)contiguous_data(in);
    char* op = (char*)contiguous_data(out);
    if (!ip || !op) {
        TRTC_ERR("%s: vectors must be DVVector or DVVectorAdaptor", who);
        return false;
    }
    if (canon_type(in.name_elem_cls()) != canon_type(out.name_elem_cls())) {
        TRTC_ERR("%s: input and output element types differ", who);
        return false;
    }
    if (!check_offsets(offsets, num_segments, std::min(in.size(), out.size()), who)) return false;
    sp.T = canon_type(in.name_elem_cls());
    size_t es = in.elem_size();
    for (size_t s = 0; s < num_segments; s++)
        sp.segs.push_back({ip + offsets[s] * es, op + offsets[s] * es, offsets[s + 1] - offsets[s]});
    return true;
}

struct ReduceSpec
{
    std::vector<ReduceSeg> segs;
    std::string T;
};

bool reduce_spec(const DVVectorLike& in, const size_t* offsets, size_t num_segments, DVVectorLike& results,
                 ReduceSpec& sp, const char* who)
{
    char* ip = (char*)contiguous_data(in);
    char* rp = (char*)contiguous_data(results);
    if (!ip || !rp) {
        TRTC_ERR("%s: vectors must be DVVector or DVVectorAdaptor", who);
        return false;
    }
    if (canon_type(in.name_elem_cls()) != canon_type(results.name_elem_cls()) || results.size() < num_segments) {
        TRTC_ERR("%s: results must have the input's element type and num_segments elements", who);
        return false;
    }
    if (!check_offsets(offsets, num_segments, in.size(), who)) return false;
    sp.T = canon_type(in.name_elem_cls());
    size_t es = in.elem_size();
    for (size_t s = 0; s < num_segments; s++)
        sp.segs.push_back({ip + offsets[s] * es, offsets[s + 1] - offsets[s], rp + s * es});
    return true;
}

bool segmented_sort(DVVectorLike& keys, DVVectorLike* values, const size_t* offsets, size_t num_segments,
                    const Functor& comp, const char* who)
{
    if (!ensure_init()) return false;
    SortSpec sp;
    if (!sort_spec(keys, values, offsets, num_segments, comp, sp, who)) return false;
    hipStream_t s = cur_stream();
    SortSchedule sched = schedule_for(sp.segs, sp.KB, sp.VB, sp.kind, sp.desc, s);
    return run_sort_batch(sp.segs, sp.KB, sp.VB, sp.kind, sp.desc, sched, s);
}

}  // namespace

bool TRTC_Segmented_Sort_By_Key(DVVectorLike& keys, DVVectorLike& values, const size_t* offsets,
                                size_t num_segments, const Functor& comp)
{
    return segmented_sort(keys, &values, offsets, num_segments, comp, "Segmented_Sort_By_Key");
}

bool TRTC_Segmented_Sort(DVVectorLike& keys, const size_t* offsets, size_t num_segments, const Functor& comp)
{
    return segmented_sort(keys, nullptr, offsets, num_segments, comp, "Segmented_Sort");
}

static bool segmented_scan(const DVVectorLike& in, DVVectorLike& out, const size_t* offsets, size_t num_segments,
                           bool exclusive, const char* who)
{
    if (!ensure_init()) return false;
    ScanSpec sp;
    if (!scan_spec(in, out, offsets, num_segments, sp, who)) return false;
    const ScalarInfo* si = scalar_info(sp.T);
    if (!si) {
        TRTC_ERR("%s: element type %s is not a scalar", who, sp.T.c_str());
        return false;
    }
    ViewBuf init(si->size, 0);
    return run_scan_batch(sp.segs, sp.T, Functor("Plus"), init, exclusive, default_scan_items(si->size), cur_stream());
}

bool TRTC_Segmented_Inclusive_Scan(const DVVectorLike& vec_in, DVVectorLike& vec_out, const size_t* offsets,
                                   size_t num_segments)
{
    return segmented_scan(vec_in, vec_out, offsets, num_segments, false, "Segmented_Inclusive_Scan");
}

bool TRTC_Segmented_Exclusive_Scan(const DVVectorLike& vec_in, DVVectorLike& vec_out, const size_t* offsets,
                                   size_t num_segments)
{
    return segmented_scan(vec_in, vec_out, offsets, num_segments, true, "Segmented_Exclusive_Scan");
}

bool TRTC_Segmented_Reduce(const DVVectorLike& vec_in, const size_t* offsets, size_t num_segments,
                           DVVectorLike& results, const Functor& op)
{
    if (!ensure_init()) return false;
    ReduceSpec sp;
    if (!reduce_spec(vec_in, offsets, num_segments, results, sp, "Segmented_Reduce")) return false;
    return run_reduce_batch(sp.segs, sp.T, op, cur_stream());
}

void TRTC_Set_Batch_Autotune(bool on) { set_autotune(on); }
bool TRTC_Set_Batch_Schedule(const char* schedule) { return set_forced_schedule(schedule); }
std::string TRTC_Batch_Tuning_Report() { return tuning_report(); }

// ---------------------------------------------------------------- TRTC_Batch
class TRTC_Batch_Impl
{
public:
    struct Op
    {
        int type = 0;  // 0 sort, 1 scan, 2 reduce
        SortSpec sort;
        SortSchedule sched;
        SortPlan sort_plan;
        ScanSpec scan;
        ScanPlan scan_plan;
        ReduceSpec red;
        ReducePlan red_plan;
        std::unique_ptr<Functor> op;
        void* scratch = nullptr;
        void* tk = nullptr;
        void* tv = nullptr;
    };
    std::vector<Op> ops;
    hipGraphExec_t exec = nullptr;
    bool ready = false;

    ~TRTC_Batch_Impl() { release(); }
    void release()
    {
        if (exec) {
            hipDeviceSynchronize();
            hipGraphExecDestroy(exec);
            exec = nullptr;
        }
        for (auto& o : ops) {
            if (o.scratch) hipFree(o.scratch);
            if (o.tk) hipFree(o.tk);
            if (o.tv) hipFree(o.tv);
            o.scratch = o.tk = o.tv = nullptr;
        }
        ready = false;
    }
    bool alloc(void** p, size_t bytes)
    {
        if (!bytes) return true;
        return check_hip(hipMalloc(p, bytes), "hipMalloc");
    }
};

TRTC_Batch::TRTC_Batch() : m_impl(new TRTC_Batch_Impl) {}
TRTC_Batch::~TRTC_Batch() { delete m_impl; }

bool TRTC_Batch::Add_Sort(DVVectorLike& keys, const size_t* offsets, size_t num_segments, const Functor& comp)
{
    if (!ensure_init()) return false;
    TRTC_Batch_Impl::Op o;
    o.type = 0;
    if (!sort_spec(keys, nullptr, offsets, num_segments, comp, o.sort, "Batch::Add_Sort")) return false;
    m_impl->release();
    m_impl->ops.push_back(std::move(o));
    return true;
}

bool TRTC_Batch::Add_Sort_By_Key(DVVectorLike& keys, DVVectorLike& values, const size_t* offsets,
                                 size_t num_segments, const Functor& comp)
{
    if (!ensure_init()) return false;
    TRTC_Batch_Impl::Op o;
    o.type = 0;
    if (!sort_spec(keys, &values, offsets, num_segments, comp, o.sort, "Batch::Add_Sort_By_Key")) return false;
    m_impl->release();
    m_impl->ops.push_back(std::move(o));
    return true;
}

bool TRTC_Batch::Add_Inclusive_Scan(const DVVectorLike& vec_in, DVVectorLike& vec_out, const size_t* offsets,
                                    size_t num_segments, const Functor& op)
{
    if (!ensure_init()) return false;
    TRTC_Batch_Impl::Op o;
    o.type = 1;
    if (!scan_spec(vec_in, vec_out, offsets, num_segments, o.scan, "Batch::Add_Inclusive_Scan")) return false;
    o.op.reset(new Functor(op));
    m_impl->release();
    m_impl->ops.push_back(std::move(o));
    return true;
}

bool TRTC_Batch::Add_Reduce(const DVVectorLike& vec_in, const size_t* offsets, size_t num_segments,
                            DVVectorLike& results, const Functor& op)
{
    if (!ensure_init()) return false;
    TRTC_Batch_Impl::Op o;
    o.type = 2;
    if (!reduce_spec(vec_in, offsets, num_segments, results, o.red, "Batch::Add_Reduce")) return false;
    o.op.reset(new Functor(op));
    m_impl->release();
    m_impl->ops.push_back(std::move(o));
    return true;
}

bool TRTC_Batch::Instantiate()
{
    if (!ensure_init()) return false;
    TRTC_Batch_Impl& B = *m_impl;
    B.release();
    hipStream_t s = cur_stream();
    // plans, persistent buffers and descriptor tables (uploaded before capture)
    for (auto& o : B.ops) {
        std::vector<char> table;
        if (o.type == 0) {
            o.sched = schedule_for(o.sort.segs, o.sort.KB, o.sort.VB, o.sort.kind, o.sort.desc, s);
            if (!plan_sort(o.sort.segs, o.sort.KB, o.sort.VB, o.sort.kind, o.sort.desc, o.sched, o.sort_plan))
                return false;
            const SortPlan& P = o.sort_plan;
            if (!B.alloc(&o.scratch, P.scratch_bytes) || !B.alloc(&o.tk, P.tmp_elems * P.KB) ||
                !B.alloc(&o.tv, P.tmp_elems * P.VB))
                return false;
            table.resize(P.table_bytes);
            fill_sort_tables(P, (char*)o.scratch, (char*)o.tk, (char*)o.tv, table.data());
        } else if (o.type == 1) {
            const ScalarInfo* si = scalar_info(o.scan.T);
            if (!si) return false;
            ViewBuf init(si->size, 0);
            if (!plan_scan(o.scan.segs, o.scan.T, *o.op, init, false, default_scan_items(si->size), o.scan_plan))
                return false;
            if (!B.alloc(&o.scratch, o.scan_plan.scratch_bytes)) return false;
            table.resize(o.scan_plan.table_bytes);
            fill_scan_tables(o.scan_plan, (char*)o.scratch, table.data());
        } else {
            if (!plan_reduce(o.red.segs, o.red.T, *o.op, o.red_plan)) return false;
            if (!B.alloc(&o.scratch, o.red_plan.scratch_bytes)) return false;
            table.resize(o.red_plan.table_bytes);
            fill_reduce_tables(o.red_plan, (char*)o.scratch, table.data());
        }
        if (!table.empty() &&
            !check_hip(hipMemcpy(o.scratch, table.data(), table.size(), hipMemcpyHostToDevice), "hipMemcpy"))
            return false;
        // kernels resolved (compiled and loaded, for non-built-in types) before capture
        bool prepared = o.type == 0 ? prepare_sort(o.sort_plan)
                                    : (o.type == 1 ? prepare_scan(o.scan_plan) : prepare_reduce(o.red_plan));
        if (!prepared) return false;
    }
    hipStream_t cs = nullptr;
    if (!check_hip(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking), "hipStreamCreate")) return false;
    hipGraph_t graph = nullptr;
    bool ok = check_hip(hipDeviceSynchronize(), "hipDeviceSynchronize");
    if (ok) ok = check_hip(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
    if (ok) {
        set_capturing(true);
        for (auto& o : B.ops) {
            if (!ok) break;
            if (o.type == 0) {
                ok = enqueue_sort(o.sort_plan, (char*)o.scratch, cs);
                if (ok && (o.sort_plan.KB & 1)) {
                    for (const auto& C : o.sort_plan.cls)
                        for (size_t i = 0; ok && i < C.segs.size(); i++) {
                            const SortSeg& g = C.segs[i];
                            if (g.n <= C.tile) continue;
                            ok = check_hip(hipMemcpyAsync(g.keys, (char*)o.tk + C.tmp_off[i] * o.sort_plan.KB,
                                                          g.n * o.sort_plan.KB, hipMemcpyDeviceToDevice, cs),
                                           "hipMemcpyAsync");
                            if (ok && o.sort_plan.VB)
                                ok = check_hip(hipMemcpyAsync(g.vals, (char*)o.tv + C.tmp_off[i] * o.sort_plan.VB,
                                                              g.n * o.sort_plan.VB, hipMemcpyDeviceToDevice, cs),
                                               "hipMemcpyAsync");
                        }
                }
            } else if (o.type == 1) {
                ok = enqueue_scan(o.scan_plan, (char*)o.scratch, cs);
            } else {
                ok = enqueue_reduce(o.red_plan, (char*)o.scratch, cs);
            }
        }
        set_capturing(false);
        hipError_t e = hipStreamEndCapture(cs, &graph);
        ok = ok && check_hip(e, "hipStreamEndCapture");
    }
    if (ok) ok = check_hip(hipGraphInstantiate(&B.exec, graph, nullptr, nullptr, 0), "hipGraphInstantiate");
    if (graph) hipGraphDestroy(graph);
    hipStreamDestroy(cs);
    if (!ok) {
        B.release();
        return false;
    }
    B.ready = true;
    return true;
}

bool TRTC_Batch::Launch()
{
    if (!m_impl->ready && !Instantiate()) return false;
    return launch_graph(m_impl->exec, "trtc_batch_graph", cur_stream());
}

std::string TRTC_Batch::Describe() const
{
    std::string r = "[";
    bool first = true;
    for (auto& o : m_impl->ops) {
        r += first ? "" : ", ";
        first = false;
        if (o.type == 0)
            r += "{\"op\": \"sort\", \"segments\": " + std::to_string(o.sort.segs.size()) + ", \"schedule\": \"" +
                 describe(o.sched) + "\"}";
        else if (o.type == 1)
            r += "{\"op\": \"inclusive_scan\", \"segments\": " + std::to_string(o.scan.segs.size()) +
                 ", \"items\": " + std::to_string(o.scan_plan.items) + "}";
        else
            r += "{\"op\": \"reduce\", \"segments\": " + std::to_string(o.red.segs.size()) + "}";
    }
    return r + "]";
}
