// Request server: HBM-budget admission, continuous batching of mixed
// sort / scan / reduce requests into segmented launches on several streams,
// and the one-request-at-a-time per-call mode.
#include "server.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <thread>
#include <unordered_map>

#include "reduce.h"
#include "scan.h"
#include "trtc_batch.h"

using namespace trtc_host;
typedef std::chrono::steady_clock Clock;

namespace {

struct Req
{
    int op = 0;
    int KB = 0, VB = 0, kind = 0;  // sorts
    std::string T;                 // scans / reduces: element type
    size_t tsz = 0;
    const char* in_a = nullptr;
    const char* in_b = nullptr;
    size_t n = 0;
    void* out[2] = {nullptr, nullptr};
    size_t out_bytes[2] = {0, 0};
    hipStream_t out_stream = nullptr;
    size_t footprint = 0;  // outputs + temporaries, charged on admission
    Clock::time_point t_submit, t_done;
    int state = 0;  // 0 queued, 1 running, 2 done, 3 failed, 4 released
};

size_t round_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

// Device memory of the server: regions of at least `grow` bytes obtained from
// hipMalloc on demand (up to `limit` in total), handed out best-fit in
// 256-byte units and coalesced on free, so steady-state serving never calls
// hipMalloc.  Blocks are only freed after the GPU work using them completed.
class Arena
{
public:
    Arena(size_t limit, size_t grow) : m_limit(limit), m_grow(grow) {}
    ~Arena()
    {
        for (auto& c : m_regions) hipFree(c.first);
    }
    void* alloc(size_t bytes)
    {
        bytes = round_up(std::max<size_t>(bytes, 1), 256);
        std::lock_guard<std::mutex> lk(m_mtx);
        for (int attempt = 0; attempt < 2; attempt++) {
            auto it = m_by_size.lower_bound(bytes);
            if (it != m_by_size.end()) {
                char* p = it->second;
                size_t sz = it->first;
                int region = m_free[p].second;
                m_by_size.erase(it);
                m_free.erase(p);
                if (sz > bytes) add_free(p + bytes, sz - bytes, region);
                m_used[p] = {bytes, region};
                return p;
            }
            if (attempt) break;
            size_t rs = std::max(bytes, m_grow);
            if (m_total + rs > m_limit) rs = bytes;
            if (m_total + rs > m_limit) return nullptr;
            void* q = nullptr;
            if (hipMalloc(&q, rs) != hipSuccess) {
                (void)hipGetLastError();
                return nullptr;
            }
            m_regions.push_back({(char*)q, rs});
            m_total += rs;
            add_free((char*)q, rs, (int)m_regions.size() - 1);
        }
        return nullptr;
    }
    void free(void* vp)
    {
        char* p = (char*)vp;
        std::lock_guard<std::mutex> lk(m_mtx);
        auto u = m_used.find(p);
        if (u == m_used.end()) return;
        size_t sz = u->second.first;
        int region = u->second.second;
        m_used.erase(u);
        // merge with the following and the preceding free block of the same region
        auto nx = m_free.find(p + sz);
        if (nx != m_free.end() && nx->second.second == region) {
            size_t nsz = nx->second.first;
            erase_free(nx->first, nsz);
            sz += nsz;
        }
        auto pv = m_free.lower_bound(p);
        if (pv != m_free.begin()) {
            --pv;
            if (pv->first + pv->second.first == p && pv->second.second == region) {
                char* q = pv->first;
                size_t qsz = pv->second.first;
                erase_free(q, qsz);
                p = q;
                sz += qsz;
            }
        }
        add_free(p, sz, region);
    }
    bool owns(void* p)
    {
        std::lock_guard<std::mutex> lk(m_mtx);
        return m_used.count((char*)p) != 0;
    }
    // Device memory obtained up front as one region of `bytes` (a batch's
    // ping-pong buffers can take several GB in one block, and blocks never
    // span regions); if hipMalloc refuses, the arena grows on demand as
    // before.
    void reserve(size_t bytes)
    {
        std::lock_guard<std::mutex> lk(m_mtx);
        bytes = round_up(std::min(bytes, m_limit), 256);
        if (m_total >= bytes) return;
        void* q = nullptr;
        if (hipMalloc(&q, bytes) != hipSuccess) {
            (void)hipGetLastError();
            return;
        }
        m_regions.push_back({(char*)q, bytes});
        m_total += bytes;
        add_free((char*)q, bytes, (int)m_regions.size() - 1);
    }
    size_t reserved() const { return m_total; }

private:
    void add_free(char* p, size_t sz, int region)
    {
        m_free[p] = {sz, region};
        m_by_size.emplace(sz, p);
    }
    void erase_free(char* p, size_t sz)
    {
        auto r = m_by_size.equal_range(sz);
        for (auto it = r.first; it != r.second; ++it)
            if (it->second == p) {
                m_by_size.erase(it);
                break;
            }
        m_free.erase(p);
    }
    size_t m_limit, m_grow, m_total = 0;
    std::mutex m_mtx;
    std::vector<std::pair<char*, size_t>> m_regions;
    std::map<char*, std::pair<size_t, int>> m_free;  // address -> (size, region)
    std::multimap<size_t, char*> m_by_size;
    std::unordered_map<char*, std::pair<size_t, int>> m_used;
};

}  // namespace

class TRTC_Server_Impl
{
public:
    size_t budget;
    int nslots;
    bool batching;
    // a batch closes at max_batch requests or once its elements would pass
    // max_batch_elems (it always takes one request); by default only the
    // budget limits it (on the 3000-request stream, batches capped at 2^28
    // elements served 56-60k req/s against 62.5-63.6k uncapped)
    size_t max_batch = 4096;
    size_t max_batch_elems = ~(size_t)0;

    std::mutex mtx;
    std::condition_variable cv_work;  // dispatcher: new request or a free slot
    std::condition_variable cv_done;  // waiters: a request finished
    std::deque<Req> reqs;             // ticket = index
    std::deque<long long> queue;      // queued tickets, arrival order
    std::vector<hipStream_t> streams;
    std::vector<int> busy;
    size_t held = 0, peak = 0;
    unsigned long long batches = 0, batched_reqs = 0, served = 0, failed = 0;
    bool stop = false;
    bool init_ok = false;
    std::thread dispatcher;
    Arena arena;

    struct SlotWork
    {
        std::vector<long long> reqs;
        hipEvent_t ev = nullptr;
        bool ok = true;
        size_t temp_bytes = 0;
        std::vector<void*> temps;  // batch temporaries, freed once the batch is done
    };
    // one waiter thread per slot: blocks on the batch's event and retires it
    std::vector<std::deque<SlotWork>> slot_q;
    std::vector<std::thread> waiters;
    std::condition_variable cv_slot;

    TRTC_Server_Impl(size_t b, int s, bool bt)
        : budget(b), nslots(std::max(1, s)), batching(bt), arena(b + b / 8 + (256 << 20), std::min<size_t>(b, 4ull << 30))
    {
        if (!batching) nslots = 1;
        init_ok = ensure_init();
        if (!init_ok) return;
        // the device-resident buffer pool: the whole admission budget is
        // reserved here, so that serving never waits on hipMalloc (a cold
        // arena grew by multi-GB hipMalloc calls inside the first large batch)
        arena.reserve(budget);
        streams.resize(nslots);
        busy.assign(nslots, 0);
        slot_q.resize(nslots);
        for (int i = 0; i < nslots; i++) hipStreamCreateWithFlags(&streams[i], hipStreamNonBlocking);
        dispatcher = std::thread([this] { run(); });
        if (batching)
            for (int i = 0; i < nslots; i++) waiters.emplace_back([this, i] { wait_slot(i); });
    }

    ~TRTC_Server_Impl()
    {
        {
            std::lock_guard<std::mutex> lk(mtx);
            stop = true;
        }
        cv_work.notify_all();
        cv_slot.notify_all();
        if (dispatcher.joinable()) dispatcher.join();
        for (auto& t : waiters) t.join();
        for (auto s : streams) hipStreamSynchronize(s);
        for (auto& r : reqs)
            for (int k = 0; k < 2; k++)
                if (r.out[k] && r.state != 4) free_mem(r.out[k], r.out_stream);
        for (auto s : streams) {
            hipStreamSynchronize(s);
            hipStreamDestroy(s);
        }
    }

    size_t estimate_footprint(Req& r)
    {
        size_t temps = 64 << 10;
        switch (r.op) {
        case TRTC_REQ_SORT_BY_KEY:
        case TRTC_REQ_SORT:
            r.out_bytes[0] = r.n * r.KB;
            r.out_bytes[1] = r.op == TRTC_REQ_SORT_BY_KEY ? r.n * r.VB : 0;
            if (r.n > 8192) temps += r.n * (r.KB + r.VB);       // ping-pong buffers
            temps += (r.n / 8192 + 1) * 256 * 4 * (r.KB + 1);  // lookback rows, histograms
            break;
        case TRTC_REQ_INCLUSIVE_SCAN:
            r.out_bytes[0] = r.n * r.tsz;
            temps += (r.n / 2048 + 1) * 2 * 8 * ((r.tsz + 3) / 4);
            break;
        default:
            r.out_bytes[0] = r.tsz;
            temps += (r.n / 16384 + 1) * 16;
        }
        return round_up(r.out_bytes[0], 256) + round_up(r.out_bytes[1], 256) + temps;
    }

    // Server memory: the arena, or the stream-ordered pool when the arena
    // cannot serve a block (fragmentation).
    void* get_mem(size_t bytes, hipStream_t s)
    {
        void* p = arena.alloc(bytes);
        return p ? p : pool_alloc(bytes, s);
    }
    void free_mem(void* p, hipStream_t s)
    {
        if (arena.owns(p)) arena.free(p);
        else pool_free(p, s);
    }

    // Allocates the outputs and enqueues the work of `take` on slot `si`;
    // batch temporaries are appended to `temps`.
    bool launch(const std::vector<long long>& take, int si, std::vector<void*>& temps)
    {
        hipStream_t s = streams[si];
        set_stream(s);
        std::vector<Req*> rs;
        {
            std::lock_guard<std::mutex> lk(mtx);
            for (long long t : take) rs.push_back(&reqs[t]);
        }
        for (Req* r : rs) {
            for (int k = 0; k < 2; k++)
                if (r->out_bytes[k]) {
                    r->out[k] = get_mem(r->out_bytes[k], s);
                    if (!r->out[k]) return false;
                }
            r->out_stream = s;
        }
        if (!batching) return launch_single(*rs[0], s);
        // group by operation and element types; one segmented launch per group
        std::map<std::string, std::vector<Req*>> groups;
        for (Req* r : rs) {
            std::string key = std::to_string(r->op) + "/" + std::to_string(r->KB) + "/" + std::to_string(r->VB) + "/" +
                              std::to_string(r->kind) + "/" + r->T;
            groups[key].push_back(r);
        }
        for (auto& g : groups) {
            Req& f = *g.second[0];
            if (f.op == TRTC_REQ_SORT_BY_KEY || f.op == TRTC_REQ_SORT) {
                std::vector<SortSeg> segs;
                for (Req* r : g.second) {
                    SortSeg x;
                    x.keys = (char*)r->out[0];
                    x.vals = (char*)r->out[1];
                    x.n = r->n;
                    x.src_keys = r->in_a;
                    x.src_vals = r->op == TRTC_REQ_SORT_BY_KEY ? r->in_b : nullptr;
                    segs.push_back(x);
                }
                int vb = f.op == TRTC_REQ_SORT_BY_KEY ? f.VB : 0;
                SortPlan plan;
                if (!plan_sort(segs, f.KB, vb, f.kind, false, served_sort_schedule(f.KB, vb), plan)) return false;
                char* sc = (char*)get_mem(plan.scratch_bytes, s);
                char* tk = plan.tmp_elems ? (char*)get_mem(plan.tmp_elems * f.KB, s) : nullptr;
                char* tv = plan.tmp_elems && vb ? (char*)get_mem(plan.tmp_elems * vb, s) : nullptr;
                for (void* p : {(void*)sc, (void*)tk, (void*)tv})
                    if (p) temps.push_back(p);
                if (!sc || (plan.tmp_elems && !tk) || (plan.tmp_elems && vb && !tv)) return false;
                if (!launch_sort_plan(plan, sc, tk, tv, s)) return false;
            } else if (f.op == TRTC_REQ_INCLUSIVE_SCAN) {
                std::vector<ScanSeg> segs;
                for (Req* r : g.second) segs.push_back({r->in_a, (char*)r->out[0], r->n});
                ViewBuf init(f.tsz, 0);
                ScanPlan plan;
                if (!plan_scan(segs, f.T, Functor("Plus"), init, false, default_scan_items(f.tsz), plan)) return false;
                char* sc = (char*)get_mem(plan.scratch_bytes, s);
                if (!sc) return false;
                temps.push_back(sc);
                if (!launch_scan_plan(plan, sc, s)) return false;
            } else {
                std::vector<ReduceSeg> segs;
                for (Req* r : g.second) segs.push_back({r->in_a, r->n, (char*)r->out[0]});
                ReducePlan plan;
                if (!plan_reduce(segs, f.T, Functor("Plus"), plan)) return false;
                char* sc = (char*)get_mem(plan.scratch_bytes, s);
                if (!sc) return false;
                temps.push_back(sc);
                if (!launch_reduce_plan(plan, sc, s)) return false;
            }
        }
        return true;
    }

    // The per-call path: the single-vector functions, one request at a time.
    bool launch_single(Req& r, hipStream_t s)
    {
        switch (r.op) {
        case TRTC_REQ_SORT_BY_KEY:
            return radix_sort_from(r.in_a, r.in_b, r.out[0], r.out[1], r.n, r.KB, r.kind, false, r.VB, s);
        case TRTC_REQ_SORT:
            return radix_sort_from(r.in_a, nullptr, r.out[0], nullptr, r.n, r.KB, r.kind, false, 0, s);
        case TRTC_REQ_INCLUSIVE_SCAN: {
            DVVectorAdaptor in(r.T.c_str(), r.n, (void*)r.in_a);
            DVVectorAdaptor out(r.T.c_str(), r.n, r.out[0]);
            return TRTC_Inclusive_Scan(in, out);
        }
        default: {
            DVVectorAdaptor in(r.T.c_str(), r.n, (void*)r.in_a);
            ViewBuf ret;
            if (!TRTC_Reduce(in, ret)) return false;
            // ret is pageable and local: finish the copy before it goes away
            return check_hip(hipMemcpyAsync(r.out[0], ret.data(), r.tsz, hipMemcpyHostToDevice, s),
                             "hipMemcpyAsync") &&
                   check_hip(hipStreamSynchronize(s), "hipStreamSynchronize");
        }
        }
    }

    void finish(const std::vector<long long>& take, bool ok, size_t temp_bytes, Clock::time_point now)
    {
        // caller holds mtx
        for (long long t : take) {
            Req& r = reqs[t];
            r.t_done = now;
            r.state = ok ? 2 : 3;
            if (ok) served++;
            else failed++;
        }
        held -= temp_bytes;
    }

    void run()
    {
        ensure_init();
        std::unique_lock<std::mutex> lk(mtx);
        while (true) {
            if (stop && queue.empty()) break;
            int slot = -1;
            for (int i = 0; i < nslots; i++)
                if (!busy[i]) {
                    slot = i;
                    break;
                }
            if (slot < 0 || queue.empty()) {
                cv_work.wait(lk);
                continue;
            }
            // admission in arrival order while the footprint fits the budget
            // (an idle server always admits the oldest request)
            std::vector<long long> take;
            size_t add = 0, temps = 0, elems = 0;
            while (!queue.empty() && take.size() < (batching ? max_batch : 1)) {
                Req& r = reqs[queue.front()];
                if (held + add + r.footprint > budget && !(take.empty() && held == 0)) break;
                if (!take.empty() && elems + r.n > max_batch_elems) break;
                elems += r.n;
                add += r.footprint;
                temps += r.footprint - round_up(r.out_bytes[0], 256) - round_up(r.out_bytes[1], 256);
                r.state = 1;
                take.push_back(queue.front());
                queue.pop_front();
            }
            if (take.empty()) {
                cv_work.wait(lk);  // a release or a completion frees budget
                continue;
            }
            held += add;
            peak = std::max(peak, held);
            busy[slot] = 1;
            batches++;
            batched_reqs += take.size();
            lk.unlock();
            std::vector<void*> tmp_mem;
            const Clock::time_point t_launch = Clock::now();
            bool ok = launch(take, slot, tmp_mem);
            if (context().verbose)
                fprintf(stderr, "[server] batch %llu: %zu requests on slot %d, host launch %.1f us\n", batches,
                        take.size(), slot,
                        std::chrono::duration<double, std::micro>(Clock::now() - t_launch).count());
            if (!batching) {
                ok = check_hip(hipStreamSynchronize(streams[slot]), "hipStreamSynchronize") && ok;
                Clock::time_point now = Clock::now();
                lk.lock();
                finish(take, ok, temps, now);
                busy[slot] = 0;
                cv_done.notify_all();
                continue;
            }
            SlotWork w;
            w.reqs = take;
            w.ok = ok;
            w.temp_bytes = temps;
            w.temps = std::move(tmp_mem);
            hipEventCreateWithFlags(&w.ev, hipEventDisableTiming);
            hipEventRecord(w.ev, streams[slot]);
            lk.lock();
            slot_q[slot].push_back(std::move(w));
            cv_slot.notify_all();
        }
    }

    void wait_slot(int i)
    {
        ensure_init();
        std::unique_lock<std::mutex> lk(mtx);
        while (true) {
            while (slot_q[i].empty() && !(stop && queue.empty() && !busy[i])) cv_slot.wait(lk);
            if (slot_q[i].empty()) break;
            SlotWork w = std::move(slot_q[i].front());
            slot_q[i].pop_front();
            lk.unlock();
            bool ok = hipEventSynchronize(w.ev) == hipSuccess && w.ok;
            Clock::time_point now = Clock::now();
            hipEventDestroy(w.ev);
            for (void* p : w.temps) free_mem(p, streams[i]);
            lk.lock();
            finish(w.reqs, ok, w.temp_bytes, now);
            busy[i] = 0;
            cv_done.notify_all();
            cv_work.notify_all();
        }
    }
};

TRTC_Server::TRTC_Server(size_t hbm_budget, int slots, bool batching)
    : m_impl(new TRTC_Server_Impl(hbm_budget, slots, batching))
{
}

TRTC_Server::~TRTC_Server() { delete m_impl; }

long long TRTC_Server::Submit(int op, const char* key_elem, const char* val_elem, const void* in_a, const void* in_b,
                              size_t n)
{
    TRTC_Server_Impl& S = *m_impl;
    if (!S.init_ok || !key_elem || !in_a || n == 0) return -1;
    Req r;
    r.op = op;
    r.in_a = (const char*)in_a;
    r.in_b = (const char*)in_b;
    r.n = n;
    if (op == TRTC_REQ_SORT_BY_KEY || op == TRTC_REQ_SORT) {
        if (!radix_key_info(key_elem, r.KB, r.kind)) return -1;
        if (op == TRTC_REQ_SORT_BY_KEY) {
            const ScalarInfo* vi = val_elem ? scalar_info(val_elem) : nullptr;
            if (!vi || !in_b || !(vi->size == 1 || vi->size == 2 || vi->size == 4 || vi->size == 8)) return -1;
            r.VB = (int)vi->size;
        }
        if (n >= (1ull << 30)) return -1;
    } else if (op == TRTC_REQ_INCLUSIVE_SCAN || op == TRTC_REQ_REDUCE) {
        const ScalarInfo* si = scalar_info(key_elem);
        if (!si || si->code == TC_BOOL) return -1;
        r.T = si->canon;
        r.tsz = si->size;
    } else {
        return -1;
    }
    r.footprint = S.estimate_footprint(r);
    r.t_submit = Clock::now();
    long long t;
    {
        std::lock_guard<std::mutex> lk(S.mtx);
        t = (long long)S.reqs.size();
        S.reqs.push_back(r);
        S.queue.push_back(t);
    }
    S.cv_work.notify_one();
    return t;
}

bool TRTC_Server::Wait(long long ticket)
{
    TRTC_Server_Impl& S = *m_impl;
    std::unique_lock<std::mutex> lk(S.mtx);
    if (ticket < 0 || ticket >= (long long)S.reqs.size()) return false;
    S.cv_done.wait(lk, [&] { return S.reqs[ticket].state >= 2; });
    return S.reqs[ticket].state != 3;
}

bool TRTC_Server::Done(long long ticket)
{
    TRTC_Server_Impl& S = *m_impl;
    std::lock_guard<std::mutex> lk(S.mtx);
    return ticket >= 0 && ticket < (long long)S.reqs.size() && S.reqs[ticket].state >= 2;
}

void* TRTC_Server::Output(long long ticket, int which)
{
    TRTC_Server_Impl& S = *m_impl;
    std::lock_guard<std::mutex> lk(S.mtx);
    if (ticket < 0 || ticket >= (long long)S.reqs.size() || which < 0 || which > 1) return nullptr;
    const Req& r = S.reqs[ticket];
    return r.state == 2 ? r.out[which] : nullptr;
}

double TRTC_Server::Latency_Us(long long ticket)
{
    TRTC_Server_Impl& S = *m_impl;
    std::lock_guard<std::mutex> lk(S.mtx);
    if (ticket < 0 || ticket >= (long long)S.reqs.size()) return -1;
    const Req& r = S.reqs[ticket];
    if (r.state < 2) return -1;
    return std::chrono::duration<double, std::micro>(r.t_done - r.t_submit).count();
}

void TRTC_Server::Release(long long ticket)
{
    TRTC_Server_Impl& S = *m_impl;
    {
        std::lock_guard<std::mutex> lk(S.mtx);
        if (ticket < 0 || ticket >= (long long)S.reqs.size()) return;
        Req& r = S.reqs[ticket];
        if (r.state != 2 && r.state != 3) return;
        for (int k = 0; k < 2; k++) {
            if (r.out[k]) S.free_mem(r.out[k], r.out_stream);
            r.out[k] = nullptr;
        }
        S.held -= round_up(r.out_bytes[0], 256) + round_up(r.out_bytes[1], 256);
        r.state = 4;
    }
    S.cv_work.notify_one();
}

void TRTC_Server::Set_Batch_Limits(size_t max_requests, size_t max_elements)
{
    TRTC_Server_Impl& S = *m_impl;
    std::lock_guard<std::mutex> lk(S.mtx);
    S.max_batch = std::max<size_t>(1, max_requests);
    S.max_batch_elems = std::max<size_t>(1, max_elements);
}

std::string TRTC_Server::Stats()
{
    TRTC_Server_Impl& S = *m_impl;
    std::lock_guard<std::mutex> lk(S.mtx);
    char b[512];
    snprintf(b, sizeof(b),
             "{\"batching\": %s, \"slots\": %d, \"budget_bytes\": %zu, \"requests\": %zu, \"served\": %llu, "
             "\"failed\": %llu, \"batches\": %llu, \"mean_batch\": %.2f, \"peak_held_bytes\": %zu, "
             "\"arena_reserved_bytes\": %zu}",
             S.batching ? "true" : "false", S.nslots, S.budget, S.reqs.size(), S.served, S.failed, S.batches,
             S.batches ? (double)S.batched_reqs / S.batches : 0.0, S.peak, S.arena.reserved());
    return b;
}

double TRTC_Server_Replay(TRTC_Server& server, size_t count, const double* arrival_us, const int* ops,
                          const char* const* key_elem, const char* const* val_elem, const void* const* in_a,
                          const void* const* in_b, const size_t* n, double* latency_us)
{
    std::vector<std::atomic<long long>> tickets(count);
    for (auto& t : tickets) t.store(-2);
    Clock::time_point start = Clock::now() + std::chrono::milliseconds(2);
    std::atomic<double> last_done_us(0.0);
    std::thread consumer([&] {
        double last = 0;
        for (size_t i = 0; i < count; i++) {
            long long t;
            while ((t = tickets[i].load()) == -2) std::this_thread::yield();
            if (t < 0) {
                latency_us[i] = -1;
                continue;
            }
            bool ok = server.Wait(t);
            double lat = server.Latency_Us(t);
            latency_us[i] = ok ? lat : -1;
            last = std::max(last, arrival_us[i] + lat);
            server.Release(t);
        }
        last_done_us.store(last);
    });
    for (size_t i = 0; i < count; i++) {
        Clock::time_point due = start + std::chrono::nanoseconds((long long)(arrival_us[i] * 1000.0));
        while (true) {
            Clock::time_point now = Clock::now();
            if (now >= due) break;
            auto left = due - now;
            if (left > std::chrono::microseconds(300)) std::this_thread::sleep_for(left - std::chrono::microseconds(200));
            else std::this_thread::yield();
        }
        tickets[i].store(server.Submit(ops[i], key_elem[i], val_elem ? val_elem[i] : nullptr, in_a[i],
                                       in_b ? in_b[i] : nullptr, n[i]));
    }
    consumer.join();
    return last_done_us.load() * 1e-6;
}
