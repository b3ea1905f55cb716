// Serving: a request loop with continuous batching.
//
//   submit -> [prep threads] each takes the queued new graphs (up to
//             opts.group_graphs graphs and kGroupPairs pairs), canonicalises
//             them on the GPU in one pass as separate vertex ranges and plans
//             them as one block-diagonal unit (a single graph is planned on
//             its own); admission keeps the resident set under the HBM
//             budget, evicting idle units least recently used first
//          -> [GPU worker] every ready request joins the next launch: the
//             single-graph units in one persistent wedge-kernel launch over
//             all of their plans, each multi-graph unit in one launch that
//             yields all of its graphs' counts (per-vertex sums reduced per
//             graph)
//          -> completion queue (btc_server_poll)
//
// With captured preparation (the default) a prep thread instead builds and
// counts each new graph with one replay of a captured HIP graph
// (btc_pipeline, one per size class per thread), answers the requests
// waiting for it right away, and then publishes a copy of its L and plan
// as a resident unit for later requests.
//
// Requests for a graph key that is resident (or being prepared) reuse it.
// While a launch is on the GPU, new arrivals are prepared concurrently on
// the prep streams.  Under light load units hold one graph; under heavy
// load the queue is drained in groups, so preparation keeps up.  Every count
// is exact and independent of what it is batched or grouped with (a graph's
// count is the sum over its own vertex range).  No reference counterpart
// (the reference is a one-shot CLI).
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <list>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>

#include <omp.h>

#include "internal.h"

namespace {

using Clock = std::chrono::steady_clock;

// A prep thread groups at most opts.group_graphs (<= kGroupGraphsMax) new
// graphs and kGroupPairs input pairs.
constexpr int kGroupGraphsMax = 64;
constexpr size_t kGroupPairs = size_t(1) << 22;
constexpr int64_t kGroupVertices = int64_t(1) << 25;  // keeps a group's vertex ids well inside int32
// Captured preparation: graphs up to these sizes go through a pipeline of
// their size class; a thread keeps at most kPipeCacheBytes of pipelines.
constexpr int64_t kPipeMaxPairs = int64_t(1) << 23;
constexpr int64_t kPipeMaxIds = int64_t(1) << 24;
constexpr int64_t kPipeCacheBytes = int64_t(4) << 30;
// Captured preparation turns to groups beyond this many queued graphs per
// prep thread.
constexpr size_t kBacklogPerThread = 8;

int64_t pow2_at_least(int64_t x, int64_t lo) {
    int64_t c = lo;
    while (c < x) c <<= 1;
    return c;
}

// A prep thread's pipelines by size class (ids, pairs, graphs per run),
// least recently used dropped first beyond kPipeCacheBytes.
struct PipeCache {
    struct Slot {
        btc_pipeline p = nullptr;
        int64_t bytes = 0;
        uint64_t used = 0;
    };
    struct Key {
        int32_t ncap;
        int64_t ecap;
        int32_t graphs;
        bool operator==(const Key& o) const { return ncap == o.ncap && ecap == o.ecap && graphs == o.graphs; }
    };
    std::vector<std::pair<Key, Slot>> slots;
    int64_t bytes = 0;
    uint64_t tick = 0;
    PipeCache() = default;
    PipeCache(const PipeCache&) = delete;
    PipeCache& operator=(const PipeCache&) = delete;
    ~PipeCache() {
        for (auto& x : slots) btc_pipeline_free(x.second.p);
    }
    // The pipeline of class (ncap, ecap, graphs), created on s when missing.
    int get(int32_t ncap, int64_t ecap, int32_t graphs, uint32_t flags, const btc_count_opts& co, hipStream_t s,
            btc_pipeline* out) {
        const Key key{ncap, ecap, graphs};
        for (auto& x : slots)
            if (x.first == key) {
                x.second.used = ++tick;
                *out = x.second.p;
                return BTC_OK;
            }
        btc_pipeline p = nullptr;
        int rc = btc_pipeline_create_batch(ncap, ecap, graphs, flags, &co, s, &p);
        if (rc != BTC_OK) return rc;
        btc_pipeline_info pi;
        btc_pipeline_get_info(p, &pi);
        while (!slots.empty() && bytes + pi.device_bytes > kPipeCacheBytes) {
            size_t old = 0;
            for (size_t k = 1; k < slots.size(); ++k)
                if (slots[k].second.used < slots[old].second.used) old = k;
            bytes -= slots[old].second.bytes;
            btc_pipeline_free(slots[old].second.p);
            slots.erase(slots.begin() + (long)old);
        }
        Slot sl;
        sl.p = p;
        sl.bytes = pi.device_bytes;
        sl.used = ++tick;
        slots.emplace_back(key, sl);
        bytes += sl.bytes;
        *out = p;
        return BTC_OK;
    }
};

struct Unit {  // what is prepared, resident and evicted together
    btc_graph graph = nullptr;  // a single graph: its L and plan
    btc_plan plan = nullptr;
    btc_batch batch = nullptr;  // several graphs: block-diagonal L, plan and per-graph sums
    std::vector<uint64_t> keys; // member graph keys, in vertex-range order
    int64_t bytes = 0;
    std::list<Unit*>::iterator lru;
    bool in_lru = false;
    Unit() = default;
    Unit(const Unit&) = delete;
    Unit& operator=(const Unit&) = delete;
    ~Unit() {
        if (plan) btc_plan_free(plan);
        if (graph) btc_graph_free(graph);
        if (batch) btc_batch_free(batch);
    }
};

struct Entry {  // one graph known to the server
    std::shared_ptr<Unit> unit;
    int idx = 0;    // its range in a multi-graph unit
    int state = 0;  // 0 preparing, 1 resident, 2 failed
    int status = BTC_OK;
    int pins = 0;   // requests holding the graph (waiting, ready or launched)
};

struct Request {
    uint64_t id = 0;
    uint64_t key = 0;
    bool resident_hit = false;
    double t_submit = 0, t_start = 0;
};

struct PrepJob {
    uint64_t key = 0;
    std::vector<int32_t> src, dst;
    int32_t n = 0;
};

// Vertex count of a job (n, or max id + 1 when n <= 0); -1 when an id is
// negative or not below n.
int64_t job_vertices(const PrepJob& j) {
    int32_t lo = 0, hi = -1;
    for (size_t e = 0; e < j.src.size(); ++e) {
        lo = std::min(lo, std::min(j.src[e], j.dst[e]));
        hi = std::max(hi, std::max(j.src[e], j.dst[e]));
    }
    const int64_t n = j.n > 0 ? j.n : (int64_t)hi + 1;
    return (lo < 0 || hi >= n) ? -1 : n;
}

}  // namespace

struct btc_server_s {
    btc_server_opts opts{};
    int device = 0;
    Clock::time_point t0 = Clock::now();
    std::mutex mu;
    std::condition_variable cv_prep, cv_gpu, cv_done;
    bool stop = false;
    uint64_t next_id = 1;
    std::unordered_map<uint64_t, Entry> cache;
    std::list<Unit*> lru;  // front = most recently used resident unit
    int64_t resident_bytes = 0;
    std::deque<PrepJob> prep_queue;
    std::unordered_map<uint64_t, std::vector<Request>> waiting;  // key -> requests waiting for its plan
    std::deque<Request> ready;
    std::deque<btc_result> done;
    std::vector<std::thread> workers;
    btc_server_stats stats{};

    double now() const { return std::chrono::duration<double>(Clock::now() - t0).count(); }

    bool unit_idle(const Unit* u) {
        for (uint64_t k : u->keys)
            if (cache[k].pins > 0) return false;
        return true;
    }

    // Caller holds mu.  Evicts idle resident units (LRU first) until `need`
    // more bytes fit; returns whether they fit.
    bool make_room(int64_t need) {
        auto it = lru.end();
        while (resident_bytes + need > opts.hbm_budget_bytes && it != lru.begin()) {
            --it;
            Unit* u = *it;
            if (!unit_idle(u)) continue;
            resident_bytes -= u->bytes;
            it = lru.erase(it);
            const std::vector<uint64_t> keys = u->keys;
            for (uint64_t k : keys) cache.erase(k);  // the last reference frees the unit
            stats.evictions += (int64_t)keys.size();
        }
        return resident_bytes + need <= opts.hbm_budget_bytes;
    }

    void finish(const Request& r, int status, uint64_t tri, int batch) {
        btc_result x{};
        x.request_id = r.id;
        x.graph_key = r.key;
        x.triangles = tri;
        x.status = status;
        x.resident_hit = r.resident_hit ? 1 : 0;
        x.batch_size = batch;
        x.t_submit = r.t_submit;
        x.t_start = r.t_start;
        x.t_done = now();
        done.push_back(x);
        ++stats.completed;
        if (status != BTC_OK) ++stats.failed;
    }

    // Caller holds mu: the graph `key` failed to prepare.
    void fail_key(uint64_t key, int rc) {
        Entry& e = cache[key];
        e.state = 2;
        e.status = rc;
        for (Request& r : waiting[key]) {
            --e.pins;
            finish(r, rc, 0, 0);
        }
        waiting.erase(key);
        cv_done.notify_all();
    }

    // Caller holds mu: unit u (prepared for `keys`) becomes resident.
    void publish(std::shared_ptr<Unit> u) {
        resident_bytes += u->bytes;
        lru.push_front(u.get());
        u->lru = lru.begin();
        u->in_lru = true;
        for (size_t k = 0; k < u->keys.size(); ++k) {
            const uint64_t key = u->keys[k];
            Entry& e = cache[key];
            e.unit = u;
            e.idx = (int)k;
            e.state = 1;
            ++stats.prepared;
            for (Request& r : waiting[key]) ready.push_back(r);
            waiting.erase(key);
        }
        cv_gpu.notify_one();
    }

    // Canonicalises (host or GPU, per the server's flags) and plans one graph.
    int prepare_single(PrepJob& job, uint32_t cf, const btc_count_opts& co, hipStream_t s, btcx::PinnedBuf& pin,
                       std::shared_ptr<Unit>& out) {
        btc_graph g = nullptr;
        int rc = BTC_OK;
        if (cf & BTC_CANON_DEVICE) {
            rc = btcx::guard([&] {
                const size_t nnz = job.src.size();
                btcx::TmpBuf<int32_t> ds, dd;
                ds.alloc(nnz, s);
                dd.alloc(nnz, s);
                if (nnz) {
                    // staged through pinned memory: asynchronous copies (the
                    // canonicalisation synchronises s before the next use)
                    int32_t* h = reinterpret_cast<int32_t*>(pin.reserve(nnz * 8));
                    std::memcpy(h, job.src.data(), nnz * 4);
                    std::memcpy(h + nnz, job.dst.data(), nnz * 4);
                    BTC_HIP(hipMemcpyAsync(ds.p, h, nnz * 4, hipMemcpyHostToDevice, s));
                    BTC_HIP(hipMemcpyAsync(dd.p, h + nnz, nnz * 4, hipMemcpyHostToDevice, s));
                }
                btcx::check_status(btc_graph_from_coo_device(ds.p, dd.p, (int64_t)nnz, job.n,
                                                             cf & ~BTC_CANON_DEVICE, s, &g));
            });
        } else {
            rc = btc_graph_from_coo(job.src.data(), job.dst.data(), (int64_t)job.src.size(), job.n, cf, &g);
        }
        std::vector<int32_t>().swap(job.src);
        std::vector<int32_t>().swap(job.dst);
        if (rc != BTC_OK) return rc;
        btc_graph_info gi;
        btc_graph_get_info(g, &gi);
        const int64_t est = 24 * gi.m + 8 * (int64_t)gi.n + (1 << 16);  // lists + items, before the plan exists
        if (!admit(est)) {
            btc_graph_free(g);
            return BTC_ERR_STATE;
        }
        btc_plan p = nullptr;
        rc = btc_plan_create(g, &co, s, &p);
        {
            std::lock_guard<std::mutex> lk(mu);
            resident_bytes -= est;
        }
        cv_prep.notify_all();  // the reservation is gone: admissions waiting for room re-check
        if (rc != BTC_OK) {
            btc_graph_free(g);
            return rc;
        }
        btc_plan_info pi;
        btc_plan_get_info(p, &pi);
        out = std::make_shared<Unit>();
        out->graph = g;
        out->plan = p;
        out->bytes = pi.device_bytes + 4 * (gi.m + gi.n + 1);
        out->keys = {job.key};
        return BTC_OK;
    }

    // Captured preparation of one graph that fits a pipeline: the count
    // answers the waiting requests at once; the exported L and plan become
    // the resident unit.  Returns false when the graph is too large for a
    // pipeline (nothing done then).
    bool prepare_captured(PrepJob& job, uint32_t cf, const btc_count_opts& co, hipStream_t s, PipeCache& cache_p,
                          std::shared_ptr<Unit>& out, int& rc) {
        const int64_t nnz = (int64_t)job.src.size();
        // with n given the pipeline checks the ids itself (on the device)
        const int64_t nv = job.n > 0 ? (int64_t)job.n : job_vertices(job);
        if (nnz > kPipeMaxPairs || nv > kPipeMaxIds) return false;
        if (nv < 0) {
            rc = BTC_ERR_ARG;
            return true;
        }
        btc_pipeline pipe = nullptr;
        rc = cache_p.get((int32_t)pow2_at_least(nv, 1024), pow2_at_least(nnz, 4096), 1, cf & BTC_CANON_DEFAULT, co, s,
                         &pipe);
        if (rc != BTC_OK) return true;
        const double t_start = now();
        uint64_t tri = 0;
        rc = btc_pipeline_count(pipe, job.src.data(), job.dst.data(), nnz, job.n, 0, s, &tri, nullptr);
        std::vector<int32_t>().swap(job.src);
        std::vector<int32_t>().swap(job.dst);
        if (rc != BTC_OK) return true;
        {
            // the requests waiting so far get their answer now
            std::lock_guard<std::mutex> lk(mu);
            Entry& e = cache[job.key];
            auto w = waiting.find(job.key);
            if (w != waiting.end()) {
                for (Request& r : w->second) {
                    --e.pins;
                    r.t_start = t_start;
                    finish(r, BTC_OK, tri, 1);
                    ++stats.launched_requests;
                }
                waiting.erase(w);
            }
            ++stats.launches;
            cv_done.notify_all();
        }
        btc_pipeline_info pi;
        btc_pipeline_get_info(pipe, &pi);
        const int64_t est = 24 * pi.m + 36 * (int64_t)pi.work_items + 12 * (int64_t)pi.n + (1 << 16);
        if (!admit(est)) {
            rc = BTC_ERR_STATE;
            return true;
        }
        btc_graph g = nullptr;
        btc_plan p = nullptr;
        rc = btc_pipeline_export(pipe, s, &g, &p);
        {
            std::lock_guard<std::mutex> lk(mu);
            resident_bytes -= est;
        }
        cv_prep.notify_all();
        if (rc != BTC_OK) return true;
        btc_plan_info info;
        btc_plan_get_info(p, &info);
        btc_graph_info gi;
        btc_graph_get_info(g, &gi);
        out = std::make_shared<Unit>();
        out->graph = g;
        out->plan = p;
        out->bytes = info.device_bytes + 4 * (gi.m + 2 * (int64_t)gi.n + 1);
        out->keys = {job.key};
        return true;
    }

    // Captured preparation of several checked graphs (ids within nv[k]) in
    // one pipeline run: the per-graph counts answer their waiting requests at
    // once; the exported batch becomes the resident unit.
    int prepare_captured_group(std::vector<PrepJob>& jobs, const std::vector<int64_t>& nv, uint32_t cf,
                               const btc_count_opts& co, hipStream_t s, PipeCache& cache_p,
                               std::shared_ptr<Unit>& out) {
        const int32_t k = (int32_t)jobs.size();
        std::vector<const int32_t*> sp((size_t)k), dp((size_t)k);
        std::vector<int64_t> sizes((size_t)k);
        int64_t nsum = 0, pairs = 0;
        for (int32_t g = 0; g < k; ++g) {
            sp[(size_t)g] = jobs[(size_t)g].src.data();
            dp[(size_t)g] = jobs[(size_t)g].dst.data();
            sizes[(size_t)g] = (int64_t)jobs[(size_t)g].src.size();
            nsum += nv[(size_t)g];
            pairs += sizes[(size_t)g];
        }
        btc_pipeline pipe = nullptr;
        int rc = cache_p.get((int32_t)pow2_at_least(nsum, 1 << 14), pow2_at_least(pairs, 1 << 16), opts.group_graphs,
                             cf & BTC_CANON_DEFAULT, co, s, &pipe);
        if (rc != BTC_OK) return rc;
        const double t_start = now();
        std::vector<uint64_t> per((size_t)k);
        uint64_t tot = 0;
        // ids were checked against nv when the group was formed
        rc = btcx::guard([&] {
            btcx::pipeline_count_checked(pipe, k, sp.data(), dp.data(), sizes.data(), nv.data(), s, per.data(), &tot);
        });
        for (PrepJob& j : jobs) {
            std::vector<int32_t>().swap(j.src);
            std::vector<int32_t>().swap(j.dst);
        }
        if (rc != BTC_OK) return rc;
        {
            std::lock_guard<std::mutex> lk(mu);
            for (int32_t g = 0; g < k; ++g) {
                Entry& e = cache[jobs[(size_t)g].key];
                auto w = waiting.find(jobs[(size_t)g].key);
                if (w == waiting.end()) continue;
                for (Request& r : w->second) {
                    --e.pins;
                    r.t_start = t_start;
                    finish(r, BTC_OK, per[(size_t)g], k);
                    ++stats.launched_requests;
                }
                waiting.erase(w);
            }
            ++stats.launches;
            cv_done.notify_all();
        }
        btc_pipeline_info pi;
        btc_pipeline_get_info(pipe, &pi);
        const int64_t est = 24 * pi.m + 36 * (int64_t)pi.work_items + 20 * (int64_t)pi.n + (1 << 16);
        if (!admit(est)) return BTC_ERR_STATE;
        btc_batch b = nullptr;
        rc = btc_pipeline_export_batch(pipe, s, &b);
        if (rc == BTC_OK) btcx::batch_set_side_streams(b, false);
        {
            std::lock_guard<std::mutex> lk(mu);
            resident_bytes -= est;
        }
        cv_prep.notify_all();
        if (rc != BTC_OK) return rc;
        btc_batch_info bi;
        btc_batch_get_info(b, &bi);
        out = std::make_shared<Unit>();
        out->batch = b;
        out->bytes = bi.plan.device_bytes + 4 * (bi.m + 2 * (int64_t)bi.n + 1) + 8 * bi.graphs;
        for (const PrepJob& j : jobs) out->keys.push_back(j.key);
        return BTC_OK;
    }

    // One GPU canonicalisation and one block-diagonal plan for several graphs.
    int prepare_group(std::vector<PrepJob>& jobs, const std::vector<int64_t>& nv, uint32_t cf,
                      const btc_count_opts& co, hipStream_t s, btcx::PinnedBuf& pin, std::shared_ptr<Unit>& out) {
        std::vector<int64_t> segs(jobs.size() + 1, 0), csegs;
        size_t pairs = 0;
        for (size_t k = 0; k < jobs.size(); ++k) {
            segs[k + 1] = segs[k] + nv[k];
            pairs += jobs[k].src.size();
        }
        btc_graph_s G;
        int rc = btcx::guard([&] {
            BTC_REQUIRE(segs.back() < INT32_MAX, BTC_ERR_OVERFLOW, "group too large");
            // the shifted pairs go straight into pinned memory
            int32_t* src = reinterpret_cast<int32_t*>(pin.reserve(std::max<size_t>(pairs, 1) * 8));
            int32_t* dst = src + pairs;
            size_t o = 0;
            for (size_t k = 0; k < jobs.size(); ++k) {
                const int32_t v0 = (int32_t)segs[k];
                for (size_t e = 0; e < jobs[k].src.size(); ++e, ++o) {
                    src[o] = jobs[k].src[e] + v0;
                    dst[o] = jobs[k].dst[e] + v0;
                }
                std::vector<int32_t>().swap(jobs[k].src);
                std::vector<int32_t>().swap(jobs[k].dst);
            }
            btcx::TmpBuf<int32_t> ds, dd;
            ds.alloc(pairs, s);
            dd.alloc(pairs, s);
            if (pairs) {
                BTC_HIP(hipMemcpyAsync(ds.p, src, pairs * 4, hipMemcpyHostToDevice, s));
                BTC_HIP(hipMemcpyAsync(dd.p, dst, pairs * 4, hipMemcpyHostToDevice, s));
            }
            btcx::canonicalise_device(ds.p, dd.p, (int64_t)pairs, (int32_t)segs.back(), cf & ~BTC_CANON_DEVICE, s,
                                      &G, &segs, &csegs);
        });
        if (rc != BTC_OK) return rc;
        const int64_t est = 24 * G.m + 8 * (int64_t)G.n + (1 << 16);
        if (!admit(est)) return BTC_ERR_STATE;
        btc_batch b = nullptr;
        rc = btcx::guard([&] {
            b = btcx::batch_from_graph(std::move(G), csegs, &co, s);
            btcx::batch_set_side_streams(b, false);  // no stream creation in the worker's count
        });
        {
            std::lock_guard<std::mutex> lk(mu);
            resident_bytes -= est;
        }
        cv_prep.notify_all();  // the reservation is gone: admissions waiting for room re-check
        if (rc != BTC_OK) return rc;
        btc_batch_info bi;
        btc_batch_get_info(b, &bi);
        out = std::make_shared<Unit>();
        out->batch = b;
        out->bytes = bi.plan.device_bytes + 4 * (bi.m + 2 * (int64_t)bi.n + 1) + 8 * bi.graphs;
        for (const PrepJob& j : jobs) out->keys.push_back(j.key);
        return BTC_OK;
    }

    // Waits until `est` more bytes fit next to the pinned units and reserves
    // them; false when the server stops meanwhile.
    bool admit(int64_t est) {
        std::unique_lock<std::mutex> lk(mu);
        cv_prep.wait(lk, [&] { return stop || make_room(est) || resident_bytes == 0; });
        if (stop) return false;
        resident_bytes += est;  // reserved until the unit's real size is known
        return true;
    }

    void prep_loop() {
        // prep threads already run side by side: no nested OpenMP teams
        omp_set_num_threads(1);
        (void)hipSetDevice(device);
        hipStream_t s = nullptr;
        (void)hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
        btc_count_opts co;
        btc_count_opts_default(&co);
        const uint32_t cf = opts.canon_flags ? (uint32_t)opts.canon_flags : BTC_CANON_DEFAULT | BTC_CANON_DEVICE;
        {
            // device arena of the thread's temporary buffers and pinned staging
            // of the pairs, released before their stream
            btcx::Arena arena;
            btcx::PinnedBuf pin;
            PipeCache pipes;
            const bool captured = opts.captured_prep != 0 && (cf & BTC_CANON_DEVICE);
            for (;;) {
                std::vector<PrepJob> jobs;
                {
                    std::unique_lock<std::mutex> lk(mu);
                    cv_prep.wait(lk, [&] { return stop || !prep_queue.empty(); });
                    if (stop) break;
                    size_t pairs = 0;
                    // graphs are grouped only when canonicalised on the GPU; with
                    // captured preparation only under a backlog (more than
                    // kBacklogPerThread queued graphs per prep thread): a pipeline
                    // answers one graph soonest, a group prepares many graphs
                    // fastest
                    const bool backlog = prep_queue.size() > kBacklogPerThread * (size_t)opts.prep_threads;
                    const size_t cap =
                        (cf & BTC_CANON_DEVICE) && (!captured || backlog) ? (size_t)opts.group_graphs : 1;
                    while (!prep_queue.empty() && jobs.size() < cap &&
                           (jobs.empty() || pairs + prep_queue.front().src.size() <= kGroupPairs)) {
                        pairs += prep_queue.front().src.size();
                        jobs.push_back(std::move(prep_queue.front()));
                        prep_queue.pop_front();
                    }
                }
                std::vector<std::shared_ptr<Unit>> units;
                std::vector<std::pair<uint64_t, int>> failed;
                std::unique_ptr<btcx::ArenaScope> scope(new btcx::ArenaScope(arena, s));
                if (jobs.size() == 1) {
                    std::shared_ptr<Unit> u;
                    int rc = BTC_OK;
                    if (!captured || !prepare_captured(jobs[0], cf, co, s, pipes, u, rc))
                        rc = prepare_single(jobs[0], cf, co, s, pin, u);
                    if (rc == BTC_OK) units.push_back(u);
                    else failed.emplace_back(jobs[0].key, rc);
                } else {
                    // bad graphs fail alone; the rest form one unit
                    std::vector<PrepJob> good;
                    std::vector<int64_t> nv;
                    for (PrepJob& j : jobs) {
                        const int64_t n = job_vertices(j);
                        if (n < 0) {
                            failed.emplace_back(j.key, BTC_ERR_ARG);
                            continue;
                        }
                        nv.push_back(n);
                        good.push_back(std::move(j));
                    }
                    int64_t nsum = 0;
                    for (int64_t n : nv) nsum += n;
                    if (good.size() == 1 || nsum >= kGroupVertices) {
                        // too few or too many vertices for one unit: one graph at a time
                        for (PrepJob& j : good) {
                            std::shared_ptr<Unit> u;
                            const int rc = prepare_single(j, cf, co, s, pin, u);
                            if (rc == BTC_OK) units.push_back(u);
                            else failed.emplace_back(j.key, rc);
                        }
                    } else if (!good.empty()) {
                        std::vector<uint64_t> keys;
                        for (const PrepJob& j : good) keys.push_back(j.key);
                        std::shared_ptr<Unit> u;
                        int64_t pairs = 0;
                        for (const PrepJob& j : good) pairs += (int64_t)j.src.size();
                        const int rc = captured && pairs <= kPipeMaxPairs
                                           ? prepare_captured_group(good, nv, cf, co, s, pipes, u)
                                           : prepare_group(good, nv, cf, co, s, pin, u);
                        if (rc == BTC_OK) units.push_back(u);
                        else
                            for (uint64_t k : keys) failed.emplace_back(k, rc);
                    }
                }
                scope.reset();
                std::lock_guard<std::mutex> lk(mu);
                for (auto& u : units) publish(u);
                for (auto& f : failed) fail_key(f.first, f.second);
            }
        }
        if (s) (void)hipStreamDestroy(s);
    }

    void gpu_loop() {
        (void)hipSetDevice(device);
        hipStream_t s = nullptr;
        (void)hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
        btcx::MultiWorkspace ws;
        std::vector<Request> batch;
        std::vector<std::shared_ptr<Unit>> units;   // distinct units of the launch
        std::vector<int> unit_of;                   // request -> units index
        std::vector<btc_plan_s*> plans;
        std::vector<uint64_t> totals;
        std::vector<std::vector<uint64_t>> per_graph;
        for (;;) {
            batch.clear();
            units.clear();
            unit_of.clear();
            {
                std::unique_lock<std::mutex> lk(mu);
                cv_gpu.wait(lk, [&] { return stop || !ready.empty(); });
                if (stop) break;
                const double t = now();
                while (!ready.empty() && (int)batch.size() < opts.max_batch) {
                    batch.push_back(ready.front());
                    ready.pop_front();
                    batch.back().t_start = t;
                }
                for (const Request& r : batch) {
                    const std::shared_ptr<Unit>& u = cache[r.key].unit;
                    size_t k = 0;
                    while (k < units.size() && units[k] != u) ++k;
                    if (k == units.size()) units.push_back(u);
                    unit_of.push_back((int)k);
                }
            }
            // single-graph units: one multi-plan launch; groups: one launch each
            plans.clear();
            std::vector<int> plan_slot(units.size(), -1);
            for (size_t k = 0; k < units.size(); ++k)
                if (units[k]->plan) {
                    plan_slot[k] = (int)plans.size();
                    plans.push_back(units[k]->plan);
                }
            totals.assign(plans.size(), 0);
            per_graph.assign(units.size(), {});
            int rc = BTC_OK;
            int64_t launches = 0;
            std::vector<int> plan_rc(plans.size(), BTC_OK);
            if (!plans.empty()) {
                rc = btcx::guard([&] { btcx::plans_count(plans.data(), (int)plans.size(), s, totals.data(), ws); });
                ++launches;
                if (rc != BTC_OK && plans.size() > 1) {
                    // one plan's failure must not fail the others: count each alone
                    for (size_t j = 0; j < plans.size(); ++j) {
                        plan_rc[j] = btcx::guard([&] { btcx::plans_count(&plans[j], 1, s, &totals[j], ws); });
                        ++launches;
                    }
                } else {
                    std::fill(plan_rc.begin(), plan_rc.end(), rc);
                }
            }
            std::vector<int> unit_rc(units.size(), BTC_OK);
            for (size_t k = 0; k < units.size(); ++k)
                if (plan_slot[k] >= 0) unit_rc[k] = plan_rc[(size_t)plan_slot[k]];
            for (size_t k = 0; k < units.size(); ++k) {
                if (!units[k]->batch) continue;
                per_graph[k].assign(units[k]->keys.size(), 0);
                uint64_t t = 0;
                unit_rc[k] = btc_batch_count(units[k]->batch, s, per_graph[k].data(), &t, nullptr);
                ++launches;
            }
            std::lock_guard<std::mutex> lk(mu);
            stats.launches += launches;
            stats.launched_requests += (int64_t)batch.size();
            for (size_t i = 0; i < batch.size(); ++i) {
                const Request& r = batch[i];
                const int k = unit_of[i];
                Unit* u = units[(size_t)k].get();
                Entry& e = cache[r.key];
                --e.pins;
                if (u->in_lru) lru.splice(lru.begin(), lru, u->lru);
                const int st = unit_rc[(size_t)k];
                uint64_t tri = 0;
                if (st == BTC_OK) tri = u->plan ? totals[(size_t)plan_slot[(size_t)k]] : per_graph[(size_t)k][(size_t)e.idx];
                finish(r, st, tri, (int)batch.size());
            }
            cv_done.notify_all();
            cv_prep.notify_all();  // unpinned units may now be evicted for waiting admissions
        }
        if (s) (void)hipStreamDestroy(s);
    }
};

extern "C" {

void btc_server_opts_default(btc_server_opts* o) {
    if (!o) return;
    std::memset(o, 0, sizeof(*o));
    o->hbm_budget_bytes = (int64_t)64 << 30;
    o->max_batch = 512;
    // 4 x 16 measured best at R-MAT serving (burst saturation and p99) once
    // preparation uses per-thread arenas: 3 x 8 1.52 K graphs/s, p99 37 ms;
    // 4 x 8 1.80 K, p99 48 ms; 3 x 16 1.46 K, 20 ms; 4 x 16 1.61 K, 25 ms
    o->prep_threads = 4;
    o->canon_flags = BTC_CANON_DEFAULT | BTC_CANON_DEVICE;
    o->group_graphs = 16;
    o->captured_prep = 1;
}

int btc_server_create(const btc_server_opts* opts, btc_server* out) {
    return btcx::guard([&] {
        BTC_REQUIRE(out, BTC_ERR_ARG, "NULL argument");
        *out = nullptr;
        auto sv = std::make_unique<btc_server_s>();
        btc_server_opts_default(&sv->opts);
        if (opts) sv->opts = *opts;
        BTC_REQUIRE(sv->opts.hbm_budget_bytes > 0 && sv->opts.max_batch > 0 && sv->opts.prep_threads > 0,
                    BTC_ERR_ARG, "server options must be positive");
        BTC_REQUIRE(sv->opts.group_graphs >= 1 && sv->opts.group_graphs <= kGroupGraphsMax, BTC_ERR_ARG,
                    "group_graphs %d (expected 1..%d)", sv->opts.group_graphs, kGroupGraphsMax);
        BTC_HIP(hipGetDevice(&sv->device));
        btc_server_s* raw = sv.get();
        for (int i = 0; i < sv->opts.prep_threads; ++i) sv->workers.emplace_back([raw] { raw->prep_loop(); });
        sv->workers.emplace_back([raw] { raw->gpu_loop(); });
        *out = sv.release();
    });
}

int btc_server_submit(btc_server sv, uint64_t graph_key, const int32_t* src, const int32_t* dst, int64_t nnz,
                      int32_t n, uint64_t* request_id) {
    return btcx::guard([&] {
        BTC_REQUIRE(sv && request_id, BTC_ERR_ARG, "NULL argument");
        BTC_REQUIRE(nnz >= 0 && (nnz == 0 || (src && dst)) , BTC_ERR_ARG, "bad edge arrays");
        // the pairs are copied before the lock is taken: preparation threads
        // and the GPU worker keep running meanwhile
        PrepJob job;
        job.key = graph_key;
        job.src.assign(src, src + nnz);
        job.dst.assign(dst, dst + nnz);
        job.n = n;
        std::lock_guard<std::mutex> lk(sv->mu);
        BTC_REQUIRE(!sv->stop, BTC_ERR_STATE, "server is stopping");
        Request r;
        r.id = sv->next_id++;
        r.key = graph_key;
        r.t_submit = sv->now();
        *request_id = r.id;
        ++sv->stats.submitted;
        auto it = sv->cache.find(graph_key);
        if (it == sv->cache.end()) {
            BTC_REQUIRE(nnz > 0 || n > 0, BTC_ERR_ARG, "graph %llu is not resident and no graph was given",
                        (unsigned long long)graph_key);
            Entry& e = sv->cache[graph_key];
            e.pins = 1;
            sv->prep_queue.push_back(std::move(job));
            sv->waiting[graph_key].push_back(r);
            // prep threads and admission waiters share cv_prep: wake them all so
            // an idle prep thread sees the job
            sv->cv_prep.notify_all();
            return;
        }
        Entry& e = it->second;
        r.resident_hit = true;
        ++sv->stats.resident_hits;
        if (e.state == 2) {
            sv->finish(r, e.status, 0, 0);
            sv->cv_done.notify_all();
            return;
        }
        ++e.pins;
        if (e.state == 0) {
            sv->waiting[graph_key].push_back(r);
        } else {
            sv->ready.push_back(r);
            sv->cv_gpu.notify_one();
        }
    });
}

int btc_server_poll(btc_server sv, btc_result* out, int32_t capacity, int32_t wait_ms, int32_t* count) {
    return btcx::guard([&] {
        BTC_REQUIRE(sv && out && count && capacity > 0, BTC_ERR_ARG, "bad argument");
        std::unique_lock<std::mutex> lk(sv->mu);
        if (sv->done.empty() && wait_ms > 0)
            sv->cv_done.wait_for(lk, std::chrono::milliseconds(wait_ms), [&] { return !sv->done.empty(); });
        int32_t k = 0;
        while (k < capacity && !sv->done.empty()) {
            out[k++] = sv->done.front();
            sv->done.pop_front();
        }
        *count = k;
    });
}

int btc_server_get_stats(btc_server sv, btc_server_stats* st) {
    return btcx::guard([&] {
        BTC_REQUIRE(sv && st, BTC_ERR_ARG, "NULL argument");
        std::lock_guard<std::mutex> lk(sv->mu);
        *st = sv->stats;
        st->resident_bytes = sv->resident_bytes;
        st->resident_graphs = 0;
        for (const Unit* u : sv->lru) st->resident_graphs += (int64_t)u->keys.size();
        st->clock = sv->now();
    });
}

void btc_server_free(btc_server sv) {
    if (!sv) return;
    {
        std::lock_guard<std::mutex> lk(sv->mu);
        sv->stop = true;
    }
    sv->cv_prep.notify_all();
    sv->cv_gpu.notify_all();
    for (auto& t : sv->workers) t.join();
    delete sv;  // cache entries release their units
}

}  // extern "C"
