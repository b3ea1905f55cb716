// Serving: a request loop with continuous batching.
//
//   submit -> [prep threads] canonicalise + plan on their own HIP streams
//             (admission: the resident set stays under the HBM budget; idle
//             graphs are evicted least recently used first)
//          -> [GPU worker] every request whose plan is resident joins the
//             next launch: one persistent wedge-kernel launch over all of
//             their plans (plus dense-tile launches) on the worker's stream
//          -> completion queue (btc_server_poll)
//
// Requests for a graph key that is resident (or being prepared) reuse it.
// While a launch is on the GPU, new arrivals are prepared concurrently on
// the prep streams and join the next launch.  Every count is exact and
// independent of what it is batched with: plans are per graph and each plan
// has its own total.  No reference counterpart (the reference is a one-shot
// CLI).
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

struct Entry {  // one graph known to the server
    btc_graph graph = nullptr;
    btc_plan plan = nullptr;
    int64_t bytes = 0;
    int state = 0;  // 0 preparing, 1 resident, 2 failed
    int status = BTC_OK;
    int pins = 0;   // requests holding the entry (waiting, ready or launched)
    std::list<uint64_t>::iterator lru;
    bool in_lru = false;
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
    std::list<uint64_t> lru;  // front = most recently used resident graph
    int64_t resident_bytes = 0;
    std::deque<PrepJob> prep_queue;
    std::unordered_map<uint64_t, std::vector<Request>> waiting;  // key -> requests waiting for its plan
    std::deque<Request> ready;
    std::deque<btc_result> done;
    std::vector<std::thread> workers;
    btc_server_stats stats{};

    double now() const { return std::chrono::duration<double>(Clock::now() - t0).count(); }

    // Caller holds mu.  Evicts idle resident graphs (LRU first) until `need`
    // more bytes fit; returns whether they fit.
    bool make_room(int64_t need) {
        auto it = lru.end();
        while (resident_bytes + need > opts.hbm_budget_bytes && it != lru.begin()) {
            --it;
            Entry& e = cache[*it];
            if (e.pins > 0) continue;
            resident_bytes -= e.bytes;
            btc_plan_free(e.plan);
            btc_graph_free(e.graph);
            const uint64_t key = *it;
            it = lru.erase(it);
            cache.erase(key);
            ++stats.evictions;
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

    void prep_loop() {
        // prep threads already run side by side: no nested OpenMP teams
        omp_set_num_threads(1);
        (void)hipSetDevice(device);
        hipStream_t s = nullptr;
        (void)hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
        btc_count_opts co;
        btc_count_opts_default(&co);
        for (;;) {
            PrepJob job;
            {
                std::unique_lock<std::mutex> lk(mu);
                cv_prep.wait(lk, [&] { return stop || !prep_queue.empty(); });
                if (stop) break;
                job = std::move(prep_queue.front());
                prep_queue.pop_front();
            }
            btc_graph g = nullptr;
            btc_plan p = nullptr;
            const uint32_t cf = opts.canon_flags ? (uint32_t)opts.canon_flags : BTC_CANON_DEFAULT | BTC_CANON_DEVICE;
            int rc = BTC_OK;
            if (cf & BTC_CANON_DEVICE) {
                // canonicalise (and then plan) on the GPU, on this thread's stream
                rc = btcx::guard([&] {
                    const size_t nnz = job.src.size();
                    btcx::DevBuf<int32_t> ds, dd;
                    ds.alloc(nnz);
                    dd.alloc(nnz);
                    if (nnz) {
                        BTC_HIP(hipMemcpyAsync(ds.p, job.src.data(), nnz * 4, hipMemcpyHostToDevice, s));
                        BTC_HIP(hipMemcpyAsync(dd.p, job.dst.data(), nnz * 4, hipMemcpyHostToDevice, s));
                    }
                    btcx::check_status(btc_graph_from_coo_device(ds.p, dd.p, (int64_t)nnz, job.n, cf & ~BTC_CANON_DEVICE,
                                                           s, &g));
                });
            } else {
                rc = btc_graph_from_coo(job.src.data(), job.dst.data(), (int64_t)job.src.size(), job.n, cf, &g);
            }
            std::vector<int32_t>().swap(job.src);
            std::vector<int32_t>().swap(job.dst);
            int64_t est = 0;
            if (rc == BTC_OK) {
                btc_graph_info gi;
                btc_graph_get_info(g, &gi);
                est = 24 * gi.m + 8 * (int64_t)gi.n + (1 << 16);  // CSR + lists + items, before the plan exists
                std::unique_lock<std::mutex> lk(mu);
                // admission: wait until the graph fits next to the pinned ones
                cv_prep.wait(lk, [&] { return stop || make_room(est) || resident_bytes == 0; });
                if (stop) {
                    btc_graph_free(g);
                    break;
                }
                resident_bytes += est;  // reserved until the plan's real size is known
            }
            if (rc == BTC_OK) rc = btc_plan_create(g, &co, s, &p);
            std::lock_guard<std::mutex> lk(mu);
            Entry& e = cache[job.key];
            if (est) resident_bytes -= est;
            if (rc == BTC_OK) {
                btc_plan_info pi;
                btc_plan_get_info(p, &pi);
                btc_graph_info gi;
                btc_graph_get_info(g, &gi);
                e.graph = g;
                e.plan = p;
                e.bytes = pi.device_bytes + 4 * (gi.m + gi.n + 1);
                e.state = 1;
                resident_bytes += e.bytes;
                lru.push_front(job.key);
                e.lru = lru.begin();
                e.in_lru = true;
                ++stats.prepared;
                for (Request& r : waiting[job.key]) ready.push_back(r);
                waiting.erase(job.key);
                cv_gpu.notify_one();
            } else {
                if (g) btc_graph_free(g);
                e.state = 2;
                e.status = rc;
                for (Request& r : waiting[job.key]) {
                    --e.pins;
                    finish(r, rc, 0, 0);
                }
                waiting.erase(job.key);
                cv_done.notify_all();
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
        std::vector<btc_plan_s*> plans;
        std::vector<uint64_t> keys, totals;
        for (;;) {
            batch.clear();
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
                // one plan per distinct graph in the launch
                plans.clear();
                keys.clear();
                for (const Request& r : batch) {
                    if (std::find(keys.begin(), keys.end(), r.key) != keys.end()) continue;
                    keys.push_back(r.key);
                    plans.push_back(cache[r.key].plan);
                }
            }
            totals.assign(plans.size(), 0);
            int rc = btcx::guard([&] { btcx::plans_count(plans.data(), (int)plans.size(), s, totals.data(), ws); });
            std::lock_guard<std::mutex> lk(mu);
            ++stats.launches;
            stats.launched_requests += (int64_t)batch.size();
            for (const Request& r : batch) {
                const size_t k = (size_t)(std::find(keys.begin(), keys.end(), r.key) - keys.begin());
                Entry& e = cache[r.key];
                --e.pins;
                if (e.in_lru) lru.splice(lru.begin(), lru, e.lru);
                finish(r, rc, rc == BTC_OK ? totals[k] : 0, (int)batch.size());
            }
            cv_done.notify_all();
            cv_prep.notify_all();  // unpinned graphs may now be evicted for waiting admissions
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
    o->prep_threads = 8;
    o->canon_flags = BTC_CANON_DEFAULT | BTC_CANON_DEVICE;
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
            PrepJob job;
            job.key = graph_key;
            job.src.assign(src, src + nnz);
            job.dst.assign(dst, dst + nnz);
            job.n = n;
            sv->prep_queue.push_back(std::move(job));
            sv->waiting[graph_key].push_back(r);
            sv->cv_prep.notify_one();
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
        st->resident_graphs = (int64_t)sv->lru.size();
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
    for (auto& kv : sv->cache) {
        if (kv.second.plan) btc_plan_free(kv.second.plan);
        if (kv.second.graph) btc_graph_free(kv.second.graph);
    }
    delete sv;
}

}  // extern "C"
