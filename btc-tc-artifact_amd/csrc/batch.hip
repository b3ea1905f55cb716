// Batched counting: many graphs in one persistent launch.
//
// The canonical L of every graph is placed on the diagonal of one
// block-diagonal L (graph g's vertices start at a multiple of 128, so no
// 16 x K tile, K <= 128, holds rows or columns of two graphs), one plan is
// built for it, and one enqueue of the dense-tile and wedge kernels counts
// all graphs; the wedge kernel's static stride over work-sorted items
// balances graphs of very different sizes.  Every wedge and tiny-pivot work
// item belongs to one graph (its pivot's), so those kernels add each item's
// count to its graph's sum (the item record carries the graph index); the
// dense-tile kernel keeps per-vertex sums, and a per-graph total is its
// graph's item sum plus the segment sum of its vertices' dense counts, one
// workgroup per graph.  No reference counterpart (the reference counts one
// graph per process).
#include <algorithm>
#include <cstring>

#include "device_util.h"
#include "internal.h"

namespace btcx {
constexpr int64_t kGraphAlign = 128;
}

struct btc_batch_s {
    int32_t count = 0;
    std::vector<int64_t> voff;  // count + 1 vertex offsets (multiples of 128)
    btc_graph_s graph;          // the block-diagonal L
    btc_plan plan = nullptr;
    btcx::DevBuf<uint32_t> d_voff;
    btcx::PvBuf d_pv;  // per-vertex sums (64-bit slots when a row is longer than 65536)
    btcx::DevBuf<unsigned long long> d_pg;
    btcx::DevBuf<unsigned long long> d_gsum;  // item sums per graph (zero between counts)
    std::vector<uint64_t> h_pg;
    btcx::CapturedGraph count_graph;  // zero + count + per-graph sums, replayed when captured
    // counts may put the dense-tile and tiny-pivot kernels on side streams
    // of the plan (created on first use); a server's units count on the
    // worker's stream alone
    bool side_streams = true;
    ~btc_batch_s() {
        if (plan) btc_plan_free(plan);
    }
    // dense-tile + wedge kernels with per-vertex sums, then per-graph sums
    void enqueue_count(hipStream_t s);
};

namespace btcx {
namespace {

// per_graph[g] = gsum[g] + sum of pv over [voff[g], voff[g + 1]) (padding
// vertices are 0); gsum[g] is cleared for the next count
__global__ __launch_bounds__(256) void graph_sums_kernel(PvOut pv, const uint32_t* __restrict__ voff,
                                                         unsigned long long* __restrict__ gsum,
                                                         unsigned long long* __restrict__ per_graph) {
    __shared__ unsigned long long part[4];
    const uint32_t g = blockIdx.x;
    uint64_t s = 0;
    for (uint32_t v = voff[g] + threadIdx.x; v < voff[g + 1]; v += 256)
        s += pv.p64 ? (uint64_t)pv.p64[v] : (uint64_t)(uint32_t)pv.p32[v];
    s = wave_sum_u64(s);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        per_graph[g] = gsum[g] + part[0] + part[1] + part[2] + part[3];
        gsum[g] = 0;
    }
}

// lseg[g] = first position of graph g's rows in L (g = 0..count)
__global__ void graph_starts_kernel(const uint32_t* __restrict__ rowptr, const uint32_t* __restrict__ voff,
                                    uint32_t count, uint32_t* __restrict__ lseg) {
    for (uint32_t g = blockIdx.x * blockDim.x + threadIdx.x; g <= count; g += gridDim.x * blockDim.x)
        lseg[g] = rowptr[voff[g]];
}

// Writes each item's graph (the one whose rows hold its pivot row's start
// t0: the last g with lseg[g] <= t0) into the item's B.w.
__global__ void item_graphs_kernel(uint4* __restrict__ items, uint32_t nitems, const uint32_t* __restrict__ lseg,
                                   uint32_t count) {
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < nitems; k += gridDim.x * blockDim.x) {
        const uint32_t t0 = items[2 * k].x;
        uint32_t lo = 0, hi = count;  // lseg[lo] <= t0 < lseg[hi] when lseg[count] > t0
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (lseg[mid] <= t0) lo = mid;
            else hi = mid;
        }
        items[2 * k + 1].w = lo;
    }
}

}  // namespace

// Per-graph item sums for a batch's plan: graph indices into its items and
// the zeroed sums the wedge and tiny kernels add to.
void batch_item_graphs(btc_batch_s* b, hipStream_t s) {
    btc_plan_s* p = b->plan;
    b->d_gsum.alloc((size_t)b->count);
    BTC_HIP(hipMemsetAsync(b->d_gsum.p, 0, (size_t)b->count * 8, s));
    if (p->nitems) {
        TmpBuf<uint32_t> lseg;
        lseg.alloc((size_t)b->count + 1, s);
        const unsigned g1 = (unsigned)std::min<int64_t>(((int64_t)b->count + 256) / 256, 1024);
        hipLaunchKernelGGL(graph_starts_kernel, dim3(g1), dim3(256), 0, s, b->graph.d_rowptr.p, b->d_voff.p,
                           (uint32_t)b->count, lseg.p);
        BTC_CHECK_LAUNCH("graph_starts_kernel");
        const unsigned g2 = (unsigned)std::min<int64_t>(((int64_t)p->nitems + 255) / 256, 4096);
        hipLaunchKernelGGL(item_graphs_kernel, dim3(g2), dim3(256), 0, s, const_cast<uint4*>(p->items_p),
                           (uint32_t)p->nitems, lseg.p, (uint32_t)b->count);
        BTC_CHECK_LAUNCH("item_graphs_kernel");
    }
    p->item_graph_sums = b->d_gsum.p;
}

// Best of three timed batch counts on s (after one untimed), in ms.
float batch_count_ms(btc_batch_s* b, hipStream_t s) {
    float best = 1e30f;
    for (int r = 0; r < 4; ++r) {
        EventTimer tm(true);
        launch_zero_u32(reinterpret_cast<uint32_t*>(b->d_pv.data()), b->d_pv.bytes() / 4, s);
        tm.start(s);
        b->enqueue_count(s);
        tm.stop(s);
        const float ms = tm.ms();
        if (r) best = std::min(best, ms);
    }
    return best;
}

}  // namespace btcx

namespace btcx {

// A batch's plan options: opts (or the defaults) with tiny_items 0 taken as
// 2, since the pivots of small graphs hold a few short lists each and a
// wave then waits on one pivot's round trips where a tiny-item group of 8
// lanes keeps 8 pivots in flight (4096-graph workload: 6.60 -> 6.23 ms).
btc_count_opts batch_opts(const btc_count_opts* opts) {
    btc_count_opts o;
    if (opts) o = *opts;
    else btc_count_opts_default(&o);
    if (o.tiny_items == 0) o.tiny_items = 2;
    return o;
}

// Plans the batch's block-diagonal graph and allocates its sums.
void batch_finish(btc_batch_s* b, const btc_count_opts* opts, hipStream_t s) {
    const btc_count_opts o = batch_opts(opts);
    check_status(btc_plan_create(&b->graph, &o, s, &b->plan));
    std::vector<uint32_t> vo32(b->voff.begin(), b->voff.end());
    b->d_voff.alloc(vo32.size());
    b->d_pv.alloc((size_t)b->graph.n, b->graph.max_row);
    b->d_pg.alloc((size_t)b->count);
    b->h_pg.assign((size_t)b->count, 0);
    BTC_HIP(hipMemcpyAsync(b->d_voff.p, vo32.data(), vo32.size() * 4, hipMemcpyHostToDevice, s));
    batch_item_graphs(b, s);
    BTC_HIP(hipStreamSynchronize(s));
}

btc_batch_s* batch_from_graph(btc_graph_s&& g, const std::vector<int64_t>& voff, const btc_count_opts* opts,
                              hipStream_t s) {
    BTC_REQUIRE(voff.size() >= 2 && voff.front() == 0 && voff.back() == g.n, BTC_ERR_ARG, "bad graph ranges");
    auto b = std::make_unique<btc_batch_s>();
    b->count = (int32_t)voff.size() - 1;
    b->voff = voff;
    b->graph = std::move(g);
    batch_finish(b.get(), opts, s);
    return b.release();
}

btc_batch_s* batch_from_plan(btc_graph_s&& g, btc_plan_s* plan, const std::vector<int64_t>& voff, hipStream_t s) {
    std::unique_ptr<btc_plan_s> own(plan);
    BTC_REQUIRE(voff.size() >= 2 && voff.front() == 0 && voff.back() == g.n, BTC_ERR_ARG, "bad graph ranges");
    auto b = std::make_unique<btc_batch_s>();
    b->count = (int32_t)voff.size() - 1;
    b->voff = voff;
    b->graph = std::move(g);
    b->plan = own.release();
    b->plan->g = &b->graph;
    std::vector<uint32_t> vo32(voff.begin(), voff.end());
    b->d_voff.alloc(vo32.size());
    b->d_pv.alloc((size_t)b->graph.n, b->graph.max_row);
    b->d_pg.alloc((size_t)b->count);
    b->h_pg.assign((size_t)b->count, 0);
    BTC_HIP(hipMemcpyAsync(b->d_voff.p, vo32.data(), vo32.size() * 4, hipMemcpyHostToDevice, s));
    batch_item_graphs(b.get(), s);
    BTC_HIP(hipStreamSynchronize(s));
    return b.release();
}

void batch_set_side_streams(btc_batch_s* b, bool on) { b->side_streams = on; }

}  // namespace btcx

void btc_batch_s::enqueue_count(hipStream_t s) {
    btcx::plan_enqueue(plan, s, d_pv.out(), nullptr, nullptr, nullptr, side_streams);
    hipLaunchKernelGGL(btcx::graph_sums_kernel, dim3((unsigned)count), dim3(256), 0, s, d_pv.out(), d_voff.p,
                       d_gsum.p, d_pg.p);
    BTC_CHECK_LAUNCH("graph_sums_kernel");
}

extern "C" {

int btc_batch_create(const btc_graph* graphs, int32_t count, const btc_count_opts* opts, void* stream,
                     btc_batch* out) {
    return btcx::guard([&] {
        BTC_REQUIRE(out && (graphs || count == 0), BTC_ERR_ARG, "NULL argument");
        *out = nullptr;
        BTC_REQUIRE(count > 0, BTC_ERR_ARG, "a batch needs at least one graph");
        auto b = std::make_unique<btc_batch_s>();
        b->count = count;
        b->voff.assign((size_t)count + 1, 0);
        int64_t m = 0;
        for (int32_t k = 0; k < count; ++k) {
            BTC_REQUIRE(graphs[k], BTC_ERR_ARG, "graph %d is NULL", k);
            btcx::ensure_host(graphs[k]);
            const int64_t n = graphs[k]->n;
            b->voff[k + 1] = b->voff[k] + (n + btcx::kGraphAlign - 1) / btcx::kGraphAlign * btcx::kGraphAlign;
            m += graphs[k]->m;
        }
        const int64_t N = b->voff[count];
        BTC_REQUIRE(N < (int64_t)INT32_MAX, BTC_ERR_OVERFLOW, "batch has %lld vertices (limit 2^31 - 1)",
                    (long long)N);
        btc_graph_s& G = b->graph;
        G.n = G.n_input = (int32_t)N;
        G.flags = 0;
        G.rowptr.assign((size_t)N + 1, 0);
        G.col.resize((size_t)m);
        for (int32_t k = 0; k < count; ++k) {
            const btc_graph_s* g = graphs[k];
            const int64_t v0 = b->voff[k];
            for (int32_t i = 0; i < g->n; ++i) G.rowptr[v0 + i + 1] = g->rowptr[i + 1] - g->rowptr[i];
        }
        for (int64_t v = 0; v < N; ++v) G.rowptr[v + 1] += G.rowptr[v];
#pragma omp parallel for schedule(dynamic, 16)
        for (int32_t k = 0; k < count; ++k) {
            const btc_graph_s* g = graphs[k];
            const int64_t v0 = b->voff[k];
            int64_t o = G.rowptr[v0];
            for (int64_t p = 0; p < g->m; ++p) G.col[o + p] = g->col[p] + (int32_t)v0;
        }
        G.labels.resize((size_t)N);
        for (int64_t v = 0; v < N; ++v) G.labels[v] = (int32_t)v;
        btcx::graph_finalize(&G);
        btcx::batch_finish(b.get(), opts, btcx::as_stream(stream));
        *out = b.release();
    });
}

int btc_batch_from_coo(const int32_t* src, const int32_t* dst, const int64_t* edge_offsets,
                       const int32_t* n_per_graph, int32_t count, uint32_t flags, const btc_count_opts* opts,
                       void* stream, btc_batch* out) {
    return btcx::guard([&] {
        BTC_REQUIRE(out && edge_offsets && n_per_graph && count > 0, BTC_ERR_ARG, "bad argument");
        *out = nullptr;
        BTC_REQUIRE((flags & ~(BTC_CANON_DEFAULT | BTC_CANON_DEVICE)) == 0, BTC_ERR_ARG, "unknown flags 0x%x", flags);
        BTC_REQUIRE(edge_offsets[0] == 0, BTC_ERR_ARG, "edge_offsets[0] != 0");
        const int64_t pairs = edge_offsets[count];
        BTC_REQUIRE(pairs == 0 || (src && dst), BTC_ERR_ARG, "src/dst is NULL");
        // vertex ranges of the graphs (n, or max id + 1 when n <= 0); ids checked per graph
        std::vector<int64_t> segs((size_t)count + 1, 0);
        for (int32_t k = 0; k < count; ++k) {
            const int64_t e0 = edge_offsets[k], e1 = edge_offsets[k + 1];
            BTC_REQUIRE(e0 <= e1, BTC_ERR_ARG, "edge_offsets not ascending at graph %d", k);
            int32_t lo = 0, hi = -1;
            for (int64_t e = e0; e < e1; ++e) {
                lo = std::min(lo, std::min(src[e], dst[e]));
                hi = std::max(hi, std::max(src[e], dst[e]));
            }
            const int64_t n = n_per_graph[k] > 0 ? n_per_graph[k] : (int64_t)hi + 1;
            BTC_REQUIRE(lo >= 0 && hi < n, BTC_ERR_ARG, "graph %d: vertex id out of [0, %lld)", k, (long long)n);
            segs[(size_t)k + 1] = segs[(size_t)k] + n;
        }
        BTC_REQUIRE(segs.back() < INT32_MAX, BTC_ERR_OVERFLOW, "batch has %lld vertices (limit 2^31 - 1)",
                    (long long)segs.back());
        hipStream_t s = btcx::as_stream(stream);
        std::vector<int32_t> hs((size_t)pairs), hd((size_t)pairs);
#pragma omp parallel for schedule(dynamic, 16)
        for (int32_t k = 0; k < count; ++k) {
            const int32_t v0 = (int32_t)segs[(size_t)k];
            for (int64_t e = edge_offsets[k]; e < edge_offsets[k + 1]; ++e) {
                hs[(size_t)e] = src[e] + v0;
                hd[(size_t)e] = dst[e] + v0;
            }
        }
        btc_graph_s G;
        std::vector<int64_t> csegs;
        {
            btcx::TmpBuf<int32_t> ds, dd;
            ds.alloc((size_t)pairs, s);
            dd.alloc((size_t)pairs, s);
            if (pairs) {
                const void* h[2] = {hs.data(), hd.data()};
                void* d[2] = {ds.p, dd.p};
                btcx::upload_arrays(h, d, 2, (size_t)pairs * 4, s);
            }
            btcx::canonicalise_device(ds.p, dd.p, pairs, (int32_t)segs.back(), flags & ~BTC_CANON_DEVICE, s, &G,
                                      &segs, &csegs);
        }
        *out = btcx::batch_from_graph(std::move(G), csegs, opts, s);
    });
}

int btc_batch_get_info(btc_batch b, btc_batch_info* info) {
    return btcx::guard([&] {
        BTC_REQUIRE(b && info, BTC_ERR_ARG, "NULL argument");
        info->graphs = b->count;
        info->n = b->graph.n;
        info->m = b->graph.m;
        btcx::check_status(btc_plan_get_info(b->plan, &info->plan));
    });
}

int btc_batch_tune(btc_batch b, const btc_count_opts* opts) {
    return btcx::guard([&] {
        BTC_REQUIRE(b, BTC_ERR_ARG, "NULL handle");
        btcx::check_status(btc_plan_tune(b->plan, opts));
        b->count_graph.reset();
    });
}

int btc_batch_autotune(btc_batch b, const btc_tune_opts* topts, void* stream, btc_count_opts* best,
                       float* best_ms) {
    return btcx::guard([&] {
        BTC_REQUIRE(b && best, BTC_ERR_ARG, "NULL argument");
        btcx::check_status(btc_autotune(&b->graph, topts, stream, best, best_ms));
        // re-plan the batch with the winner (tile width and tau need a new
        // plan); the tuner times counts of the whole block-diagonal graph,
        // so the batch's own counts decide whether the new plan is kept
        hipStream_t s = btcx::as_stream(stream);
        b->count_graph.reset();
        const float before = btcx::batch_count_ms(b, s);
        btc_plan old = b->plan;
        btc_plan p = nullptr;
        const btc_count_opts o = btcx::batch_opts(best);
        btcx::check_status(btc_plan_create(&b->graph, &o, stream, &p));
        b->plan = p;
        btcx::batch_item_graphs(b, s);
        const float after = btcx::batch_count_ms(b, s);
        if (after <= before) {
            btc_plan_free(old);
        } else {  // the plan the batch had counts faster: keep it
            btc_plan_free(p);
            b->plan = old;
            btcx::batch_item_graphs(b, s);
            *best = old->launch;
        }
        BTC_HIP(hipStreamSynchronize(s));
    });
}

int btc_batch_capture(btc_batch b, void* stream) {
    return btcx::guard([&] {
        BTC_REQUIRE(b, BTC_ERR_ARG, "NULL handle");
        (void)stream;
        b->count_graph.capture([&](hipStream_t cs) {
            btcx::launch_zero_u32(reinterpret_cast<uint32_t*>(b->d_pv.data()), b->d_pv.bytes() / 4, cs);
            b->enqueue_count(cs);
        });
    });
}

int btc_batch_count(btc_batch b, void* stream, uint64_t* per_graph, uint64_t* total, float* kernel_ms) {
    return btcx::guard([&] {
        BTC_REQUIRE(b && total, BTC_ERR_ARG, "NULL argument");
        hipStream_t s = btcx::as_stream(stream);
        btcx::EventTimer tm(kernel_ms != nullptr);
        if (b->count_graph.exec) {
            tm.start(s);
            b->count_graph.launch(s);
        } else {
            btcx::launch_zero_u32(reinterpret_cast<uint32_t*>(b->d_pv.data()), b->d_pv.bytes() / 4, s);
            tm.start(s);
            b->enqueue_count(s);
        }
        if (kernel_ms) tm.stop(s);  // both kernels and the per-graph sums
        unsigned long long t = 0;
        BTC_HIP(hipMemcpyAsync(&t, b->plan->scratch.p, 8, hipMemcpyDeviceToHost, s));
        BTC_HIP(hipMemcpyAsync(b->h_pg.data(), b->d_pg.p, (size_t)b->count * 8, hipMemcpyDeviceToHost, s));
        BTC_HIP(hipStreamSynchronize(s));
        if (kernel_ms) *kernel_ms = tm.ms();
        uint64_t sum = 0;
        for (auto x : b->h_pg) sum += x;
        BTC_REQUIRE(sum == t, BTC_ERR_STATE, "per-graph sums %llu != total %llu", (unsigned long long)sum,
                    (unsigned long long)t);
        if (per_graph) std::copy(b->h_pg.begin(), b->h_pg.end(), per_graph);
        *total = t;
    });
}

void btc_batch_free(btc_batch b) { delete b; }

}  // extern "C"
