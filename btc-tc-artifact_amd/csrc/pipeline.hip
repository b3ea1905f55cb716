// btc_pipeline: COO pairs -> triangle total as one captured HIP graph.
//
// The graph holds the device canonicalisation (canonicalise_capacity), the
// wedge-kernel planner (plan_capacity) and the count
// (pipeline_count_enqueue), each built from kernels that read their live
// sizes -- pairs, nnz(L), vertices, lists, items -- from a control block in
// HBM (PipeCtl) that earlier kernels of the same graph fill.  Every buffer
// is carved at creation from one allocation sized for n_cap ids and nnz_cap
// pairs, so a run is: copy the pairs and two header words in, one
// hipGraphLaunch, copy the control block and the total out.  The steps are
// those of the reference's preprocess_for_triangle_counting
// (btc/operations/convert.h:601-692) followed by planning and counting; the
// reference runs them from the host for every graph.
#include <algorithm>
#include <cstring>
#include <mutex>

#include "internal.h"
#include "prim.h"

struct btc_pipeline_s {
    int32_t n_cap = 0;
    int64_t nnz_cap = 0;
    int32_t max_graphs = 1;  // graphs one run may hold (vertex ranges of one L)
    uint32_t flags = 0;
    btc_count_opts opts{};
    btcx::DevBuf<unsigned char> mem;
    btcx::PipeCtl* ctl = nullptr;
    unsigned long long* work = nullptr;
    int32_t* src = nullptr;
    int32_t* dst = nullptr;
    uint16_t* col16 = nullptr;
    int32_t* labels = nullptr;  // the canonicalisation's label buffer
    // several graphs: input ranges (host-written), ranges in L, per-vertex
    // and per-graph sums
    uint32_t* segs = nullptr;
    uint32_t* csegs = nullptr;
    unsigned long long* pv = nullptr;
    unsigned long long* per_graph = nullptr;
    btcx::CanonCap canon{};
    btcx::PlanCap plan{};
    btcx::CapturedGraph graph;
    int32_t graph_nodes = 0;
    btcx::PinnedBuf pin;  // [0, 4 KiB): control block and total; then staged pairs
    btcx::PipeCtl last{};
    size_t ws_bytes = 0;  // the sorts' and scans' workspace (zeroed at creation)
    // a run's end is awaited on this event, which blocks the host thread
    // instead of spinning (threads that run pipelines side by side leave
    // their cores to each other meanwhile)
    hipEvent_t done = nullptr;
    ~btc_pipeline_s() {
        if (done) (void)hipEventDestroy(done);
    }
    bool ran = false;     // `last` holds an accepted run
    std::mutex mu;        // one run at a time
};

namespace btcx {
namespace {

// Pinned staging: control block, total and error word from 0, the graphs'
// ranges at kPinSegs, per-graph sums at kPinPerGraph, pairs from kPinHeader.
constexpr size_t kPinHeader = 4096;
constexpr size_t kPinSegs = 1024, kPinPerGraph = 2048;
constexpr int kMaxPipelineGraphs = 64;
static_assert(sizeof(PipeCtl) + 16 <= kPinSegs && kPinSegs + 4 * (kMaxPipelineGraphs + 1) <= kPinPerGraph &&
                  kPinPerGraph + 8 * kMaxPipelineGraphs <= kPinHeader,
              "pinned header layout");

// Bump layout over one allocation (base NULL: sizing pass).
struct Carve {
    unsigned char* base = nullptr;
    size_t top = 0;
    template <class T>
    T* take(size_t count) {
        const size_t at = top;
        top += (std::max<size_t>(count, 1) * sizeof(T) + 255) & ~size_t(255);
        return base ? reinterpret_cast<T*>(base + at) : nullptr;
    }
};

int id_bits(int64_t n) {
    int b = 1;
    while (b < 31 && (1ll << b) < n) ++b;
    return b;
}

void layout(btc_pipeline_s* p, Carve& c) {
    const size_t E = (size_t)p->nnz_cap, N = (size_t)p->n_cap, M = E;
    const uint32_t C = p->opts.core_ids == 0 ? kCoreIdsDefault : p->opts.core_ids < 0 ? 0u : (uint32_t)p->opts.core_ids;
    p->ctl = c.take<PipeCtl>(1);
    p->work = c.take<unsigned long long>(kWorkWords);
    p->src = c.take<int32_t>(E);
    p->dst = c.take<int32_t>(E);
    CanonCap& k = p->canon;
    k.src = p->src;
    k.dst = p->dst;
    k.nnz_cap = (int64_t)E;
    k.n_cap = p->n_cap;
    k.flags = p->flags;
    k.ctl = p->ctl;
    k.ka = c.take<uint64_t>(E);
    k.kb = c.take<uint64_t>(E);
    k.flag = c.take<uint32_t>(E);
    k.pos = c.take<uint32_t>(E + 1);
    k.used = c.take<uint32_t>(N);
    k.newid = c.take<uint32_t>(N + 1);
    k.deg = c.take<uint32_t>(N);
    k.k0 = c.take<uint32_t>(N);
    k.v0 = c.take<uint32_t>(N);
    k.k1 = c.take<uint32_t>(N);
    k.v1 = c.take<uint32_t>(N);
    k.rank = c.take<uint32_t>(N);
    k.lab0 = c.take<int32_t>(N);
    k.lab1 = c.take<int32_t>(N);
    k.rowptr = c.take<uint32_t>(N + 1);
    k.col = c.take<int32_t>(E + kColPad);
    p->col16 = c.take<uint16_t>(E + kColPad);
    k.col16 = p->col16;
    PlanCap& q = p->plan;
    q.rowptr = k.rowptr;
    q.col = k.col;
    q.m_cap = (int64_t)M;
    q.n_cap = p->n_cap;
    q.core_ids = C;
    q.cus = device_cu_count();
    q.ctl = p->ctl;
    // the canonicalisation's keys and flags are dead once L is written:
    // the planner's list keys and flags reuse them
    q.k0 = reinterpret_cast<uint32_t*>(k.ka);
    q.v0 = q.k0 ? q.k0 + M : nullptr;
    q.k1 = reinterpret_cast<uint32_t*>(k.kb);
    q.v1 = q.k1 ? q.k1 + M : nullptr;
    q.lflag = k.flag;
    q.lpos = k.pos;
    q.erow = c.take<uint32_t>(M);
    q.off = c.take<uint32_t>(N + 1);
    q.cend = c.take<uint32_t>(N);
    q.has = c.take<uint32_t>(N);
    q.rbi = c.take<uint32_t>(N + 1);
    q.lists = c.take<uint4>(M);
    q.rowbits = c.take<uint32_t>(C ? pipe_bitmap_rows((int64_t)M, p->n_cap) * (C / 32) : 1);
    q.ur = c.take<int32_t>(M);
    q.cost = c.take<uint64_t>(M);
    q.P = c.take<uint64_t>(M + 1);
    q.start = c.take<uint32_t>(M);
    q.spos = c.take<uint32_t>(M + 1);
    q.items_cap = M;
    q.first = c.take<uint32_t>(M);
    q.tmp_items = c.take<uint4>(2 * M);
    q.sk0 = c.take<uint32_t>(M);
    q.sv0 = c.take<uint32_t>(M);
    q.sk1 = c.take<uint32_t>(M);
    q.sv1 = c.take<uint32_t>(M);
    q.items = c.take<uint4>(2 * M);
    q.work = p->work;
    const size_t big = std::max(E, N);
    p->ws_bytes = dprim_ws_bytes(big);
    const DevWs ws{c.take<unsigned char>(p->ws_bytes), big};
    k.ws = ws;
    q.ws = ws;
    if (p->max_graphs > 1) {
        const size_t G = (size_t)p->max_graphs;
        p->segs = c.take<uint32_t>(G + 1);
        p->csegs = c.take<uint32_t>(G + 1);
        p->per_graph = c.take<unsigned long long>(G);
        p->pv = c.take<unsigned long long>(N);
        k.max_graphs = p->max_graphs;
        k.segs = p->segs;
        k.csegs = p->csegs;
    }
}

// Carves the buffers and captures the graph.
void pipeline_build(btc_pipeline_s* p, hipStream_t s) {
    Carve sizing;
    layout(p, sizing);
    p->mem.alloc(sizing.top);
    Carve real;
    real.base = p->mem.p;
    layout(p, real);
    // the count's words and the sorts' and scans' workspace start cleared;
    // every replay leaves them cleared again
    BTC_HIP(hipMemsetAsync(p->work, 0, kWorkWords * 8, s));
    BTC_HIP(hipMemsetAsync(p->canon.ws.p, 0, p->ws_bytes, s));
    BTC_HIP(hipStreamSynchronize(s));
    p->graph.capture([&](hipStream_t cs) {
        p->labels = canonicalise_capacity(p->canon, cs);
        plan_capacity(p->plan, cs);
        pipeline_count_enqueue(p->plan, p->col16, p->opts, cs, p->pv, p->csegs, p->per_graph, p->max_graphs);
    });
    BTC_HIP(hipEventCreateWithFlags(&p->done, hipEventBlockingSync | hipEventDisableTiming));
    size_t nodes = 0;
    BTC_HIP(hipGraphGetNodes(p->graph.graph, nullptr, &nodes));
    p->graph_nodes = (int32_t)nodes;
    p->pin.reserve(kPinHeader);
}

// One replay, the pairs already in p->src / p->dst (and, for several graphs,
// their ranges at pin + kPinSegs): header words in, control block, total,
// error word and per-graph sums out.  Caller holds p->mu.
void replay(btc_pipeline_s* p, char* pin, uint32_t nnz, int32_t n_in, uint32_t graphs, hipStream_t s,
            uint64_t* total, uint64_t* per_graph, float* kernel_ms) {
    PipeCtl* h = reinterpret_cast<PipeCtl*>(pin);
    std::memset(h, 0, sizeof(PipeCtl));
    h->nnz = nnz;
    h->n_in = n_in;
    h->graphs = graphs;
    BTC_HIP(hipMemcpyAsync(p->ctl, h, kPipeHostWords * 4, hipMemcpyHostToDevice, s));
    if (p->max_graphs > 1)
        BTC_HIP(hipMemcpyAsync(p->segs, pin + kPinSegs, ((size_t)graphs + 1) * 4, hipMemcpyHostToDevice, s));
    EventTimer tm(kernel_ms != nullptr);
    tm.start(s);
    p->graph.launch(s);
    tm.stop(s);
    unsigned long long* htotal = reinterpret_cast<unsigned long long*>(pin + sizeof(PipeCtl));
    uint32_t* herr = reinterpret_cast<uint32_t*>(htotal + 1);
    unsigned long long* hpg = reinterpret_cast<unsigned long long*>(pin + kPinPerGraph);
    BTC_HIP(hipMemcpyAsync(h, p->ctl, sizeof(PipeCtl), hipMemcpyDeviceToHost, s));
    BTC_HIP(hipMemcpyAsync(htotal, p->work, 8, hipMemcpyDeviceToHost, s));
    BTC_HIP(hipMemcpyAsync(herr, dsort_error_word(p->canon.ws), 4, hipMemcpyDeviceToHost, s));
    if (p->max_graphs > 1)
        BTC_HIP(hipMemcpyAsync(hpg, p->per_graph, (size_t)graphs * 8, hipMemcpyDeviceToHost, s));
    BTC_HIP(hipEventRecord(p->done, s));
    BTC_HIP(hipEventSynchronize(p->done));
    BTC_REQUIRE(*herr == 0, BTC_ERR_HIP, "a device sort of the pipeline failed (look-back spin limit)");
    const PipeCtl c = *h;
    BTC_REQUIRE(!(c.status & kPipeNegativeId), BTC_ERR_ARG, "negative vertex id %d", c.lo);
    BTC_REQUIRE(!(c.status & kPipeIdRange), BTC_ERR_ARG, "vertex id %d >= n = %d", c.hi, n_in);
    BTC_REQUIRE(!(c.status & kPipeTooManyIds), BTC_ERR_ARG, "vertex id %d: more than the pipeline's %d ids", c.hi,
                p->n_cap);
    if (kernel_ms) *kernel_ms = tm.ms();
    *total = *htotal;
    if (per_graph) {
        if (p->max_graphs > 1) std::memcpy(per_graph, hpg, (size_t)graphs * 8);
        else per_graph[0] = *htotal;
    }
    p->last = c;
    p->ran = true;
}

void pipeline_run_batch(btc_pipeline_s* p, const int32_t* src, const int32_t* dst, const int64_t* eoff,
                        const int32_t* npg, int32_t count, hipStream_t s, uint64_t* per_graph, uint64_t* total,
                        float* kernel_ms);

void pipeline_run(btc_pipeline_s* p, const int32_t* src, const int32_t* dst, int64_t nnz, int32_t n,
                  bool device_pairs, hipStream_t s, uint64_t* total, float* kernel_ms) {
    BTC_REQUIRE(nnz >= 0 && nnz <= p->nnz_cap, BTC_ERR_ARG, "%lld pairs (pipeline capacity %lld)", (long long)nnz,
                (long long)p->nnz_cap);
    BTC_REQUIRE(n <= p->n_cap, BTC_ERR_ARG, "n = %d exceeds the pipeline's %d ids", n, p->n_cap);
    BTC_REQUIRE(nnz == 0 || (src && dst), BTC_ERR_ARG, "src/dst is NULL");
    if (p->max_graphs > 1) {  // one graph as a run of one range (ids checked on the host)
        BTC_REQUIRE(!device_pairs, BTC_ERR_ARG, "a pipeline for several graphs takes host pairs");
        const int64_t eoff[2] = {0, nnz};
        pipeline_run_batch(p, src, dst, eoff, &n, 1, s, nullptr, total, kernel_ms);
        return;
    }
    std::lock_guard<std::mutex> lk(p->mu);
    p->ran = false;
    const size_t pair_bytes = (size_t)nnz * 4;
    char* pin = p->pin.reserve(kPinHeader + (device_pairs ? 0 : 2 * pair_bytes));
    if (nnz) {
        if (device_pairs) {
            BTC_HIP(hipMemcpyAsync(p->src, src, pair_bytes, hipMemcpyDeviceToDevice, s));
            BTC_HIP(hipMemcpyAsync(p->dst, dst, pair_bytes, hipMemcpyDeviceToDevice, s));
        } else {
            std::memcpy(pin + kPinHeader, src, pair_bytes);
            std::memcpy(pin + kPinHeader + pair_bytes, dst, pair_bytes);
            BTC_HIP(hipMemcpyAsync(p->src, pin + kPinHeader, pair_bytes, hipMemcpyHostToDevice, s));
            BTC_HIP(hipMemcpyAsync(p->dst, pin + kPinHeader + pair_bytes, pair_bytes, hipMemcpyHostToDevice, s));
        }
    }
    replay(p, pin, (uint32_t)nnz, n, 1, s, total, nullptr, kernel_ms);
}

// Several graphs in one run, their ids already checked: graph g's pairs are
// src[g] / dst[g][0, sizes[g]) over ids [0, nv[g]); they are shifted into
// one id space while they are staged.
void pipeline_run_checked(btc_pipeline_s* p, int32_t count, const int32_t* const* src, const int32_t* const* dst,
                          const int64_t* sizes, const int64_t* nv, hipStream_t s, uint64_t* per_graph,
                          uint64_t* total, float* kernel_ms) {
    BTC_REQUIRE(count >= 1 && count <= p->max_graphs && p->max_graphs > 1, BTC_ERR_ARG,
                "%d graphs (the pipeline takes %d)", count, p->max_graphs);
    std::vector<int64_t> eoff((size_t)count + 1, 0), segs((size_t)count + 1, 0);
    for (int32_t g = 0; g < count; ++g) {
        eoff[(size_t)g + 1] = eoff[(size_t)g] + sizes[g];
        segs[(size_t)g + 1] = segs[(size_t)g] + nv[g];
    }
    const int64_t pairs = eoff[(size_t)count];
    BTC_REQUIRE(pairs <= p->nnz_cap, BTC_ERR_ARG, "%lld pairs (pipeline capacity %lld)", (long long)pairs,
                (long long)p->nnz_cap);
    BTC_REQUIRE(segs[(size_t)count] <= p->n_cap, BTC_ERR_ARG, "%lld ids in all (the pipeline takes %d)",
                (long long)segs[(size_t)count], p->n_cap);
    std::lock_guard<std::mutex> lk(p->mu);
    p->ran = false;
    const size_t pair_bytes = (size_t)pairs * 4;
    char* pin = p->pin.reserve(kPinHeader + 2 * pair_bytes);
    int32_t* hs = reinterpret_cast<int32_t*>(pin + kPinHeader);
    int32_t* hd = hs + pairs;
#pragma omp parallel for schedule(dynamic, 1)
    for (int32_t g = 0; g < count; ++g) {
        const int32_t v0 = (int32_t)segs[(size_t)g];
        const int64_t e0 = eoff[(size_t)g];
        for (int64_t e = 0; e < sizes[g]; ++e) {
            hs[e0 + e] = src[g][e] + v0;
            hd[e0 + e] = dst[g][e] + v0;
        }
    }
    uint32_t* hsegs = reinterpret_cast<uint32_t*>(pin + kPinSegs);
    for (int32_t g = 0; g <= count; ++g) hsegs[g] = (uint32_t)segs[(size_t)g];
    if (pairs) {
        BTC_HIP(hipMemcpyAsync(p->src, hs, pair_bytes, hipMemcpyHostToDevice, s));
        BTC_HIP(hipMemcpyAsync(p->dst, hd, pair_bytes, hipMemcpyHostToDevice, s));
    }
    replay(p, pin, (uint32_t)pairs, (int32_t)segs[(size_t)count], (uint32_t)count, s, total, per_graph, kernel_ms);
}

// Several graphs in one run: graph g's pairs are src/dst[eoff[g], eoff[g +
// 1]) over ids [0, npg[g]) (<= 0: its largest id + 1), checked here.
void pipeline_run_batch(btc_pipeline_s* p, const int32_t* src, const int32_t* dst, const int64_t* eoff,
                        const int32_t* npg, int32_t count, hipStream_t s, uint64_t* per_graph, uint64_t* total,
                        float* kernel_ms) {
    BTC_REQUIRE(eoff && npg && count >= 1, BTC_ERR_ARG, "bad argument");
    BTC_REQUIRE(count <= p->max_graphs, BTC_ERR_ARG, "%d graphs (the pipeline takes %d)", count, p->max_graphs);
    BTC_REQUIRE(eoff[0] == 0, BTC_ERR_ARG, "edge_offsets[0] != 0");
    const int64_t pairs = eoff[count];
    BTC_REQUIRE(pairs >= 0 && pairs <= p->nnz_cap, BTC_ERR_ARG, "%lld pairs (pipeline capacity %lld)",
                (long long)pairs, (long long)p->nnz_cap);
    BTC_REQUIRE(pairs == 0 || (src && dst), BTC_ERR_ARG, "src/dst is NULL");
    if (p->max_graphs == 1) {
        uint64_t t = 0;
        pipeline_run(p, src, dst, pairs, npg[0], false, s, &t, kernel_ms);
        *total = t;
        if (per_graph) per_graph[0] = t;
        return;
    }
    std::vector<const int32_t*> sp((size_t)count), dp((size_t)count);
    std::vector<int64_t> sizes((size_t)count), nv((size_t)count);
    for (int32_t g = 0; g < count; ++g) {
        const int64_t e0 = eoff[g], e1 = eoff[g + 1];
        BTC_REQUIRE(e0 <= e1, BTC_ERR_ARG, "edge_offsets not ascending at graph %d", g);
        int32_t lo = 0, hi = -1;
        for (int64_t e = e0; e < e1; ++e) {
            lo = std::min(lo, std::min(src[e], dst[e]));
            hi = std::max(hi, std::max(src[e], dst[e]));
        }
        const int64_t n = npg[g] > 0 ? npg[g] : (int64_t)hi + 1;
        BTC_REQUIRE(lo >= 0 && hi < n, BTC_ERR_ARG, "graph %d: vertex id out of [0, %lld)", g, (long long)n);
        sp[(size_t)g] = src + e0;
        dp[(size_t)g] = dst + e0;
        sizes[(size_t)g] = e1 - e0;
        nv[(size_t)g] = n;
    }
    pipeline_run_checked(p, count, sp.data(), dp.data(), sizes.data(), nv.data(), s, per_graph, total, kernel_ms);
}

// Copies of the last accepted run: its graph (with_plan false) or its graph
// and wedge-kernel plan, carved from one allocation that the graph owns (the
// plan's arrays are views into it, so the plan must go first).  The plan
// counts like any other and later runs of the pipeline do not touch it.
void pipeline_copy_locked(btc_pipeline_s* p, hipStream_t s, bool with_plan, btc_graph_s** gout,
                          btc_plan_s** pout) {
    BTC_REQUIRE(p->ran, BTC_ERR_STATE, "no accepted run yet");
    const PipeCtl& c = p->last;
    const PlanCap& pc = p->plan;
    const size_t n = c.nc, m = c.m, ne = c.ne, ni = c.nitems, CW = pc.core_ids / 32;
    const size_t lb = ne * sizeof(uint4), bb = (size_t)c.nbm * CW * 4, rb = ne * 4;
    auto g = std::make_unique<btc_graph_s>();
    std::unique_ptr<btc_plan_s> q;
    if (with_plan) q = std::make_unique<btc_plan_s>();
    auto carve = [&](Carve& cv) {
        uint32_t* rowptr = cv.take<uint32_t>(n + 1);
        int32_t* col = cv.take<int32_t>(m + kColPad);
        int32_t* labels = cv.take<int32_t>(std::max<size_t>(n, 1));
        if (cv.base) {
            g->d_rowptr.view(rowptr, n + 1);
            g->d_col.view(col, m + kColPad);
            g->d_labels.view(labels, std::max<size_t>(n, 1));
        }
        if (!with_plan) return;
        unsigned char* L = cv.take<unsigned char>(lb + bb + rb);
        uint4* items = cv.take<uint4>(std::max<size_t>(2 * ni, 1));
        uint16_t* col16 = cv.take<uint16_t>(m + kColPad);
        unsigned long long* scratch = cv.take<unsigned long long>(kWorkWords);
        if (cv.base) {
            q->lists.view(L, lb + bb + rb);
            q->items_dev.view(items, std::max<size_t>(2 * ni, 1));
            q->col16.view(col16, m + kColPad);
            q->scratch.view(scratch, kWorkWords);
        }
    };
    Carve sizing;
    carve(sizing);
    g->block.alloc(sizing.top);
    Carve real;
    real.base = g->block.p;
    carve(real);
    g->n = (int32_t)n;
    g->n_input = (int32_t)c.n;
    g->m = (int64_t)m;
    g->flags = p->flags;
    g->max_row = (int32_t)c.rs[0];
    g->wedges = (int64_t)c.rs[1];
    g->merge_work = (int64_t)c.rs[2];
    g->on_host = false;
    BTC_HIP(hipMemcpyAsync(g->d_rowptr.p, p->canon.rowptr, (n + 1) * 4, hipMemcpyDeviceToDevice, s));
    BTC_HIP(hipMemcpyAsync(g->d_col.p, p->canon.col, (m + kColPad) * 4, hipMemcpyDeviceToDevice, s));
    if (n) BTC_HIP(hipMemcpyAsync(g->d_labels.p, p->labels, n * 4, hipMemcpyDeviceToDevice, s));
    if (with_plan) {
        unsigned char* L = q->lists.p;
        if (lb) BTC_HIP(hipMemcpyAsync(L, pc.lists, lb, hipMemcpyDeviceToDevice, s));
        if (bb) BTC_HIP(hipMemcpyAsync(L + lb, pc.rowbits, bb, hipMemcpyDeviceToDevice, s));
        if (rb) BTC_HIP(hipMemcpyAsync(L + lb + bb, pc.ur, rb, hipMemcpyDeviceToDevice, s));
        if (ni) BTC_HIP(hipMemcpyAsync(q->items_dev.p, pc.items, 2 * ni * sizeof(uint4), hipMemcpyDeviceToDevice, s));
        BTC_HIP(hipMemcpyAsync(q->col16.p, p->col16, (m + kColPad) * 2, hipMemcpyDeviceToDevice, s));
        BTC_HIP(hipMemsetAsync(q->scratch.p, 0, q->scratch.bytes(), s));
    }
    BTC_HIP(hipEventRecord(p->done, s));
    BTC_HIP(hipEventSynchronize(p->done));
    if (with_plan) {
        q->g = g.get();
        q->K = 64;
        q->tau = 16 * q->K;  // no dense tiles
        q->core_ids = pc.core_ids;
        q->use_tiny = false;
        q->launch = p->opts;
        q->list_p = reinterpret_cast<const uint4*>(q->lists.p);
        q->rowbits_p = reinterpret_cast<const uint4*>(q->lists.p + lb);
        q->urow_p = reinterpret_cast<const int32_t*>(q->lists.p + lb + bb);
        q->items_p = q->items_dev.p;
        q->nitems = (uint32_t)ni;
        q->n_block_items = c.item_counts[0];
        q->n_tiny_items = 0;
        btc_plan_info& in = q->info;
        in.sparse_edges = (int64_t)c.st[1];
        in.sparse_merge_work = (int64_t)c.st[2];
        in.core_lists = (int64_t)c.st[3];
        in.core_rows = c.nbm;
        in.wedge_checks = (int64_t)c.checks;
        in.work_items = (int64_t)ni;
        in.tile_k = q->K;
        in.tau = q->tau;
        in.core_ids = pc.core_ids;
        in.stream_bytes = 4 * ((int64_t)c.checks + (int64_t)c.st[4]) - 2 * (int64_t)c.st[6] + 16 * (int64_t)ne +
                          32 * (int64_t)ni;
        in.device_bytes = (int64_t)(q->lists.bytes() + q->items_dev.bytes() + q->col16.bytes() + q->scratch.bytes());
        *pout = q.release();
    }
    *gout = g.release();
}

btc_graph_s* pipeline_graph(btc_pipeline_s* p, hipStream_t s) {
    std::lock_guard<std::mutex> lk(p->mu);
    btc_graph_s* g = nullptr;
    pipeline_copy_locked(p, s, false, &g, nullptr);
    return g;
}

void pipeline_export(btc_pipeline_s* p, hipStream_t s, btc_graph_s** gout, btc_plan_s** pout) {
    std::lock_guard<std::mutex> lk(p->mu);
    pipeline_copy_locked(p, s, true, gout, pout);
}

// The last run of several graphs as a batch (its L, plan and the graphs'
// vertex ranges in L).
btc_batch_s* pipeline_export_batch(btc_pipeline_s* p, hipStream_t s) {
    std::lock_guard<std::mutex> lk(p->mu);
    BTC_REQUIRE(p->max_graphs > 1, BTC_ERR_STATE, "a pipeline of single graphs (btc_pipeline_export)");
    BTC_REQUIRE(p->ran, BTC_ERR_STATE, "no accepted run yet");
    const uint32_t k = p->last.graphs;
    std::vector<uint32_t> cs((size_t)k + 1);
    BTC_HIP(hipMemcpyAsync(cs.data(), p->csegs, cs.size() * 4, hipMemcpyDeviceToHost, s));
    BTC_HIP(hipStreamSynchronize(s));
    btc_graph_s* g = nullptr;
    btc_plan_s* q = nullptr;
    pipeline_copy_locked(p, s, true, &g, &q);
    std::unique_ptr<btc_graph_s> own_g(g);
    const std::vector<int64_t> voff(cs.begin(), cs.end());
    return batch_from_plan(std::move(*own_g), q, voff, s);
}

}  // namespace

void pipeline_count_checked(btc_pipeline_s* p, int32_t count, const int32_t* const* src, const int32_t* const* dst,
                            const int64_t* sizes, const int64_t* nv, hipStream_t s, uint64_t* per_graph,
                            uint64_t* total) {
    pipeline_run_checked(p, count, src, dst, sizes, nv, s, per_graph, total, nullptr);
}

}  // namespace btcx

extern "C" {

int btc_pipeline_create_batch(int32_t n_cap, int64_t nnz_cap, int32_t max_graphs, uint32_t flags,
                              const btc_count_opts* opts, void* stream, btc_pipeline* out) {
    return btcx::guard([&] {
        BTC_REQUIRE(out, BTC_ERR_ARG, "out is NULL");
        *out = nullptr;
        BTC_REQUIRE(n_cap >= 1, BTC_ERR_ARG, "n_cap %d < 1", n_cap);
        BTC_REQUIRE(nnz_cap >= 1 && nnz_cap + (int64_t)btcx::kColPad < (int64_t)0xFFFFFFFFLL, BTC_ERR_ARG,
                    "nnz_cap %lld (expected 1 .. 2^32 - %zu)", (long long)nnz_cap, btcx::kColPad + 1);
        BTC_REQUIRE(max_graphs >= 1 && max_graphs <= btcx::kMaxPipelineGraphs, BTC_ERR_ARG,
                    "max_graphs %d (expected 1 .. %d)", max_graphs, btcx::kMaxPipelineGraphs);
        BTC_REQUIRE((flags & ~(BTC_CANON_DEFAULT | BTC_CANON_DEVICE)) == 0, BTC_ERR_ARG, "unknown flags 0x%x",
                    flags);
        if (max_graphs > 1 && (flags & BTC_CANON_DEGREE_ORDER)) {
            int b = 1, gb = 1;
            while (b < 31 && (1ll << b) < n_cap) ++b;
            while ((1 << gb) < max_graphs) ++gb;
            BTC_REQUIRE(b + gb <= 32, BTC_ERR_OVERFLOW, "%d ids and %d graphs exceed 32-bit relabel keys", n_cap,
                        max_graphs);
        }
        auto p = std::make_unique<btc_pipeline_s>();
        p->n_cap = n_cap;
        p->nnz_cap = nnz_cap;
        p->max_graphs = max_graphs;
        p->flags = flags & BTC_CANON_DEFAULT;
        btc_count_opts_default(&p->opts);
        if (opts) p->opts = *opts;
        BTC_REQUIRE(p->opts.core_ids <= 0 ||
                        (p->opts.core_ids % 128 == 0 && (uint32_t)p->opts.core_ids <= btcx::kCoreIdsMax),
                    BTC_ERR_ARG, "core_ids %d (expected a multiple of 128 up to %u, 0 = default, < 0 = none)",
                    p->opts.core_ids, btcx::kCoreIdsMax);
        btcx::check_launch_opts(p->opts);
        btcx::pipeline_build(p.get(), btcx::as_stream(stream));
        *out = p.release();
    });
}

int btc_pipeline_create(int32_t n_cap, int64_t nnz_cap, uint32_t flags, const btc_count_opts* opts, void* stream,
                        btc_pipeline* out) {
    return btc_pipeline_create_batch(n_cap, nnz_cap, 1, flags, opts, stream, out);
}

int btc_pipeline_count(btc_pipeline p, const int32_t* src, const int32_t* dst, int64_t nnz, int32_t n,
                       int32_t device_pairs, void* stream, uint64_t* total, float* kernel_ms) {
    return btcx::guard([&] {
        BTC_REQUIRE(p && total, BTC_ERR_ARG, "NULL argument");
        btcx::pipeline_run(p, src, dst, nnz, n, device_pairs != 0, btcx::as_stream(stream), total, kernel_ms);
    });
}

int btc_pipeline_count_batch(btc_pipeline p, const int32_t* src, const int32_t* dst, const int64_t* edge_offsets,
                             const int32_t* n_per_graph, int32_t count, void* stream, uint64_t* per_graph,
                             uint64_t* total, float* kernel_ms) {
    return btcx::guard([&] {
        BTC_REQUIRE(p && total, BTC_ERR_ARG, "NULL argument");
        btcx::pipeline_run_batch(p, src, dst, edge_offsets, n_per_graph, count, btcx::as_stream(stream), per_graph,
                                 total, kernel_ms);
    });
}

int btc_pipeline_export_batch(btc_pipeline p, void* stream, btc_batch* out) {
    return btcx::guard([&] {
        BTC_REQUIRE(p && out, BTC_ERR_ARG, "NULL argument");
        *out = nullptr;
        *out = btcx::pipeline_export_batch(p, btcx::as_stream(stream));
    });
}

int btc_pipeline_graph(btc_pipeline p, void* stream, btc_graph* out) {
    return btcx::guard([&] {
        BTC_REQUIRE(p && out, BTC_ERR_ARG, "NULL argument");
        *out = nullptr;
        *out = btcx::pipeline_graph(p, btcx::as_stream(stream));
    });
}

int btc_pipeline_export(btc_pipeline p, void* stream, btc_graph* graph, btc_plan* plan) {
    return btcx::guard([&] {
        BTC_REQUIRE(p && graph && plan, BTC_ERR_ARG, "NULL argument");
        *graph = nullptr;
        *plan = nullptr;
        btcx::pipeline_export(p, btcx::as_stream(stream), graph, plan);
    });
}

int btc_pipeline_get_info(btc_pipeline p, btc_pipeline_info* info) {
    return btcx::guard([&] {
        BTC_REQUIRE(p && info, BTC_ERR_ARG, "NULL argument");
        std::lock_guard<std::mutex> lk(p->mu);
        std::memset(info, 0, sizeof(*info));
        info->n_cap = p->n_cap;
        info->nnz_cap = p->nnz_cap;
        info->max_graphs = p->max_graphs;
        info->flags = p->flags;
        info->graph_nodes = p->graph_nodes;
        info->device_bytes = (int64_t)p->mem.bytes();
        if (p->ran) {
            const btcx::PipeCtl& c = p->last;
            info->n = (int32_t)c.nc;
            info->m = c.m;
            info->max_row = (int32_t)c.rs[0];
            info->wedges = (int64_t)c.rs[1];
            info->wedge_checks = (int64_t)c.checks;
            info->work_items = (int32_t)c.nitems;
        }
    });
}

void btc_pipeline_free(btc_pipeline p) { delete p; }

}  // extern "C"
