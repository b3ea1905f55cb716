// Hybrid count engine for L .* (L L^T) on gfx950.
//
// sum over (i, j) in L of |N_L(i) & N_L(j)| is split by edge:
//
//  * edges inside dense 16 x K mask tiles (popcount > tau) go through the
//    Bit-BSR tile path (launch_tile_count with a list of dense tiles, every
//    tile on the dense AND+popcount path);
//
//  * every other edge (i, j) is evaluated as the wedges (i; j, k) with k a
//    neighbour of i below j: k is in N_L(j) exactly when (i, j, k) is a
//    triangle.  Because N_L(i) is sorted, those k are the prefix of row i
//    in front of j, clipped to [min N_L(j), max N_L(j)].  Edges are grouped
//    by pivot j; a workgroup loads N_L(j) once into an LDS hash table and
//    streams the prefixes of all rows i that hold j through it, 64
//    consecutive column ids per wave load.  With descending-degree labels
//    this checks sum_i C(d_L(i), 2) wedges at most, far below the
//    sum of d_L(i) + d_L(j) a merge would read.
//
// Work items are (pivot, range of prefixes) chunks of bounded size, taken
// by a persistent grid through one atomic counter (work stealing), heaviest
// first.  Per-vertex sums of row i collect the prefix counts of row i with
// int32 atomics.
#include <omp.h>

#include <algorithm>
#include <atomic>
#include <cstring>
#include <numeric>

#include "device_util.h"
#include "internal.h"

namespace btcx {
namespace {

constexpr int kTableSlots = 4096;  // 16 KiB of LDS per workgroup
constexpr uint32_t kEmpty = 0xFFFFFFFFu;
constexpr int kThreads = 256;
constexpr int kWaves = kThreads / 64;

__device__ __forceinline__ uint32_t hslot(uint32_t key, int shift) { return (key * 2654435761u) >> shift; }

template <bool PV>
__global__ __launch_bounds__(kThreads) void wedge_kernel(
    const uint32_t* __restrict__ rowptr, const int32_t* __restrict__ col, const uint32_t* __restrict__ us,
    const uint32_t* __restrict__ ue, const int32_t* __restrict__ ur, const uint4* __restrict__ items,
    uint32_t nitems, uint32_t* __restrict__ counter, unsigned long long* __restrict__ result,
    int32_t* __restrict__ pv) {
    __shared__ uint32_t table[kTableSlots];
    __shared__ uint32_t s_item;
    const int lane = threadIdx.x & 63;
    const uint32_t wid = uniform(threadIdx.x >> 6);
    uint64_t acc = 0;
    for (;;) {
        if (threadIdx.x == 0) s_item = atomicAdd(counter, 1u);
        __syncthreads();
        const uint32_t it = uniform(s_item);
        if (it >= nitems) break;
        const uint4 w = items[it];
        const uint32_t j = uniform(w.x), eb = uniform(w.y), ee = uniform(w.z);
        const uint32_t t0 = uniform(rowptr[j]), t1 = uniform(rowptr[j + 1]);
        const uint32_t d = t1 - t0;  // >= 1 for every listed pivot; 0 would match nothing
        if (d == 0) {
        } else if (d <= (uint32_t)kTableSlots / 2) {
            const int logS = max(6, 32 - __clz(2 * d - 1));
            const int shift = 32 - logS;
            const uint32_t mask = (1u << logS) - 1u;
            for (uint32_t s = threadIdx.x; s <= mask; s += kThreads) table[s] = kEmpty;
            __syncthreads();
            for (uint32_t p = t0 + threadIdx.x; p < t1; p += kThreads) {
                const uint32_t key = (uint32_t)col[p];
                uint32_t h = hslot(key, shift);
                while (atomicCAS(&table[h], kEmpty, key) != kEmpty) h = (h + 1u) & mask;
            }
            __syncthreads();
            for (uint32_t e = eb + wid; e < ee; e += kWaves) {
                const uint32_t s0 = uniform(us[e]), s1 = uniform(ue[e]);
                uint32_t hits = 0;
                for (uint32_t q = s0 + lane; q < s1; q += 64) {
                    const uint32_t key = (uint32_t)col[q];
                    uint32_t h = hslot(key, shift);
                    for (;;) {
                        const uint32_t v = table[h];
                        if (v == key) {
                            ++hits;
                            break;
                        }
                        if (v == kEmpty) break;
                        h = (h + 1u) & mask;
                    }
                }
                acc += hits;
                if (PV) {
                    const uint32_t c = wave_sum_u32(hits);
                    if (lane == 0 && c) atomicAdd(pv + ur[e], (int32_t)c);
                }
            }
        } else {
            // pivot row longer than the LDS table: binary search in HBM
            for (uint32_t e = eb + wid; e < ee; e += kWaves) {
                const uint32_t s0 = uniform(us[e]), s1 = uniform(ue[e]);
                uint32_t hits = 0;
                for (uint32_t q = s0 + lane; q < s1; q += 64) {
                    const int32_t key = col[q];
                    uint32_t lo = t0, hi = t1;
                    while (lo < hi) {
                        const uint32_t mid = (lo + hi) >> 1;
                        if (col[mid] < key) lo = mid + 1;
                        else hi = mid;
                    }
                    hits += (lo < t1 && col[lo] == key) ? 1u : 0u;
                }
                acc += hits;
                if (PV) {
                    const uint32_t c = wave_sum_u32(hits);
                    if (lane == 0 && c) atomicAdd(pv + ur[e], (int32_t)c);
                }
            }
        }
        __syncthreads();
    }
    acc = wave_sum_u64(acc);
    if (lane == 0 && acc) atomicAdd(result, (unsigned long long)acc);
}

}  // namespace
}  // namespace btcx

struct btc_plan_s {
    btc_graph_s* g = nullptr;
    int K = 0;
    int tau = 0;
    std::unique_ptr<btc_bitbsr_s> bsr;
    btcx::DevBuf<uint32_t> dense_list;
    uint32_t n_dense = 0;
    btcx::DevBuf<uint32_t> ustart, uend;
    btcx::DevBuf<int32_t> urow;
    btcx::DevBuf<uint4> items;
    uint32_t nitems = 0;
    btcx::DevBuf<unsigned long long> scratch;  // [0] total, [1] work-item counter
    btc_plan_info info{};
    float last_dense_ms = 0.f;
    float last_wedge_ms = 0.f;
};

namespace btcx {
namespace {

// Marks the L entries that fall into dense tiles and uploads the dense tile
// list.  Leaves p->bsr empty when no tile is dense.
void classify_dense(btc_plan_s* p, hipStream_t s, std::vector<uint8_t>& dense) {
    btc_graph_s* g = p->g;
    dense.assign((size_t)g->m, 0);
    if (p->tau >= 16 * p->K || g->m == 0) return;
    p->bsr = std::make_unique<btc_bitbsr_s>();
    bitbsr_build(g, p->K, s, p->bsr.get());
    btc_bitbsr_s* b = p->bsr.get();
    DevBuf<uint32_t> d_pop;
    d_pop.alloc(b->nblk);
    launch_tile_popcount(b, d_pop.p, s);
    std::vector<uint32_t> pop(b->nblk), indptr((size_t)b->n_rb + 1), indices(b->nblk);
    BTC_HIP(hipMemcpyAsync(pop.data(), d_pop.p, (size_t)b->nblk * 4, hipMemcpyDeviceToHost, s));
    BTC_HIP(hipMemcpyAsync(indptr.data(), b->indptr.p, indptr.size() * 4, hipMemcpyDeviceToHost, s));
    BTC_HIP(hipMemcpyAsync(indices.data(), b->indices.p, (size_t)b->nblk * 4, hipMemcpyDeviceToHost, s));
    BTC_HIP(hipStreamSynchronize(s));
    std::vector<uint32_t> list;
    int64_t dense_edges = 0;
    for (uint32_t t = 0; t < b->nblk; ++t)
        if ((int)pop[t] > p->tau) {
            list.push_back(t);
            dense_edges += pop[t];
        }
    p->n_dense = (uint32_t)list.size();
    p->info.dense_tiles = p->n_dense;
    p->info.dense_edges = dense_edges;
    if (list.empty()) {
        p->bsr.reset();
        return;
    }
    const int K = p->K;
#pragma omp parallel for schedule(dynamic, 1024)
    for (int32_t i = 0; i < g->n; ++i) {
        const uint32_t I = (uint32_t)i / 16;
        const uint32_t* b0 = indices.data() + indptr[I];
        const uint32_t* b1 = indices.data() + indptr[I + 1];
        for (int64_t q = g->rowptr[i]; q < g->rowptr[i + 1]; ++q) {
            const uint32_t jc = (uint32_t)g->col[q] / K;
            const uint32_t* f = std::lower_bound(b0, b1, jc);
            dense[q] = (int)pop[f - indices.data()] > p->tau;
        }
    }
    p->dense_list.alloc(list.size());
    BTC_HIP(hipMemcpyAsync(p->dense_list.p, list.data(), list.size() * 4, hipMemcpyHostToDevice, s));
    BTC_HIP(hipStreamSynchronize(s));
}

// Groups the sparse edges' clipped prefixes by pivot and cuts them into work
// items of about `chunk` wedge checks.
void build_wedge_lists(btc_plan_s* p, const std::vector<uint8_t>& dense, int64_t chunk, hipStream_t s) {
    btc_graph_s* g = p->g;
    const int32_t n = g->n;
    const int64_t* rp = g->rowptr.data();
    const int32_t* cl = g->col.data();
    std::vector<std::atomic<uint32_t>> cnt((size_t)n);
#pragma omp parallel for schedule(static)
    for (int32_t j = 0; j < n; ++j) cnt[j].store(0, std::memory_order_relaxed);

    // clipped prefix of entry q (row i, pivot j = col[q]); empty when nothing can match
    auto clip = [&](int32_t i, int64_t q, uint32_t& s0, uint32_t& s1) -> bool {
        const int32_t j = cl[q];
        if (rp[j + 1] == rp[j] || q == rp[i]) return false;
        const int32_t lo = cl[rp[j]], hi = cl[rp[j + 1] - 1];
        const int32_t* a = std::lower_bound(cl + rp[i], cl + q, lo);
        const int32_t* b = std::upper_bound(a, cl + q, hi);
        if (b <= a) return false;
        s0 = (uint32_t)(a - cl);
        s1 = (uint32_t)(b - cl);
        return true;
    };
    int64_t sparse_edges = 0, sparse_work = 0;
#pragma omp parallel for schedule(dynamic, 1024) reduction(+ : sparse_edges, sparse_work)
    for (int32_t i = 0; i < n; ++i)
        for (int64_t q = rp[i]; q < rp[i + 1]; ++q) {
            if (dense[q]) continue;
            ++sparse_edges;
            sparse_work += (rp[i + 1] - rp[i]) + (rp[cl[q] + 1] - rp[cl[q]]);
            uint32_t s0, s1;
            if (clip(i, q, s0, s1)) cnt[cl[q]].fetch_add(1, std::memory_order_relaxed);
        }
    std::vector<int64_t> off((size_t)n + 1, 0);
    for (int32_t j = 0; j < n; ++j) off[j + 1] = off[j] + cnt[j].load(std::memory_order_relaxed);
    const int64_t ne = off[n];
    BTC_REQUIRE(ne < (int64_t)0xFFFFFFFFLL, BTC_ERR_OVERFLOW, "too many wedge lists");
#pragma omp parallel for schedule(static)
    for (int32_t j = 0; j < n; ++j) cnt[j].store((uint32_t)off[j], std::memory_order_relaxed);
    std::vector<uint32_t> us((size_t)ne), ue((size_t)ne);
    std::vector<int32_t> ur((size_t)ne);
#pragma omp parallel for schedule(dynamic, 1024)
    for (int32_t i = 0; i < n; ++i)
        for (int64_t q = rp[i]; q < rp[i + 1]; ++q) {
            if (dense[q]) continue;
            uint32_t s0, s1;
            if (!clip(i, q, s0, s1)) continue;
            const uint32_t e = cnt[cl[q]].fetch_add(1, std::memory_order_relaxed);
            us[e] = s0;
            ue[e] = s1;
            ur[e] = i;
        }
    // work items: consecutive lists of one pivot, about `chunk` checks each
    std::vector<uint4> items;
    int64_t checks = 0, pivot_words = 0;
    for (int32_t j = 0; j < n; ++j) {
        int64_t e = off[j], e1 = off[j + 1];
        while (e < e1) {
            int64_t w = 0, f = e;
            while (f < e1 && (w == 0 || w + (ue[f] - us[f]) <= chunk)) {
                w += ue[f] - us[f];
                ++f;
            }
            checks += w;
            pivot_words += rp[j + 1] - rp[j];
            items.push_back(make_uint4((uint32_t)j, (uint32_t)e, (uint32_t)f, (uint32_t)std::min<int64_t>(w, 0xFFFFFFFFLL)));
            e = f;
        }
    }
    std::stable_sort(items.begin(), items.end(), [](const uint4& a, const uint4& b) { return a.w > b.w; });
    p->nitems = (uint32_t)items.size();
    p->info.sparse_edges = sparse_edges;
    p->info.sparse_merge_work = sparse_work;
    p->info.wedge_checks = checks;
    // prefix ids + pivot rows (4 B each), list entries (start, end, row: 12 B), items (16 B)
    p->info.stream_bytes = 4 * (checks + pivot_words) + 12 * ne + 16 * (int64_t)items.size();
    p->info.work_items = (int64_t)items.size();
    p->ustart.alloc((size_t)ne);
    p->uend.alloc((size_t)ne);
    p->urow.alloc((size_t)ne);
    p->items.alloc(items.size());
    if (ne) {
        BTC_HIP(hipMemcpyAsync(p->ustart.p, us.data(), (size_t)ne * 4, hipMemcpyHostToDevice, s));
        BTC_HIP(hipMemcpyAsync(p->uend.p, ue.data(), (size_t)ne * 4, hipMemcpyHostToDevice, s));
        BTC_HIP(hipMemcpyAsync(p->urow.p, ur.data(), (size_t)ne * 4, hipMemcpyHostToDevice, s));
    }
    if (!items.empty())
        BTC_HIP(hipMemcpyAsync(p->items.p, items.data(), items.size() * sizeof(uint4), hipMemcpyHostToDevice, s));
    BTC_HIP(hipStreamSynchronize(s));
}

template <bool PV>
int occupancy_blocks() {
    int nb = 0;
    BTC_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, wedge_kernel<PV>, kThreads, 0));
    return std::max(1, nb);
}

void plan_count(btc_plan_s* p, hipStream_t s, uint64_t* total, int32_t* per_vertex, float* kernel_ms) {
    btc_graph_s* g = p->g;
    BTC_REQUIRE(!per_vertex || g->max_row <= 65536, BTC_ERR_OVERFLOW,
                "longest row %d > 65536: per-vertex counts could overflow int32", g->max_row);
    DevBuf<int32_t> pv;
    if (per_vertex) {
        pv.alloc((size_t)g->n);
        BTC_HIP(hipMemsetAsync(pv.p, 0, (size_t)g->n * 4, s));
    }
    BTC_HIP(hipMemsetAsync(p->scratch.p, 0, 2 * sizeof(unsigned long long), s));
    unsigned long long* d_total = p->scratch.p;
    uint32_t* d_counter = reinterpret_cast<uint32_t*>(p->scratch.p + 1);
    EventTimer tm(kernel_ms != nullptr), tw(kernel_ms != nullptr);
    tm.start(s);
    if (p->n_dense) launch_tile_count(p->bsr.get(), p->dense_list.p, p->n_dense, false, -1, d_total, pv.p, s);
    tw.start(s);
    if (p->nitems) {
        const int cus = device_cu_count();
        if (per_vertex) {
            const unsigned grid = (unsigned)(cus * occupancy_blocks<true>());
            hipLaunchKernelGGL(wedge_kernel<true>, dim3(grid), dim3(kThreads), 0, s, g->d_rowptr.p, g->d_col.p,
                               p->ustart.p, p->uend.p, p->urow.p, p->items.p, p->nitems, d_counter, d_total, pv.p);
        } else {
            const unsigned grid = (unsigned)(cus * occupancy_blocks<false>());
            hipLaunchKernelGGL(wedge_kernel<false>, dim3(grid), dim3(kThreads), 0, s, g->d_rowptr.p, g->d_col.p,
                               p->ustart.p, p->uend.p, p->urow.p, p->items.p, p->nitems, d_counter, d_total,
                               nullptr);
        }
        BTC_CHECK_LAUNCH("wedge_kernel");
    }
    tm.stop(s);
    tw.stop(s);
    unsigned long long h = 0;
    BTC_HIP(hipMemcpyAsync(&h, d_total, 8, hipMemcpyDeviceToHost, s));
    if (per_vertex && g->n)
        BTC_HIP(hipMemcpyAsync(per_vertex, pv.p, (size_t)g->n * 4, hipMemcpyDeviceToHost, s));
    BTC_HIP(hipStreamSynchronize(s));
    if (kernel_ms) {
        *kernel_ms = tm.ms();
        p->last_wedge_ms = tw.ms();
        p->last_dense_ms = *kernel_ms - p->last_wedge_ms;
    }
    *total = h;
}

}  // namespace
}  // namespace btcx

extern "C" {

void btc_count_opts_default(btc_count_opts* o) {
    if (!o) return;
    std::memset(o, 0, sizeof(*o));
    o->tile_k = 0;
    o->tau = -1;
}

int btc_plan_create(btc_graph g, const btc_count_opts* opts, void* stream, btc_plan* out) {
    return btcx::guard([&] {
        BTC_REQUIRE(g && out, BTC_ERR_ARG, "NULL argument");
        *out = nullptr;
        btc_count_opts o;
        btc_count_opts_default(&o);
        if (opts) o = *opts;
        BTC_REQUIRE(o.tile_k == 0 || o.tile_k == 32 || o.tile_k == 64 || o.tile_k == 128, BTC_ERR_ARG,
                    "tile width K = %d (expected 0, 32, 64 or 128)", o.tile_k);
        hipStream_t s = btcx::as_stream(stream);
        auto p = std::make_unique<btc_plan_s>();
        p->g = g;
        p->K = o.tile_k ? o.tile_k : 64;
        p->tau = o.tau >= 0 ? o.tau : btc_default_tau(p->K);
        btcx::ensure_device(g, s);
        std::vector<uint8_t> dense;
        btcx::classify_dense(p.get(), s, dense);
        int64_t chunk = o.chunk_work > 0 ? o.chunk_work : 0;
        if (!chunk) {
            const int64_t cus = btcx::device_cu_count();
            chunk = std::max<int64_t>(8192, std::min<int64_t>(1 << 18, g->wedges / std::max<int64_t>(1, cus * 64)));
        }
        btcx::build_wedge_lists(p.get(), dense, chunk, s);
        p->scratch.alloc(2);
        p->info.tile_k = p->K;
        p->info.tau = p->tau;
        int64_t bytes = (int64_t)(p->ustart.bytes() + p->uend.bytes() + p->urow.bytes() + p->items.bytes() +
                                  p->dense_list.bytes() + p->scratch.bytes());
        if (p->bsr)
            bytes += (int64_t)(p->bsr->indptr.bytes() + p->bsr->indices.bytes() + p->bsr->row_indices.bytes() +
                               p->bsr->tiles.bytes());
        p->info.device_bytes = bytes;
        *out = p.release();
    });
}

int btc_plan_get_info(btc_plan p, btc_plan_info* info) {
    return btcx::guard([&] {
        BTC_REQUIRE(p && info, BTC_ERR_ARG, "NULL argument");
        *info = p->info;
    });
}

int btc_plan_count(btc_plan p, void* stream, uint64_t* total, int32_t* per_vertex, float* kernel_ms) {
    return btcx::guard([&] {
        BTC_REQUIRE(p && total, BTC_ERR_ARG, "NULL argument");
        btcx::plan_count(p, btcx::as_stream(stream), total, per_vertex, kernel_ms);
    });
}

int btc_plan_last_timing(btc_plan p, float* dense_ms, float* wedge_ms) {
    return btcx::guard([&] {
        BTC_REQUIRE(p, BTC_ERR_ARG, "NULL handle");
        if (dense_ms) *dense_ms = p->last_dense_ms;
        if (wedge_ms) *wedge_ms = p->last_wedge_ms;
    });
}

void btc_plan_free(btc_plan p) { delete p; }

int btc_count(btc_graph g, const btc_count_opts* opts, void* stream, uint64_t* total, int32_t* per_vertex,
              float* kernel_ms) {
    btc_plan p = nullptr;
    int rc = btc_plan_create(g, opts, stream, &p);
    if (rc != BTC_OK) return rc;
    rc = btc_plan_count(p, stream, total, per_vertex, kernel_ms);
    btc_plan_free(p);
    return rc;
}

}  // extern "C"
