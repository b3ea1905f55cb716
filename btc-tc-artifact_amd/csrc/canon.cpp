// Host canonicalisation: COO pairs -> strictly lower triangle L in CSR.
//
// Follows btc/operations/convert.h: orient (u, v) to (max, min) and drop
// self-loops (symmetrize_and_triangular_gpu, :281-324 with
// OrientLowerTriangular, reorder.h:39-46), sort and deduplicate, then compact
// the used vertex ids in order (remap_node_ids_gpu, :387-425).  The optional
// degree relabel (descending degree, ties by id) is not in the reference.
// The work is bucketed by row (counting sort) and each row is sorted on its
// own, which is O(m log d) and parallel over rows.
#include <omp.h>

#include <algorithm>
#include <atomic>
#include <memory>
#include <numeric>

#include "common.h"

namespace {

int64_t scan_exclusive(std::vector<int64_t>& v) {
    int64_t s = 0;
    for (auto& x : v) {
        int64_t t = x;
        x = s;
        s += t;
    }
    return s;
}

// Bucket oriented pairs (row > col) by row, sort each row, optionally drop
// duplicates.  `emit(f)` must call f(row, col) for every pair, in any order
// and from any thread; it is run twice (count, then scatter).
template <class Emit>
void bucket_rows(int32_t n, bool dedupe, Emit&& emit, std::vector<int64_t>& rowptr,
                 std::vector<int32_t>& col) {
    std::vector<std::atomic<int64_t>> cnt((size_t)n);
#pragma omp parallel for schedule(static)
    for (int32_t i = 0; i < n; ++i) cnt[i].store(0, std::memory_order_relaxed);
    emit([&](int32_t r, int32_t) { cnt[r].fetch_add(1, std::memory_order_relaxed); });
    std::vector<int64_t> off((size_t)n + 1, 0);
    for (int32_t i = 0; i < n; ++i) off[i] = cnt[i].load(std::memory_order_relaxed);
    const int64_t total = scan_exclusive(off);
    off[n] = total;
#pragma omp parallel for schedule(static)
    for (int32_t i = 0; i < n; ++i) cnt[i].store(off[i], std::memory_order_relaxed);
    std::vector<int32_t> tmp((size_t)total);
    emit([&](int32_t r, int32_t c) { tmp[cnt[r].fetch_add(1, std::memory_order_relaxed)] = c; });
    std::vector<int64_t> len((size_t)n + 1, 0);
#pragma omp parallel for schedule(dynamic, 1024)
    for (int32_t i = 0; i < n; ++i) {
        int32_t* b = tmp.data() + off[i];
        int32_t* e = tmp.data() + off[i + 1];
        std::sort(b, e);
        len[i] = dedupe ? (int64_t)(std::unique(b, e) - b) : (int64_t)(e - b);
    }
    rowptr.assign((size_t)n + 1, 0);
    for (int32_t i = 0; i < n; ++i) rowptr[i + 1] = rowptr[i] + len[i];
    if (!dedupe) {
        col.swap(tmp);
        return;
    }
    col.assign((size_t)rowptr[n], 0);
#pragma omp parallel for schedule(dynamic, 1024)
    for (int32_t i = 0; i < n; ++i)
        std::copy(tmp.data() + off[i], tmp.data() + off[i] + len[i], col.data() + rowptr[i]);
}

}  // namespace

namespace btcx {

void graph_finalize(btc_graph_s* g) {
    const int32_t n = g->n;
    g->m = n ? g->rowptr[n] : 0;
    int32_t mx = 0;
    int64_t wedges = 0, work = 0;
#pragma omp parallel for reduction(max : mx) reduction(+ : wedges, work) schedule(dynamic, 4096)
    for (int32_t i = 0; i < n; ++i) {
        int64_t d = g->rowptr[i + 1] - g->rowptr[i];
        mx = std::max<int32_t>(mx, (int32_t)d);
        wedges += d * (d - 1) / 2;
        for (int64_t p = g->rowptr[i]; p < g->rowptr[i + 1]; ++p) {
            int32_t j = g->col[p];
            work += d + (g->rowptr[j + 1] - g->rowptr[j]);
        }
    }
    g->max_row = mx;
    g->wedges = wedges;
    g->merge_work = work;
}

}  // namespace btcx

extern "C" {

int btc_graph_from_coo(const int32_t* src, const int32_t* dst, int64_t nnz, int32_t n,
                       uint32_t flags, btc_graph* out) {
    return btcx::guard([&] {
        BTC_REQUIRE(out, BTC_ERR_ARG, "out is NULL");
        *out = nullptr;
        BTC_REQUIRE(nnz >= 0, BTC_ERR_ARG, "nnz %lld < 0", (long long)nnz);
        BTC_REQUIRE(nnz == 0 || (src && dst), BTC_ERR_ARG, "src/dst is NULL");
        BTC_REQUIRE((flags & ~(BTC_CANON_DEFAULT | BTC_CANON_DEVICE)) == 0, BTC_ERR_ARG, "unknown flags 0x%x", flags);
        if (flags & BTC_CANON_DEVICE) {
            // stage the pairs in HBM (stream-ordered buffers: freeing them
            // does not wait for the whole device) and canonicalise there
            btcx::TmpBuf<int32_t> ds, dd;
            ds.alloc((size_t)nnz, nullptr);
            dd.alloc((size_t)nnz, nullptr);
            if (nnz) {
                const void* h[2] = {src, dst};
                void* d[2] = {ds.p, dd.p};
                btcx::upload_arrays(h, d, 2, (size_t)nnz * 4, nullptr);
            }
            auto g = std::make_unique<btc_graph_s>();
            btcx::canonicalise_device(ds.p, dd.p, nnz, n, flags, nullptr, g.get(), nullptr, nullptr);
            *out = g.release();
            return;
        }
        int32_t lo_id = 0, hi_id = -1;
#pragma omp parallel for reduction(min : lo_id) reduction(max : hi_id) schedule(static)
        for (int64_t e = 0; e < nnz; ++e) {
            lo_id = std::min(lo_id, std::min(src[e], dst[e]));
            hi_id = std::max(hi_id, std::max(src[e], dst[e]));
        }
        BTC_REQUIRE(lo_id >= 0, BTC_ERR_ARG, "negative vertex id %d", lo_id);
        if (n <= 0) n = hi_id + 1;
        BTC_REQUIRE(hi_id < n, BTC_ERR_ARG, "vertex id %d >= n = %d", hi_id, n);

        auto g = std::make_unique<btc_graph_s>();
        g->n_input = n;
        g->flags = flags;

        // orient to (max, min), drop self-loops, sort + dedupe per row
        std::vector<int64_t> rp;
        std::vector<int32_t> cl;
        bucket_rows(n, true,
                    [&](auto&& f) {
#pragma omp parallel for schedule(static, 65536)
                        for (int64_t e = 0; e < nnz; ++e) {
                            int32_t u = src[e], v = dst[e];
                            if (u == v) continue;
                            f(std::max(u, v), std::min(u, v));
                        }
                    },
                    rp, cl);
        const int64_t m = rp[n];

        // order-preserving compaction of used ids (skipped when nothing is used)
        std::vector<int32_t> newid((size_t)n), labels;
        int32_t nc = n;
        if ((flags & BTC_CANON_COMPACT) && m > 0) {
            std::vector<uint8_t> used((size_t)n, 0);
#pragma omp parallel for schedule(dynamic, 4096)
            for (int32_t i = 0; i < n; ++i) {
                if (rp[i + 1] > rp[i]) used[i] = 1;
                for (int64_t p = rp[i]; p < rp[i + 1]; ++p) used[cl[p]] = 1;  // benign race: all write 1
            }
            int32_t k = 0;
            for (int32_t i = 0; i < n; ++i) {
                newid[i] = used[i] ? k : -1;
                k += used[i];
            }
            nc = k;
        } else {
            std::iota(newid.begin(), newid.end(), 0);
        }
        labels.assign((size_t)nc, 0);
        for (int32_t i = 0; i < n; ++i)
            if (newid[i] >= 0) labels[newid[i]] = i;

        if (!(flags & BTC_CANON_DEGREE_ORDER)) {
            // monotone relabel keeps rows sorted: just drop unused rows and map columns
            g->rowptr.assign((size_t)nc + 1, 0);
            for (int32_t i = 0; i < n; ++i)
                if (newid[i] >= 0) g->rowptr[newid[i] + 1] = rp[i + 1] - rp[i];
            for (int32_t i = 0; i < nc; ++i) g->rowptr[i + 1] += g->rowptr[i];
            g->col.resize((size_t)m);
#pragma omp parallel for schedule(dynamic, 4096)
            for (int32_t i = 0; i < n; ++i) {
                if (newid[i] < 0) continue;
                int64_t o = g->rowptr[newid[i]];
                for (int64_t p = rp[i]; p < rp[i + 1]; ++p) g->col[o++] = newid[cl[p]];
            }
            g->labels.swap(labels);
        } else {
            // degree of every compacted vertex, then rank by (-degree, id)
            std::vector<std::atomic<int32_t>> deg((size_t)nc);
#pragma omp parallel for schedule(static)
            for (int32_t v = 0; v < nc; ++v) deg[v].store(0, std::memory_order_relaxed);
#pragma omp parallel for schedule(dynamic, 4096)
            for (int32_t i = 0; i < n; ++i) {
                if (newid[i] < 0) continue;
                deg[newid[i]].fetch_add((int32_t)(rp[i + 1] - rp[i]), std::memory_order_relaxed);
                for (int64_t p = rp[i]; p < rp[i + 1]; ++p)
                    deg[newid[cl[p]]].fetch_add(1, std::memory_order_relaxed);
            }
            std::vector<int32_t> order((size_t)nc);
            std::iota(order.begin(), order.end(), 0);
            std::stable_sort(order.begin(), order.end(), [&](int32_t a, int32_t b) {
                return deg[a].load(std::memory_order_relaxed) > deg[b].load(std::memory_order_relaxed);
            });
            std::vector<int32_t> rank((size_t)nc);
            for (int32_t r = 0; r < nc; ++r) rank[order[r]] = r;
            bucket_rows(nc, false,
                        [&](auto&& f) {
#pragma omp parallel for schedule(dynamic, 4096)
                            for (int32_t i = 0; i < n; ++i) {
                                if (newid[i] < 0) continue;
                                int32_t a = rank[newid[i]];
                                for (int64_t p = rp[i]; p < rp[i + 1]; ++p) {
                                    int32_t b = rank[newid[cl[p]]];
                                    f(std::max(a, b), std::min(a, b));
                                }
                            }
                        },
                        g->rowptr, g->col);
            g->labels.resize((size_t)nc);
            for (int32_t r = 0; r < nc; ++r) g->labels[r] = labels[order[r]];
        }
        g->n = nc;
        btcx::graph_finalize(g.get());
        *out = g.release();
    });
}

int btc_graph_from_csr(const int64_t* rowptr, const int32_t* col, int32_t n, btc_graph* out) {
    return btcx::guard([&] {
        BTC_REQUIRE(out, BTC_ERR_ARG, "out is NULL");
        *out = nullptr;
        BTC_REQUIRE(n >= 0 && rowptr, BTC_ERR_ARG, "bad n or rowptr");
        BTC_REQUIRE(rowptr[0] == 0, BTC_ERR_ARG, "rowptr[0] != 0");
        const int64_t m = rowptr[n];
        BTC_REQUIRE(m == 0 || col, BTC_ERR_ARG, "col is NULL");
        int bad = 0;
#pragma omp parallel for reduction(| : bad) schedule(dynamic, 4096)
        for (int32_t i = 0; i < n; ++i) {
            if (rowptr[i + 1] < rowptr[i]) { bad |= 1; continue; }
            for (int64_t p = rowptr[i]; p < rowptr[i + 1]; ++p) {
                if (col[p] < 0 || col[p] >= i) bad |= 2;
                if (p > rowptr[i] && col[p] <= col[p - 1]) bad |= 4;
            }
        }
        BTC_REQUIRE(!(bad & 1), BTC_ERR_ARG, "rowptr is not monotone");
        BTC_REQUIRE(!(bad & 2), BTC_ERR_ARG, "L must be strictly lower (0 <= col < row)");
        BTC_REQUIRE(!(bad & 4), BTC_ERR_ARG, "rows must be strictly ascending");
        auto g = std::make_unique<btc_graph_s>();
        g->n = g->n_input = n;
        g->rowptr.assign(rowptr, rowptr + n + 1);
        g->col.assign(col, col + m);
        g->labels.resize((size_t)n);
        std::iota(g->labels.begin(), g->labels.end(), 0);
        btcx::graph_finalize(g.get());
        *out = g.release();
    });
}

int btc_graph_get_info(btc_graph g, btc_graph_info* info) {
    return btcx::guard([&] {
        BTC_REQUIRE(g && info, BTC_ERR_ARG, "NULL argument");
        info->n = g->n;
        info->n_input = g->n_input;
        info->m = g->m;
        info->max_row = g->max_row;
        info->flags = g->flags;
        info->wedges = g->wedges;
        info->merge_work = g->merge_work;
    });
}

int btc_graph_copy_csr(btc_graph g, int64_t* rowptr, int32_t* col) {
    return btcx::guard([&] {
        BTC_REQUIRE(g && rowptr && (col || g->m == 0), BTC_ERR_ARG, "NULL argument");
        btcx::ensure_host(g);
        std::copy(g->rowptr.begin(), g->rowptr.end(), rowptr);
        std::copy(g->col.begin(), g->col.end(), col);
    });
}

int btc_graph_copy_labels(btc_graph g, int32_t* canon_to_input) {
    return btcx::guard([&] {
        BTC_REQUIRE(g && (canon_to_input || g->n == 0), BTC_ERR_ARG, "NULL argument");
        btcx::ensure_host(g);
        std::copy(g->labels.begin(), g->labels.end(), canon_to_input);
    });
}

void btc_graph_free(btc_graph g) { delete g; }

// Host OpenMP merge over L: for every (i, j) in L count |N_L(i) & N_L(j)| by a
// two-pointer walk (triangle_count_cpu.h:60-101); per_vertex[i] is the row sum.
int btc_count_cpu(btc_graph g, int32_t threads, uint64_t* total, int32_t* per_vertex) {
    return btcx::guard([&] {
        BTC_REQUIRE(g && total, BTC_ERR_ARG, "NULL argument");
        btcx::ensure_host(g);
        const int32_t n = g->n;
        const int64_t* rp = g->rowptr.data();
        const int32_t* c = g->col.data();
        uint64_t sum = 0;
        int overflow = 0;
        int nt = threads > 0 ? threads : omp_get_max_threads();
        (void)nt;  // unused in the device pass, where OpenMP pragmas are ignored
#pragma omp parallel for num_threads(nt) reduction(+ : sum) reduction(| : overflow) schedule(dynamic, 256)
        for (int32_t i = 0; i < n; ++i) {
            uint64_t row = 0;
            for (int64_t p = rp[i]; p < rp[i + 1]; ++p) {
                int32_t j = c[p];
                int64_t a = rp[i], ae = rp[i + 1], b = rp[j], be = rp[j + 1];
                while (a < ae && b < be) {
                    int32_t x = c[a], y = c[b];
                    if (x < y) ++a;
                    else if (x > y) ++b;
                    else { ++row; ++a; ++b; }
                }
            }
            sum += row;
            if (per_vertex) {
                if (row > (uint64_t)INT32_MAX) overflow = 1;
                per_vertex[i] = (int32_t)row;
            }
        }
        BTC_REQUIRE(!overflow, BTC_ERR_OVERFLOW, "a per-vertex count exceeds int32");
        *total = sum;
    });
}

}  // extern "C"
