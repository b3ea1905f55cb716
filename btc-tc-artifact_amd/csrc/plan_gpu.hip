// Device planner: the wedge-kernel plan of a graph whose L is resident in
// HBM, built by kernels, radix sorts and scans instead of host loops.
//
// The plan has the same contents as the host planner's (engine.hip,
// classify_dense + build_wedge_lists): dense mask tiles (popcount > tau)
// listed for the tile kernel; for every other edge (i, j) of L the prefix
// of row i in front of j clipped to [min N_L(j), max N_L(j)], with its
// hub-core bitmap choice; lists grouped by pivot j; work items of about
// `chunk` checks, workgroup items then wave items, each class by work.  Two
// details differ and neither changes a count: lists of one pivot are in
// edge order (a stable radix sort by pivot), and a pivot's lists are cut
// where the running cost crosses a multiple of `chunk` rather than
// greedily.
#include "device_util.h"
#include "internal.h"
#include "prim.h"

namespace btcx {
namespace {

constexpr int kT = 256;
constexpr unsigned kRedBlocks = 1024;    // eager reductions
constexpr unsigned kCapRedBlocks = 256;  // capacity-sized reductions: about one per CU

unsigned grid_for(int64_t n) {
    const int64_t g = (n + kT - 1) / kT;
    return (unsigned)std::max<int64_t>(1, std::min<int64_t>(g, 65536));
}

int id_bits(int64_t n) {
    int b = 1;
    while (b < 32 && (1ll << b) < n) ++b;
    return b;
}

#define GS_LOOP(i, n) for (int64_t i = blockIdx.x * (int64_t)kT + threadIdx.x; i < (n); i += (int64_t)gridDim.x * kT)

#define BTC_LAUNCH(kernel, n, s, ...)                                               \
    do {                                                                            \
        hipLaunchKernelGGL(kernel, dim3(grid_for(n)), dim3(kT), 0, s, __VA_ARGS__); \
        BTC_CHECK_LAUNCH(#kernel);                                                  \
    } while (0)

// Kernels that end in same-address atomics (one per workgroup) run on at
// most kRedBlocks workgroups.
#define BTC_LAUNCH_RED(kernel, n, s, ...)                                               \
    do {                                                                            \
        hipLaunchKernelGGL(kernel, dim3(std::min(grid_for(n), kRedBlocks)), dim3(kT), 0, s, __VA_ARGS__); \
        BTC_CHECK_LAUNCH(#kernel);                                                  \
    } while (0)

// First index in [lo, hi) of a[] with a[x] >= v (a ascending).
template <class T>
__device__ __forceinline__ uint32_t lower_bound_dev(const T* a, uint32_t lo, uint32_t hi, T v) {
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (a[mid] < v) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

template <class T>
__device__ __forceinline__ uint32_t upper_bound_dev(const T* a, uint32_t lo, uint32_t hi, T v) {
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (a[mid] <= v) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// erow[e] = row of entry e: one wave per row, lanes stride over its entries.
__global__ void edge_rows_kernel(const uint32_t* __restrict__ rowptr, Cnt n_, uint32_t* __restrict__ erow) {
    const int64_t n = (int64_t)n_.get();
    const int64_t waves = ((int64_t)gridDim.x * kT) >> 6;
    const int lane = threadIdx.x & 63;
    for (int64_t i = (blockIdx.x * (int64_t)kT + threadIdx.x) >> 6; i < n; i += waves)
        for (uint32_t p = rowptr[i] + lane; p < rowptr[i + 1]; p += 64) erow[p] = (uint32_t)i;
}

// Dense mask tiles: flag[t] = popcount > tau; dense edges summed.
__global__ void dense_tiles_kernel(const uint32_t* __restrict__ pop, uint32_t nblk, int tau,
                                   uint32_t* __restrict__ flag, unsigned long long* __restrict__ dense_edges) {
    uint64_t sum = 0;
    GS_LOOP(t, nblk) {
        const bool d = (int)pop[t] > tau;
        flag[t] = d;
        sum += d ? pop[t] : 0u;
    }
    __shared__ uint64_t lds[kT / 64];
    sum = block_sum_u64<kT>(sum, lds);
    if (threadIdx.x == 0 && sum) atomicAdd(dense_edges, (unsigned long long)sum);
}

__global__ void select_kernel(const uint32_t* __restrict__ flag, const uint32_t* __restrict__ pos, Cnt n_,
                              uint32_t* __restrict__ out) {
    const int64_t n = (int64_t)n_.get();
    GS_LOOP(t, n) if (flag[t]) out[pos[t]] = (uint32_t)t;
}

// dense[e] = the edge lies in a dense mask tile of the Bit-BSR.
__global__ void dense_edges_kernel(const int32_t* __restrict__ col, const uint32_t* __restrict__ erow, int64_t m,
                                   int K, const uint32_t* __restrict__ indptr, const uint32_t* __restrict__ indices,
                                   const uint32_t* __restrict__ pop, int tau, uint8_t* __restrict__ dense) {
    GS_LOOP(e, m) {
        const uint32_t I = erow[e] / 16, jc = (uint32_t)col[e] / (uint32_t)K;
        const uint32_t t = lower_bound_dev(indices, indptr[I], indptr[I + 1], jc);
        dense[e] = (int)pop[t] > tau;
    }
}

// Clipped prefix of entry q (row i, pivot j = col[q]) as in the host
// planner; false when nothing in it can match.
__device__ __forceinline__ bool clip(const uint32_t* __restrict__ rowptr, const int32_t* __restrict__ col,
                                     uint32_t i, uint32_t q, uint32_t& s0, uint32_t& s1) {
    const uint32_t j = (uint32_t)col[q];
    const uint32_t r0 = rowptr[i], t0 = rowptr[j], t1 = rowptr[j + 1];
    if (t1 == t0 || q == r0) return false;
    const int32_t lo = col[t0], hi = col[t1 - 1];
    s0 = lower_bound_dev(col, r0, q, lo);
    s1 = upper_bound_dev(col, s0, q, hi);
    return s1 > s0;
}

// Pass over the edges: list flag per edge, sparse edge count and merge work.
__global__ void clip_kernel(const uint32_t* __restrict__ rowptr, const int32_t* __restrict__ col,
                            const uint32_t* __restrict__ erow, const uint8_t* __restrict__ dense, Cnt m_,
                            uint32_t* __restrict__ flag, unsigned long long* __restrict__ stats) {
    const int64_t m = (int64_t)m_.get();
    uint64_t se = 0, sw = 0;
    GS_LOOP(q, m) {
        uint32_t f = 0;
        if (!dense || !dense[q]) {
            const uint32_t i = erow[q], j = (uint32_t)col[q];
            ++se;
            sw += (rowptr[i + 1] - rowptr[i]) + (rowptr[j + 1] - rowptr[j]);
            uint32_t s0, s1;
            f = clip(rowptr, col, i, (uint32_t)q, s0, s1);
        }
        flag[q] = f;
    }
    __shared__ uint64_t lds[kT / 64];
    se = block_sum_u64<kT>(se, lds);
    sw = block_sum_u64<kT>(sw, lds);
    if (threadIdx.x == 0) {
        if (se) atomicAdd(&stats[0], (unsigned long long)se);
        if (sw) atomicAdd(&stats[1], (unsigned long long)sw);
    }
}

__global__ void list_keys_kernel(const uint32_t* __restrict__ flag, const uint32_t* __restrict__ pos,
                                 const int32_t* __restrict__ col, Cnt m_, uint32_t* __restrict__ key,
                                 uint32_t* __restrict__ val) {
    const int64_t m = (int64_t)m_.get();
    GS_LOOP(q, m) if (flag[q]) {
        key[pos[q]] = (uint32_t)col[q];
        val[pos[q]] = (uint32_t)q;
    }
}

// off[j] = first list of pivot j (j = 0..n) in the pivot-sorted keys.
__global__ void pivot_offsets_kernel(const uint32_t* __restrict__ key, Cnt ne_, Cnt n_,
                                     uint32_t* __restrict__ off) {
    const uint32_t ne = (uint32_t)ne_.get();
    const int64_t n = (int64_t)n_.get();
    GS_LOOP(j, (int64_t)n + 1) off[j] = lower_bound_dev(key, 0u, ne, (uint32_t)j);
}

// Core ids (< C) per row; rows with at least kCoreRowMin get a bitmap.
__global__ void core_rows_kernel(const uint32_t* __restrict__ rowptr, const int32_t* __restrict__ col, Cnt n_,
                                 uint32_t C, uint32_t* __restrict__ cend, uint32_t* __restrict__ has) {
    const int64_t n = (int64_t)n_.get();
    GS_LOOP(i, n) {
        const uint32_t r0 = rowptr[i];
        const uint32_t c = lower_bound_dev(col, r0, rowptr[i + 1], (int32_t)C);
        cend[i] = c;
        has[i] = c - r0 >= kCoreRowMin;
    }
}

__global__ void core_fill_kernel(const uint32_t* __restrict__ rowptr, const int32_t* __restrict__ col,
                                 const uint32_t* __restrict__ cend, const uint32_t* __restrict__ has,
                                 const uint32_t* __restrict__ rbi, Cnt n_, uint32_t CW,
                                 uint32_t* __restrict__ rowbits) {
    const int64_t n = (int64_t)n_.get();
    const int64_t waves = ((int64_t)gridDim.x * kT) >> 6;
    const int lane = threadIdx.x & 63;
    for (int64_t i = (blockIdx.x * (int64_t)kT + threadIdx.x) >> 6; i < n; i += waves) {
        if (!has[i]) continue;
        uint32_t* b = rowbits + (size_t)rbi[i] * CW;
        for (uint32_t p = rowptr[i] + lane; p < cend[i]; p += 64) {
            const uint32_t k = (uint32_t)col[p];
            atomicOr(&b[k >> 5], 1u << (k & 31u));
        }
    }
}

// List record of every pivot-sorted list e (see the wedge kernel's
// scan_lists), its row, and its cost (ids looked up + core words ANDed).
__global__ void records_kernel(const uint32_t* __restrict__ rowptr, const int32_t* __restrict__ col,
                               const uint32_t* __restrict__ erow, const uint32_t* __restrict__ val, Cnt ne_,
                               const uint32_t* __restrict__ cend, const uint32_t* __restrict__ has,
                               const uint32_t* __restrict__ rbi, uint32_t C, uint4* __restrict__ lst,
                               int32_t* __restrict__ ur, uint64_t* __restrict__ cost,
                               unsigned long long* __restrict__ ncore, unsigned long long* __restrict__ nnarrow) {
    const uint32_t ne = (uint32_t)ne_.get();
    uint64_t nc = 0, nn = 0;
    GS_LOOP(e, ne) {
        const uint32_t q = val[e], i = erow[q], j = (uint32_t)col[q];
        uint32_t s0 = 0, s1 = 0;
        clip(rowptr, col, i, q, s0, s1);
        const uint32_t lo = (uint32_t)col[rowptr[j]], hi = (uint32_t)col[rowptr[j + 1] - 1];
        const uint32_t base = lo & ~127u;
        uint32_t sc = s0, rb = kNoBitmap, words = 0;
        if (has[i] && base < C) {
            const uint32_t ce = min(s1, cend[i]);
            const uint32_t core_ids = ce > s0 ? ce - s0 : 0u;
            const uint32_t w = core_words(lo, hi, C);
            if (take_core_bitmap(core_ids, w)) {
                sc = max(s0, cend[i]);
                rb = rbi[i];
                words = w;
                ++nc;
            }
        }
        lst[e] = make_uint4(s0, sc, s1, rb);
        ur[e] = (int32_t)i;
        cost[e] = (uint64_t)(s1 > sc ? s1 - sc : 0u) + words;
        if (hi < kNarrowIds) nn += s1 > sc ? s1 - sc : 0u;
    }
    __shared__ uint64_t lds[kT / 64];
    nc = block_sum_u64<kT>(nc, lds);
    nn = block_sum_u64<kT>(nn, lds);
    if (threadIdx.x == 0 && nc) atomicAdd(ncore, (unsigned long long)nc);
    if (threadIdx.x == 0 && nn) atomicAdd(nnarrow, (unsigned long long)nn);
}

__device__ __forceinline__ bool pivot_small(const uint32_t* __restrict__ rowptr, const int32_t* __restrict__ col,
                                            const uint32_t* __restrict__ off, uint32_t j) {
    const uint32_t t0 = rowptr[j], t1 = rowptr[j + 1];
    return wave_pivot(off[j + 1] - off[j], (uint32_t)col[t0], (uint32_t)col[t1 - 1], t1 - t0);
}

// start[e] = list e opens a work item: the pivot's first list, or (large
// pivots) the running cost crosses a multiple of chunk.
__global__ void item_start_kernel(const uint32_t* __restrict__ rowptr, const int32_t* __restrict__ col,
                                  const uint32_t* __restrict__ key, const uint32_t* __restrict__ off,
                                  const uint64_t* __restrict__ P, Cnt ne_, uint64_t chunk_,
                                  const unsigned long long* __restrict__ d_checks, uint64_t target_items,
                                  uint32_t* __restrict__ start) {
    const uint32_t ne = (uint32_t)ne_.get();
    // with d_checks: about target_items items, at least 8192 checks each
    const uint64_t chunk =
        d_checks ? max((uint64_t)8192, ((uint64_t)*d_checks + target_items - 1) / target_items) : chunk_;
    GS_LOOP(e, ne) {
        const uint32_t j = key[e], eb = off[j];
        uint32_t st = e == eb;
        if (!st && !pivot_small(rowptr, col, off, j)) st = (P[e] - P[eb]) / chunk != (P[e - 1] - P[eb]) / chunk;
        start[e] = st;
    }
}

// Item k = lists [first[k], first[k + 1]); sort key puts workgroup items
// before wave items before tiny items, each by descending work.
__global__ void items_kernel(const uint32_t* __restrict__ rowptr, const int32_t* __restrict__ col,
                             const uint32_t* __restrict__ key, const uint32_t* __restrict__ off,
                             const uint64_t* __restrict__ P, const uint32_t* __restrict__ first, Cnt nitems_,
                             Cnt ne_, bool use_tiny, bool tiny_all, bool coarse, uint4* __restrict__ tmp_items,
                             uint32_t* __restrict__ skey, uint32_t* __restrict__ sval,
                             unsigned long long* __restrict__ stats) {
    const uint32_t nitems = (uint32_t)nitems_.get();
    const uint32_t ne = (uint32_t)ne_.get();
    uint64_t words = 0, small_items = 0, tiny_items = 0;
    GS_LOOP(k, nitems) {
        const uint32_t eb = first[k], ee = k + 1 < nitems ? first[k + 1] : ne;
        const uint32_t j = key[eb];
        const uint32_t t0 = rowptr[j], t1 = rowptr[j + 1];
        const uint64_t w = P[ee] - P[eb];
        const uint32_t w32 = (uint32_t)min(w, (uint64_t)0xFFFFFFFFu);
        const uint32_t lo = (uint32_t)col[t0], hi = (uint32_t)col[t1 - 1];
        const bool small = pivot_small(rowptr, col, off, j);
        const bool tiny = use_tiny && small && tiny_pivot(off[j + 1] - off[j], lo, hi, t1 - t0, tiny_all);
        tmp_items[2 * k] = make_uint4(t0, t1, eb, ee);
        tmp_items[2 * k + 1] = make_uint4(lo, hi, w32, 0u);
        const uint32_t cls = tiny ? 2u : small ? 1u : 0u;
        if (coarse) {
            // 8-bit key: class, then work by half octaves, heaviest first
            const uint32_t len = w ? 64 - __clzll(w) : 0u;
            const uint32_t half = len >= 2 ? (uint32_t)(w >> (len - 2)) & 1u : 0u;
            skey[k] = (cls << 6) | (63u - min(63u, 2 * len + half));
        } else {
            const uint32_t w30 = (uint32_t)min(w, (uint64_t)0x3FFFFFFFu);
            skey[k] = (cls << 30) | (0x3FFFFFFFu - w30);
        }
        sval[k] = (uint32_t)k;
        words += t1 - t0;
        small_items += small;
        tiny_items += tiny;
    }
    __shared__ uint64_t lds[kT / 64];
    words = block_sum_u64<kT>(words, lds);
    small_items = block_sum_u64<kT>(small_items, lds);
    tiny_items = block_sum_u64<kT>(tiny_items, lds);
    if (threadIdx.x == 0) {
        if (words) atomicAdd(&stats[0], (unsigned long long)words);
        if (small_items) atomicAdd(&stats[1], (unsigned long long)small_items);
        if (tiny_items) atomicAdd(&stats[3], (unsigned long long)tiny_items);
    }
}

// With c set (a capacity plan), workgroup 0 also records the item counts of
// the wedge kernel's launch (wave items, st[5], follow the workgroup items;
// there are no tiny items) and clears the count's total and cursors.
__global__ void place_items_kernel(const uint4* __restrict__ tmp_items, const uint32_t* __restrict__ order,
                                   Cnt nitems_, uint4* __restrict__ items, PipeCtl* __restrict__ c,
                                   unsigned long long* __restrict__ work) {
    const uint32_t nitems = (uint32_t)nitems_.get();
    if (c && blockIdx.x == 0 && threadIdx.x == 0) {
        c->item_counts[0] = c->nitems - (uint32_t)c->st[5];
        c->item_counts[1] = c->nitems;
        work[0] = 0;
        work[1] = 0;
    }
    GS_LOOP(r, nitems) {
        const uint32_t k = order[r];
        items[2 * r] = tmp_items[2 * k];
        items[2 * r + 1] = tmp_items[2 * k + 1];
    }
}

// ------------------------------------------------- capturable form kernels
// pivot_offsets_kernel and core_rows_kernel in one launch.
__global__ void pivot_core_kernel(const uint32_t* __restrict__ key, Cnt ne_, Cnt n_,
                                  const uint32_t* __restrict__ rowptr, const int32_t* __restrict__ col, uint32_t C,
                                  uint32_t* __restrict__ off, uint32_t* __restrict__ cend,
                                  uint32_t* __restrict__ has) {
    const uint32_t ne = (uint32_t)ne_.get();
    const int64_t n = n_.get();
    GS_LOOP(j, n + 1) off[j] = lower_bound_dev(key, 0u, ne, (uint32_t)j);
    GS_LOOP(i, n) {
        const uint32_t r0 = rowptr[i];
        const uint32_t c = lower_bound_dev(col, r0, rowptr[i + 1], (int32_t)C);
        cend[i] = c;
        has[i] = c - r0 >= kCoreRowMin;
    }
}

// rowbits[0, nbm * CW) = 0
__global__ void rowbits_clear_kernel(uint32_t* __restrict__ rowbits, const uint32_t* __restrict__ d_nbm,
                                     uint32_t CW) {
    const int64_t n = (int64_t)*d_nbm * CW;
    GS_LOOP(i, n) rowbits[i] = 0;
}

template <class T>
T read_dev(const T* p, hipStream_t s) {
    T h{};
    BTC_HIP(hipMemcpyAsync(&h, p, sizeof(T), hipMemcpyDeviceToHost, s));
    BTC_HIP(hipStreamSynchronize(s));
    return h;
}

}  // namespace

void plan_device(btc_plan_s* p, int64_t chunk, hipStream_t s) {
    btc_graph_s* g = p->g;
    const int32_t n = g->n;
    const int64_t m = g->m;
    const uint32_t* rowptr = g->d_rowptr.p;
    const int32_t* col = g->d_col.p;
    BTC_REQUIRE(m < (int64_t)0xFFFFFFFFLL, BTC_ERR_OVERFLOW, "nnz(L) %lld exceeds 2^32 - 1", (long long)m);

    TmpBuf<uint32_t> erow;
    erow.alloc((size_t)m, s);
    if (m) BTC_LAUNCH(edge_rows_kernel, (int64_t)n * 64, s, rowptr, n, erow.p);

    // dense mask tiles
    TmpBuf<unsigned long long> st;
    st.alloc(8, s);
    BTC_HIP(hipMemsetAsync(st.p, 0, 64, s));
    TmpBuf<uint8_t> dense;
    p->n_dense = 0;
    if (p->tau < 16 * p->K && m > 0) {
        const btc_bitbsr_s* b = plan_bsr(p, s);
        TmpBuf<uint32_t> pop, flag, pos;
        pop.alloc(b->nblk, s);
        flag.alloc(b->nblk, s);
        pos.alloc((size_t)b->nblk + 1, s);
        launch_tile_popcount(b, pop.p, s);
        BTC_LAUNCH_RED(dense_tiles_kernel, b->nblk, s, pop.p, b->nblk, p->tau, flag.p, st.p + 0);
        exclusive_scan_u32(flag.p, pos.p, b->nblk, s);
        const uint32_t nd = read_dev(pos.p + b->nblk, s);
        if (nd) {
            p->dense_list.alloc(nd);
            BTC_LAUNCH(select_kernel, b->nblk, s, flag.p, pos.p, (int64_t)b->nblk, p->dense_list.p);
            dense.alloc((size_t)m, s);
            BTC_LAUNCH(dense_edges_kernel, m, s, col, erow.p, m, p->K, b->indptr.p, b->indices.p, pop.p, p->tau,
                       dense.p);
            p->n_dense = nd;
        }
        BTC_HIP(hipStreamSynchronize(s));
        if (!nd) drop_plan_bsr(p);
    }
    p->info.dense_tiles = p->n_dense;
    p->info.dense_edges = (int64_t)read_dev(st.p + 0, s);

    // prefix lists: flag, compact (pivot, edge), stable sort by pivot
    TmpBuf<uint32_t> lflag, lpos;
    lflag.alloc((size_t)m, s);
    lpos.alloc((size_t)m + 1, s);
    if (m) BTC_LAUNCH_RED(clip_kernel, m, s, rowptr, col, erow.p, p->n_dense ? dense.p : nullptr, m, lflag.p, st.p + 1);
    exclusive_scan_u32(lflag.p, lpos.p, (size_t)m, s);
    const uint32_t ne = read_dev(lpos.p + m, s);
    TmpBuf<uint32_t> k0, v0, k1, v1;
    k0.alloc(ne, s);
    v0.alloc(ne, s);
    k1.alloc(ne, s);
    v1.alloc(ne, s);
    if (m) BTC_LAUNCH(list_keys_kernel, m, s, lflag.p, lpos.p, col, m, k0.p, v0.p);
    lflag.release();
    lpos.release();
    dense.release();
    const bool alt = sort_pairs_u32(k0.p, v0.p, k1.p, v1.p, ne, id_bits(n), s);
    const uint32_t* key = alt ? k1.p : k0.p;
    const uint32_t* val = alt ? v1.p : v0.p;
    TmpBuf<uint32_t> off;
    off.alloc((size_t)n + 1, s);
    BTC_LAUNCH(pivot_offsets_kernel, (int64_t)n + 1, s, key, ne, n, off.p);

    // hub-core bitmaps
    TmpBuf<uint32_t> cend, has, rbi;
    cend.alloc((size_t)n, s);
    has.alloc((size_t)n, s);
    rbi.alloc((size_t)n + 1, s);
    const uint32_t C = p->core_ids;
    if (n) BTC_LAUNCH(core_rows_kernel, n, s, rowptr, col, n, C, cend.p, has.p);
    exclusive_scan_u32(has.p, rbi.p, (size_t)n, s);
    const uint32_t nbm = read_dev(rbi.p + n, s);

    // one allocation: list records, row bitmaps, list rows; items apart
    const size_t lb = (size_t)ne * sizeof(uint4), bb = (size_t)nbm * (C / 32) * 4, rb = (size_t)ne * 4;
    p->lists.alloc(lb + bb + rb);
    uint4* lst = reinterpret_cast<uint4*>(p->lists.p);
    uint32_t* rowbits = reinterpret_cast<uint32_t*>(p->lists.p + lb);
    int32_t* ur = reinterpret_cast<int32_t*>(p->lists.p + lb + bb);
    if (bb) {
        BTC_HIP(hipMemsetAsync(rowbits, 0, bb, s));
        BTC_LAUNCH(core_fill_kernel, (int64_t)n * 64, s, rowptr, col, cend.p, has.p, rbi.p, n, C / 32, rowbits);
    }
    TmpBuf<uint64_t> cost, P;
    cost.alloc(ne, s);
    P.alloc((size_t)ne + 1, s);
    if (ne)
        BTC_LAUNCH_RED(records_kernel, ne, s, rowptr, col, erow.p, val, ne, cend.p, has.p, rbi.p, C, lst, ur, cost.p,
                   st.p + 3, st.p + 6);
    exclusive_scan_u64(cost.p, P.p, ne, s);
    erow.release();

    // work items
    TmpBuf<uint32_t> start, spos;
    start.alloc(ne, s);
    spos.alloc((size_t)ne + 1, s);
    if (ne) BTC_LAUNCH(item_start_kernel, ne, s, rowptr, col, key, off.p, P.p, ne, (uint64_t)chunk, nullptr, (uint64_t)1,
                       start.p);
    exclusive_scan_u32(start.p, spos.p, ne, s);
    const uint32_t nitems = read_dev(spos.p + ne, s);
    TmpBuf<uint32_t> first, sk0, sv0, sk1, sv1;
    first.alloc(nitems, s);
    if (ne) BTC_LAUNCH(select_kernel, ne, s, start.p, spos.p, (int64_t)ne, first.p);
    TmpBuf<uint4> tmp;
    tmp.alloc((size_t)nitems * 2, s);
    sk0.alloc(nitems, s);
    sv0.alloc(nitems, s);
    sk1.alloc(nitems, s);
    sv1.alloc(nitems, s);
    if (nitems)
        BTC_LAUNCH_RED(items_kernel, nitems, s, rowptr, col, key, off.p, P.p, first.p, nitems, ne,
                   p->use_tiny, p->tiny_all, false, tmp.p, sk0.p, sv0.p, st.p + 4);
    const bool ialt = sort_pairs_u32(sk0.p, sv0.p, sk1.p, sv1.p, nitems, 32, s);
    p->items_dev.alloc((size_t)nitems * 2);
    if (nitems)
        BTC_LAUNCH(place_items_kernel, nitems, s, tmp.p, ialt ? sv1.p : sv0.p, nitems, p->items_dev.p, nullptr,
                   nullptr);

    // st: [0] dense edges, [1] sparse edges, [2] their merge work, [3] core
    // lists, [4] pivot row words over items, [5] wave items (tiny ones
    // included), [6] ids looked up in the 16-bit copy, [7] tiny items
    unsigned long long h[8];
    BTC_HIP(hipMemcpyAsync(h, st.p, sizeof(h), hipMemcpyDeviceToHost, s));
    const uint64_t checks = read_dev(P.p + ne, s);  // synchronises h as well
    p->items_p = p->items_dev.p;
    p->list_p = lst;
    p->rowbits_p = reinterpret_cast<const uint4*>(rowbits);
    p->urow_p = ur;
    p->nitems = nitems;
    p->n_block_items = nitems - (uint32_t)h[5];
    p->n_tiny_items = (uint32_t)h[7];
    p->info.tiny_items = p->n_tiny_items;
    p->info.sparse_edges = (int64_t)h[1];
    p->info.sparse_merge_work = (int64_t)h[2];
    p->info.core_lists = (int64_t)h[3];
    p->info.core_rows = nbm;
    p->info.wedge_checks = (int64_t)checks;
    p->info.work_items = nitems;
    // looked-up ids (4 B, or 2 B below kNarrowIds), core words and pivot
    // rows (4 B), list records (16 B), items (32 B)
    p->info.stream_bytes = 4 * ((int64_t)checks + (int64_t)h[4]) - 2 * (int64_t)h[6] + 16 * (int64_t)ne +
                           32 * (int64_t)nitems;
}

namespace {

unsigned cap_grid(int64_t cap) {
    return (unsigned)std::max<int64_t>(1, std::min<int64_t>((cap + kT - 1) / kT, 2048));
}

}  // namespace

#define CAP_LAUNCH(kernel, cap, s, ...)                                                \
    do {                                                                               \
        hipLaunchKernelGGL(kernel, dim3(cap_grid(cap)), dim3(kT), 0, s, __VA_ARGS__);  \
        BTC_CHECK_LAUNCH(#kernel);                                                     \
    } while (0)

#define CAP_LAUNCH_RED(kernel, cap, s, ...)                                                \
    do {                                                                               \
        hipLaunchKernelGGL(kernel, dim3(std::min(cap_grid(cap), kCapRedBlocks)), dim3(kT), 0, s, __VA_ARGS__);  \
        BTC_CHECK_LAUNCH(#kernel);                                                     \
    } while (0)


void plan_capacity(const PlanCap& c, hipStream_t s) {
    PipeCtl* ctl = c.ctl;
    const int64_t M = c.m_cap, N = c.n_cap, I = (int64_t)c.items_cap;
    const Cnt m = Cnt::dev(&ctl->m), nc = Cnt::dev(&ctl->nc), ne = Cnt::dev(&ctl->ne),
              nitems = Cnt::dev(&ctl->nitems);
    const uint32_t* rowptr = c.rowptr;
    const int32_t* col = c.col;
    unsigned long long* st = ctl->st;  // cleared with the rest of *ctl

    CAP_LAUNCH(edge_rows_kernel, N * 64, s, rowptr, nc, c.erow);
    // prefix lists: flag, compact (pivot, edge), stable sort by pivot
    CAP_LAUNCH_RED(clip_kernel, M, s, rowptr, col, c.erow, nullptr, m, c.lflag, st + 1);
    dscan_u32(c.lflag, c.lpos, (size_t)M, &ctl->m, &ctl->ne, c.ws, s);
    CAP_LAUNCH(list_keys_kernel, M, s, c.lflag, c.lpos, col, m, c.k0, c.v0);
    const bool alt = dsort_pairs_u32(c.k0, c.v0, c.k1, c.v1, (size_t)M, &ctl->ne, id_bits(N), c.ws, s);
    const uint32_t* key = alt ? c.k1 : c.k0;
    const uint32_t* val = alt ? c.v1 : c.v0;

    // pivot offsets and hub-core rows, bitmaps
    const uint32_t C = c.core_ids;
    CAP_LAUNCH(pivot_core_kernel, N + 1, s, key, ne, nc, rowptr, col, C, c.off, c.cend, c.has);
    dscan_u32(c.has, c.rbi, (size_t)N, &ctl->nc, &ctl->nbm, c.ws, s);
    if (C) {
        CAP_LAUNCH(rowbits_clear_kernel, (int64_t)pipe_bitmap_rows(M, c.n_cap) * (C / 32), s, c.rowbits, &ctl->nbm,
                   C / 32);
        CAP_LAUNCH(core_fill_kernel, N * 64, s, rowptr, col, c.cend, c.has, c.rbi, nc, C / 32, c.rowbits);
    }
    CAP_LAUNCH_RED(records_kernel, M, s, rowptr, col, c.erow, val, ne, c.cend, c.has, c.rbi, C, c.lists, c.ur, c.cost,
               st + 3, st + 6);
    dscan_u64(c.cost, c.P, (size_t)M, &ctl->ne, reinterpret_cast<uint64_t*>(&ctl->checks), c.ws, s);

    // work items (each opens at a list, so there are at most ne of them):
    // about 64 per CU, sorted by class and then work in half octaves
    CAP_LAUNCH(item_start_kernel, M, s, rowptr, col, key, c.off, c.P, ne, (uint64_t)0, &ctl->checks,
               (uint64_t)64 * (uint64_t)c.cus, c.start);
    dscan_u32(c.start, c.spos, (size_t)M, &ctl->ne, &ctl->nitems, c.ws, s);
    CAP_LAUNCH(select_kernel, M, s, c.start, c.spos, ne, c.first);
    CAP_LAUNCH_RED(items_kernel, I, s, rowptr, col, key, c.off, c.P, c.first, nitems, ne, false, false, true, c.tmp_items,
               c.sk0, c.sv0, st + 4);
    const bool ialt = dsort_pairs_u32(c.sk0, c.sv0, c.sk1, c.sv1, (size_t)I, &ctl->nitems, 8, c.ws, s);
    CAP_LAUNCH(place_items_kernel, I, s, c.tmp_items, ialt ? c.sv1 : c.sv0, nitems, c.items, ctl, c.work);
}

}  // namespace btcx
