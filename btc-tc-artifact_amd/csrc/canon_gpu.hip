// Device canonicalisation: COO pairs in HBM -> strictly lower triangle L
// (CSR) in HBM, equal bit for bit to the host path (canon.cpp).
//
// Same steps as btc/operations/convert.h: orient (u, v) to (max, min) and
// drop self-loops (symmetrize_and_triangular_gpu, :281-324), sort and
// deduplicate, compact the used ids in order (remap_node_ids_gpu,
// :387-425); then the optional relabel by descending degree (ties by id).
// A pair is one key row << b | col (b = bits of n), radix-sorted on 2b
// bits; duplicates are dropped by a flag scan; CSR rows are found by a
// binary search of each row's first key.  The host copies of the CSR are
// made only when a host consumer asks (ensure_host).
#include <climits>

#include "device_util.h"
#include "internal.h"
#include "prim.h"

namespace btcx {
namespace {

constexpr int kT = 256;
constexpr unsigned kRedBlocks = 1024;    // eager reductions
constexpr unsigned kCapRedBlocks = 256;  // capacity-sized reductions: about one per CU

int id_bits(int64_t n) {
    int b = 1;
    while (b < 31 && (1ll << b) < n) ++b;
    return b;
}

unsigned grid_for(int64_t n) {
    const int64_t g = (n + kT - 1) / kT;
    return (unsigned)std::max<int64_t>(1, std::min<int64_t>(g, 65536));
}

__global__ void id_range_kernel(const int32_t* __restrict__ src, const int32_t* __restrict__ dst, Cnt nnz_,
                                int32_t* __restrict__ lohi) {
    __shared__ int32_t s_lo[kT / 64], s_hi[kT / 64];
    const int64_t nnz = nnz_.get();
    int32_t lo = INT_MAX, hi = -1;
    for (int64_t e = blockIdx.x * (int64_t)kT + threadIdx.x; e < nnz; e += (int64_t)gridDim.x * kT) {
        const int32_t u = src[e], v = dst[e];
        lo = min(lo, min(u, v));
        hi = max(hi, max(u, v));
    }
    for (int o = 32; o > 0; o >>= 1) {
        lo = min(lo, __shfl_xor(lo, o));
        hi = max(hi, __shfl_xor(hi, o));
    }
    if ((threadIdx.x & 63) == 0) {
        s_lo[threadIdx.x >> 6] = lo;
        s_hi[threadIdx.x >> 6] = hi;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < kT / 64; ++w) {
            lo = min(lo, s_lo[w]);
            hi = max(hi, s_hi[w]);
        }
        if (lo != INT_MAX) atomicMin(&lohi[0], lo);
        if (hi != -1) atomicMax(&lohi[1], hi);
    }
}

// (u, v) -> (max << b) | min; self-loops become the sentinel (all 2b bits
// set, which no oriented pair reaches since max > min).
__global__ void orient_kernel(const int32_t* __restrict__ src, const int32_t* __restrict__ dst, int64_t nnz, int b,
                              uint64_t sentinel, uint64_t* __restrict__ keys) {
    for (int64_t e = blockIdx.x * (int64_t)kT + threadIdx.x; e < nnz; e += (int64_t)gridDim.x * kT) {
        const uint32_t u = (uint32_t)src[e], v = (uint32_t)dst[e];
        keys[e] = u == v ? sentinel : ((uint64_t)max(u, v) << b) | (uint64_t)min(u, v);
    }
}

// flag[e] = key e is the first of its run and not the sentinel
__global__ void first_flag_kernel(const uint64_t* __restrict__ keys, Cnt n_, uint64_t sentinel,
                                  uint32_t* __restrict__ flag) {
    const int64_t n = n_.get();
    for (int64_t e = blockIdx.x * (int64_t)kT + threadIdx.x; e < n; e += (int64_t)gridDim.x * kT) {
        const uint64_t k = keys[e];
        flag[e] = (k != sentinel && (e == 0 || keys[e - 1] != k)) ? 1u : 0u;
    }
}

__global__ void compact_kernel(const uint64_t* __restrict__ keys, const uint32_t* __restrict__ flag,
                               const uint32_t* __restrict__ pos, Cnt n_, uint64_t* __restrict__ out) {
    const int64_t n = n_.get();
    for (int64_t e = blockIdx.x * (int64_t)kT + threadIdx.x; e < n; e += (int64_t)gridDim.x * kT)
        if (flag[e]) out[pos[e]] = keys[e];
}

__global__ void mark_used_kernel(const uint64_t* __restrict__ keys, Cnt m_, int b, uint32_t* __restrict__ used) {
    const int64_t m = m_.get();
    const uint64_t mask = (1ull << b) - 1;
    for (int64_t e = blockIdx.x * (int64_t)kT + threadIdx.x; e < m; e += (int64_t)gridDim.x * kT) {
        const uint64_t k = keys[e];
        used[k >> b] = 1u;
        used[k & mask] = 1u;
    }
}

// labels[newid[v]] = v for every used v (newid = exclusive scan of used)
__global__ void labels_kernel(const uint32_t* __restrict__ used, const uint32_t* __restrict__ newid, Cnt n_,
                              int32_t* __restrict__ labels) {
    const int64_t n = n_.get();
    for (int64_t v = blockIdx.x * (int64_t)kT + threadIdx.x; v < n; v += (int64_t)gridDim.x * kT)
        if (used[v]) labels[newid[v]] = (int32_t)v;
}

__global__ void iota_kernel(int32_t* __restrict__ x, Cnt n_) {
    const int64_t n = n_.get();
    for (int64_t v = blockIdx.x * (int64_t)kT + threadIdx.x; v < n; v += (int64_t)gridDim.x * kT) x[v] = (int32_t)v;
}

// Rewrites keys (row << b | col) through map (NULL = identity) into
// max(map row, map col) << bn | min(...).
__global__ void relabel_kernel(uint64_t* __restrict__ keys, Cnt m_, int b, const uint32_t* __restrict__ map,
                               int bn) {
    const int64_t m = m_.get();
    const uint64_t mask = (1ull << b) - 1;
    for (int64_t e = blockIdx.x * (int64_t)kT + threadIdx.x; e < m; e += (int64_t)gridDim.x * kT) {
        const uint64_t k = keys[e];
        uint32_t r = (uint32_t)(k >> b), c = (uint32_t)(k & mask);
        if (map) {
            r = map[r];
            c = map[c];
        }
        keys[e] = ((uint64_t)max(r, c) << bn) | (uint64_t)min(r, c);
    }
}

// deg[v] += entries of v in the ascending keys [0, m) (row r = k >> b,
// column c = k & mask).  Rows come in runs, so a wave adds each run of its
// 64 keys with one atomic; columns of ids below kDegWindow (where hub
// vertices of skewed graphs concentrate) are counted in LDS first, so a hub
// gets one global atomic per workgroup rather than one per edge.
constexpr uint32_t kDegWindow = 8192;

__global__ __launch_bounds__(kT) void degree_kernel(const uint64_t* __restrict__ keys, Cnt m_, int b,
                                                    uint32_t* __restrict__ deg) {
    __shared__ uint32_t h[kDegWindow];
    for (uint32_t i = threadIdx.x; i < kDegWindow; i += kT) h[i] = 0;
    __syncthreads();
    const int64_t m = m_.get();
    const uint64_t mask = (1ull << b) - 1;
    const uint32_t lane = threadIdx.x & 63;
    for (int64_t base = blockIdx.x * (int64_t)kT + (threadIdx.x & ~63u); base < m; base += (int64_t)gridDim.x * kT) {
        const int64_t e = base + lane;
        const bool ok = e < m;
        const uint64_t k = ok ? keys[e] : ~0ull;
        const uint32_t r = (uint32_t)(k >> b);
        const uint32_t prev = __shfl_up(r, 1), next = __shfl_down(r, 1);
        const bool first = lane == 0 || prev != r;
        const bool last = lane == 63 || next != r;
        const uint64_t starts = __ballot(first);
        const uint32_t run0 = 63 - __clzll(starts & ((2ull << lane) - 1));  // first lane of my run
        if (ok && last) {
            const uint32_t cnt = lane - run0 + 1;
            if (r < kDegWindow) atomicAdd(&h[r], cnt);
            else atomicAdd(&deg[r], cnt);
        }
        if (ok) {
            const uint32_t c = (uint32_t)(k & mask);
            if (c < kDegWindow) atomicAdd(&h[c], 1u);
            else atomicAdd(&deg[c], 1u);
        }
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < kDegWindow; i += kT)
        if (h[i]) atomicAdd(&deg[i], h[i]);
}

// key = top - degree (ascending = descending degree; top is at least the
// largest degree), value = vertex
__global__ void rank_keys_kernel(const uint32_t* __restrict__ deg, Cnt n_, uint32_t top, uint32_t* __restrict__ key,
                                 uint32_t* __restrict__ val) {
    const int64_t n = n_.get();
    for (int64_t v = blockIdx.x * (int64_t)kT + threadIdx.x; v < n; v += (int64_t)gridDim.x * kT) {
        key[v] = top - deg[v];
        val[v] = (uint32_t)v;
    }
}

// key[r] = segment of vertex order[r] (seg = nseg + 1 ascending vertex
// offsets; the last segment whose start is <= the vertex), val[r] = order[r]
__global__ void segment_keys_kernel(const uint32_t* __restrict__ order, int32_t n, const uint32_t* __restrict__ seg,
                                    int nseg, uint32_t* __restrict__ key, uint32_t* __restrict__ val) {
    for (int64_t r = blockIdx.x * (int64_t)kT + threadIdx.x; r < n; r += (int64_t)gridDim.x * kT) {
        const uint32_t v = order[r];
        int lo = 0, hi = nseg;
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (seg[mid] <= v) lo = mid;
            else hi = mid;
        }
        key[r] = (uint32_t)lo;
        val[r] = v;
    }
}

// out[k] = map[at[k]] for k <= nseg (compacted segment starts)
__global__ void gather_kernel(const uint32_t* __restrict__ map, const uint32_t* __restrict__ at, int cnt,
                              uint32_t* __restrict__ out) {
    const int k = blockIdx.x * kT + threadIdx.x;
    if (k < cnt) out[k] = map[at[k]];
}

// rank[order[r]] = r; labels_out[r] = labels_in[order[r]]
__global__ void rank_scatter_kernel(const uint32_t* __restrict__ order, Cnt n_, uint32_t* __restrict__ rank,
                                    const int32_t* __restrict__ labels_in, int32_t* __restrict__ labels_out) {
    const int64_t n = n_.get();
    for (int64_t r = blockIdx.x * (int64_t)kT + threadIdx.x; r < n; r += (int64_t)gridDim.x * kT) {
        const uint32_t v = order[r];
        rank[v] = (uint32_t)r;
        labels_out[r] = labels_in[v];
    }
}

// L from the ascending keys: rowptr[r] = first e with keys[e] >= r << b
// (r = 0..n), col[e] = the key's low b bits, col[m, m + kColPad) = 0; col16
// (when set) gets the low 16 bits of col over the same m + kColPad entries.
__global__ void csr_kernel(const uint64_t* __restrict__ keys, Cnt m_, int b, Cnt n_, uint32_t* __restrict__ rowptr,
                           int32_t* __restrict__ col, uint16_t* __restrict__ col16) {
    const int64_t m = m_.get();
    const int64_t n = n_.get();
    const uint64_t mask = (1ull << b) - 1;
    const int64_t stride = (int64_t)gridDim.x * kT;
    const int64_t t0 = blockIdx.x * (int64_t)kT + threadIdx.x;
    for (int64_t r = t0; r <= n; r += stride) {
        const uint64_t t = (uint64_t)r << b;
        int64_t lo = 0, hi = m;
        while (lo < hi) {
            const int64_t mid = (lo + hi) >> 1;
            if (keys[mid] < t) lo = mid + 1;
            else hi = mid;
        }
        rowptr[r] = (uint32_t)lo;
    }
    for (int64_t e = t0; e < m + (int64_t)kColPad; e += stride) {
        const int32_t c = e < m ? (int32_t)(keys[e] & mask) : 0;
        col[e] = c;
        if (col16) col16[e] = (uint16_t)c;
    }
}

// out: [0] longest row, [1] sum C(d, 2), [2] sum over edges of d(i) + d(j)
// (rows [0, n), edges [0, m) of L; one atomic per workgroup and word)
__global__ __launch_bounds__(kT) void stats_kernel(const uint32_t* __restrict__ rowptr,
                                                   const int32_t* __restrict__ col, Cnt n_, Cnt m_,
                                                   unsigned long long* __restrict__ out) {
    __shared__ uint64_t lds[kT / 64];
    const int64_t n = n_.get();
    const int64_t m = m_.get();
    const int64_t stride = (int64_t)gridDim.x * kT;
    uint32_t mx = 0;
    uint64_t w = 0, work = 0;
    for (int64_t i = blockIdx.x * (int64_t)kT + threadIdx.x; i < n; i += stride) {
        const uint64_t d = rowptr[i + 1] - rowptr[i];
        mx = max(mx, (uint32_t)d);
        w += d * (d - (d > 0)) / 2;
        work += d * d;
    }
    for (int64_t e = blockIdx.x * (int64_t)kT + threadIdx.x; e < m; e += stride) {
        const int32_t j = col[e];
        work += rowptr[j + 1] - rowptr[j];
    }
    for (int o = 32; o > 0; o >>= 1) mx = max(mx, (uint32_t)__shfl_xor((int)mx, o));
    __shared__ uint32_t s_mx[kT / 64];
    if ((threadIdx.x & 63) == 0) s_mx[threadIdx.x >> 6] = mx;
    w = block_sum_u64<kT>(w, lds);
    work = block_sum_u64<kT>(work, lds);
    if (threadIdx.x == 0) {
        for (int k = 1; k < kT / 64; ++k) mx = max(mx, s_mx[k]);
        if (mx) atomicMax(&out[0], (unsigned long long)mx);
        if (w) atomicAdd(&out[1], (unsigned long long)w);
        if (work) atomicAdd(&out[2], (unsigned long long)work);
    }
}

// ------------------------------------------------- capturable form kernels
// Clears a run's words (keeping the ones the host wrote) and opens the id
// range.
__global__ void pipe_reset_kernel(PipeCtl* __restrict__ c) {
    constexpr uint32_t kWords = sizeof(PipeCtl) / 4;
    uint32_t* w = reinterpret_cast<uint32_t*>(c);
    for (uint32_t i = threadIdx.x + kPipeHostWords; i < kWords; i += blockDim.x) w[i] = 0;
    __syncthreads();
    if (threadIdx.x == 0) {
        c->lo = INT_MAX;
        c->hi = -1;
    }
}

// orient_kernel for unchecked input.  Every workgroup resolves n itself (the
// given count, or the largest id + 1) and workgroup 0 records it with the
// refusal bits; pairs with an id outside [0, n) become the sentinel too (the
// run is refused then, but its kernels stay in bounds).
__global__ void orient_checked_kernel(const int32_t* __restrict__ src, const int32_t* __restrict__ dst,
                                      PipeCtl* __restrict__ c, int32_t n_cap, int b, uint64_t sentinel,
                                      uint64_t* __restrict__ keys) {
    const int64_t nnz = c->nnz;
    const int32_t lo = nnz ? c->lo : 0, hi = c->hi;
    int64_t nn = c->n_in > 0 ? (int64_t)c->n_in : (int64_t)hi + 1;
    uint32_t st = 0;
    if (lo < 0) st |= kPipeNegativeId;
    if (hi >= nn) st |= kPipeIdRange;
    if (nn > n_cap) {
        st |= kPipeTooManyIds;
        nn = n_cap;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        c->n = (uint32_t)nn;
        c->status = st;
    }
    const uint32_t n = (uint32_t)nn;
    for (int64_t e = blockIdx.x * (int64_t)kT + threadIdx.x; e < nnz; e += (int64_t)gridDim.x * kT) {
        const uint32_t u = (uint32_t)src[e], v = (uint32_t)dst[e];
        keys[e] = (u == v || u >= n || v >= n) ? sentinel : ((uint64_t)max(u, v) << b) | (uint64_t)min(u, v);
    }
}

__global__ void fill_u32_kernel(uint32_t* __restrict__ p, Cnt n_, uint32_t v) {
    const int64_t n = n_.get();
    for (int64_t i = blockIdx.x * (int64_t)kT + threadIdx.x; i < n; i += (int64_t)gridDim.x * kT) p[i] = v;
}

// The compacted ids: labels (labels[newid[v]] = v for used v) and the keys
// rewritten through newid (monotone, so they stay sorted).  With no edge
// nothing is compacted (labels = identity, nc = n), as in
// canonicalise_device.
__global__ void pipe_compact_ids_kernel(const uint32_t* __restrict__ used, const uint32_t* __restrict__ newid,
                                        PipeCtl* __restrict__ c, int32_t* __restrict__ labels,
                                        uint64_t* __restrict__ keys, int b) {
    const int64_t n = c->n, m = c->m;
    const int64_t stride = (int64_t)gridDim.x * kT, t0 = blockIdx.x * (int64_t)kT + threadIdx.x;
    const bool none = m == 0;
    for (int64_t v = t0; v < n; v += stride) {
        if (none) labels[v] = (int32_t)v;
        else if (used[v]) labels[newid[v]] = (int32_t)v;
    }
    if (none && t0 == 0) c->nc = (uint32_t)n;
    const uint64_t mask = (1ull << b) - 1;
    for (int64_t e = t0; e < m; e += stride) {
        const uint64_t k = keys[e];
        keys[e] = ((uint64_t)newid[k >> b] << b) | (uint64_t)newid[k & mask];
    }
}

// First vertex of L of every graph: its first input id through newid
// (none: no compaction, as when there is no edge).
__global__ void pipe_csegs_kernel(const uint32_t* __restrict__ segs, const uint32_t* __restrict__ newid,
                                  const PipeCtl* __restrict__ c, uint32_t* __restrict__ csegs) {
    const uint32_t k = c->graphs;
    const bool map = newid && c->m != 0;
    for (uint32_t g = threadIdx.x; g <= k; g += blockDim.x) csegs[g] = map ? newid[segs[g]] : segs[g];
}

// rank_keys_kernel within graphs: key = graph << b | (top - degree), so one
// sort orders by (graph, descending degree, id).
__global__ void pipe_rank_keys_kernel(const uint32_t* __restrict__ deg, const PipeCtl* __restrict__ c,
                                      const uint32_t* __restrict__ csegs, uint32_t top, int b,
                                      uint32_t* __restrict__ key, uint32_t* __restrict__ val) {
    const int64_t n = c->nc;
    const uint32_t k = c->graphs;
    for (int64_t v = blockIdx.x * (int64_t)kT + threadIdx.x; v < n; v += (int64_t)gridDim.x * kT) {
        uint32_t lo = 0, hi = k;  // the last graph whose first vertex is <= v
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (csegs[mid] <= (uint32_t)v) lo = mid;
            else hi = mid;
        }
        key[v] = (lo << b) | (top - deg[v]);
        val[v] = (uint32_t)v;
    }
}

// No compaction: labels = identity and nc = n.
__global__ void pipe_identity_kernel(PipeCtl* __restrict__ c, int32_t* __restrict__ labels) {
    const int64_t n = c->n;
    const int64_t t0 = blockIdx.x * (int64_t)kT + threadIdx.x;
    for (int64_t v = t0; v < n; v += (int64_t)gridDim.x * kT) labels[v] = (int32_t)v;
    if (t0 == 0) c->nc = (uint32_t)n;
}

#define BTC_LAUNCH(kernel, n, s, ...)                                                   \
    do {                                                                                \
        hipLaunchKernelGGL(kernel, dim3(grid_for(n)), dim3(kT), 0, s, __VA_ARGS__);     \
        BTC_CHECK_LAUNCH(#kernel);                                                      \
    } while (0)

// Kernels that end in same-address atomics (one per workgroup) run on at
// most kRedBlocks workgroups.
#define BTC_LAUNCH_RED(kernel, n, s, ...)                                                   \
    do {                                                                                \
        hipLaunchKernelGGL(kernel, dim3(std::min(grid_for(n), kRedBlocks)), dim3(kT), 0, s, __VA_ARGS__);     \
        BTC_CHECK_LAUNCH(#kernel);                                                      \
    } while (0)

}  // namespace

void canonicalise_device(const int32_t* src, const int32_t* dst, int64_t nnz, int32_t n, uint32_t flags,
                         hipStream_t s, btc_graph_s* g, const std::vector<int64_t>* segs,
                         std::vector<int64_t>* csegs) {
    BTC_REQUIRE(nnz < (int64_t)0xFFFFFFFFLL, BTC_ERR_OVERFLOW, "nnz %lld exceeds 2^32 - 1 pairs", (long long)nnz);
    const int nseg = segs ? (int)segs->size() - 1 : 0;
    if (segs) {
        BTC_REQUIRE(nseg >= 1 && csegs && (*segs)[0] == 0 && segs->back() == n, BTC_ERR_ARG,
                    "bad segments (%d segments over %d vertices)", nseg, n);
        for (int k = 0; k < nseg; ++k) BTC_REQUIRE((*segs)[k] <= (*segs)[k + 1], BTC_ERR_ARG, "segments not ascending");
    }
    // id range (and n when not given)
    int32_t h_lohi[2] = {INT_MAX, -1};
    {
        TmpBuf<int32_t> lohi;
        lohi.alloc(2, s);
        BTC_HIP(hipMemcpyAsync(lohi.p, h_lohi, 8, hipMemcpyHostToDevice, s));
        if (nnz) BTC_LAUNCH_RED(id_range_kernel, nnz, s, src, dst, nnz, lohi.p);
        BTC_HIP(hipMemcpyAsync(h_lohi, lohi.p, 8, hipMemcpyDeviceToHost, s));
        BTC_HIP(hipStreamSynchronize(s));
    }
    if (nnz == 0) h_lohi[0] = 0;
    BTC_REQUIRE(h_lohi[0] >= 0, BTC_ERR_ARG, "negative vertex id %d", h_lohi[0]);
    if (n <= 0) n = h_lohi[1] + 1;
    BTC_REQUIRE(h_lohi[1] < n, BTC_ERR_ARG, "vertex id %d >= n = %d", h_lohi[1], n);
    g->n_input = n;
    g->flags = flags;
    const int b = id_bits(n);
    const uint64_t sentinel = (1ull << (2 * b)) - 1;

    // oriented keys, sorted, deduplicated
    TmpBuf<uint64_t> ka, kb;
    ka.alloc((size_t)nnz, s);
    kb.alloc((size_t)nnz, s);
    if (nnz) BTC_LAUNCH(orient_kernel, nnz, s, src, dst, nnz, b, sentinel, ka.p);
    uint64_t* sorted = sort_u64(ka.p, kb.p, (size_t)nnz, 2 * b, s);
    uint64_t* other = sorted == ka.p ? kb.p : ka.p;
    int64_t m = 0;
    {
        TmpBuf<uint32_t> flag, pos;
        flag.alloc((size_t)nnz, s);
        pos.alloc((size_t)nnz + 1, s);
        if (nnz) BTC_LAUNCH(first_flag_kernel, nnz, s, sorted, nnz, sentinel, flag.p);
        exclusive_scan_u32(flag.p, pos.p, (size_t)nnz, s);
        uint32_t hm = 0;
        BTC_HIP(hipMemcpyAsync(&hm, pos.p + nnz, 4, hipMemcpyDeviceToHost, s));
        if (nnz) BTC_LAUNCH(compact_kernel, nnz, s, sorted, flag.p, pos.p, nnz, other);
        BTC_HIP(hipStreamSynchronize(s));
        m = hm;
    }
    uint64_t* keys = other;      // m unique oriented pairs, ascending
    uint64_t* spare = sorted;

    // order-preserving compaction of the used ids (none when no edge)
    int32_t nc = n;
    TmpBuf<uint32_t> newid;
    DevBuf<int32_t> labels;
    if ((flags & BTC_CANON_COMPACT) && m > 0) {
        TmpBuf<uint32_t> used;
        used.alloc((size_t)n, s);
        newid.alloc((size_t)n + 1, s);
        BTC_HIP(hipMemsetAsync(used.p, 0, (size_t)n * 4, s));
        BTC_LAUNCH(mark_used_kernel, m, s, keys, m, b, used.p);
        exclusive_scan_u32(used.p, newid.p, (size_t)n, s);
        uint32_t hn = 0;
        BTC_HIP(hipMemcpyAsync(&hn, newid.p + n, 4, hipMemcpyDeviceToHost, s));
        BTC_HIP(hipStreamSynchronize(s));
        nc = (int32_t)hn;
        labels.alloc((size_t)nc);
        BTC_LAUNCH(labels_kernel, n, s, used.p, newid.p, n, labels.p);
    } else {
        labels.alloc((size_t)n);
        BTC_LAUNCH(iota_kernel, n, s, labels.p, (int64_t)n);
    }
    const int bn = id_bits(nc);
    if (m) BTC_LAUNCH(relabel_kernel, m, s, keys, m, b, newid.p, bn);  // monotone: order kept
    // segment starts after compaction (the relabel is monotone, so every
    // segment stays one contiguous range)
    TmpBuf<uint32_t> dseg;
    if (segs) {
        std::vector<uint32_t> at(segs->begin(), segs->end()), cs(at.size());
        dseg.alloc(at.size(), s);
        BTC_HIP(hipMemcpyAsync(dseg.p, at.data(), at.size() * 4, hipMemcpyHostToDevice, s));
        if (newid.p) {
            TmpBuf<uint32_t> tmp;
            tmp.alloc(at.size(), s);
            hipLaunchKernelGGL(gather_kernel, dim3((unsigned)((at.size() + kT - 1) / kT)), dim3(kT), 0, s, newid.p,
                               dseg.p, (int)at.size(), tmp.p);
            BTC_CHECK_LAUNCH("gather_kernel");
            BTC_HIP(hipMemcpyAsync(dseg.p, tmp.p, at.size() * 4, hipMemcpyDeviceToDevice, s));
        }
        BTC_HIP(hipMemcpyAsync(cs.data(), dseg.p, cs.size() * 4, hipMemcpyDeviceToHost, s));
        BTC_HIP(hipStreamSynchronize(s));
        csegs->assign(cs.begin(), cs.end());
    }

    if (flags & BTC_CANON_DEGREE_ORDER) {
        TmpBuf<uint32_t> deg, k0, v0, k1, v1, rank;
        DevBuf<int32_t> lab2;
        deg.alloc((size_t)nc, s);
        k0.alloc((size_t)nc, s);
        v0.alloc((size_t)nc, s);
        k1.alloc((size_t)nc, s);
        v1.alloc((size_t)nc, s);
        rank.alloc((size_t)nc, s);
        lab2.alloc((size_t)nc);
        BTC_HIP(hipMemsetAsync(deg.p, 0, (size_t)nc * 4, s));
        if (m) {
            hipLaunchKernelGGL(degree_kernel, dim3(std::min(grid_for(m), kRedBlocks)), dim3(kT), 0, s, keys, m, bn, deg.p);
            BTC_CHECK_LAUNCH("degree_kernel");
        }
        BTC_LAUNCH(rank_keys_kernel, nc, s, deg.p, nc, 0xFFFFFFFFu, k0.p, v0.p);
        bool alt = sort_pairs_u32(k0.p, v0.p, k1.p, v1.p, (size_t)nc, 32, s);
        if (segs) {
            // then a stable sort by segment: (segment, -degree, id)
            uint32_t* ord = alt ? v1.p : v0.p;
            uint32_t* kk = alt ? k0.p : k1.p;  // free key buffer
            uint32_t* vv = alt ? v0.p : v1.p;  // free value buffer
            BTC_LAUNCH(segment_keys_kernel, nc, s, ord, nc, dseg.p, nseg, kk, vv);
            int sb = 1;
            while (sb < 32 && (1ll << sb) < (int64_t)nseg) ++sb;
            // sort (kk, vv) using (other key buffer, ord) as the spare pair
            uint32_t* kk2 = alt ? k1.p : k0.p;
            const bool alt2 = sort_pairs_u32(kk, vv, kk2, ord, (size_t)nc, sb, s);
            alt = alt2 ? (ord == v1.p) : (vv == v1.p);
        }
        BTC_LAUNCH(rank_scatter_kernel, nc, s, alt ? v1.p : v0.p, nc, rank.p, labels.p, lab2.p);
        if (m) BTC_LAUNCH(relabel_kernel, m, s, keys, m, bn, rank.p, bn);
        keys = sort_u64(keys, spare, (size_t)m, 2 * bn, s);
        labels = std::move(lab2);
    }

    // CSR of L
    g->n = nc;
    g->m = m;
    g->d_rowptr.alloc((size_t)nc + 1);
    g->d_col.alloc((size_t)m + kColPad);
    BTC_LAUNCH(csr_kernel, std::max<int64_t>((int64_t)nc + 1, m + (int64_t)kColPad), s, keys, m, bn, nc,
               g->d_rowptr.p, g->d_col.p, (uint16_t*)nullptr);
    g->d_labels = std::move(labels);

    // derived statistics (graph_finalize on the device)
    TmpBuf<unsigned long long> st;
    st.alloc(3, s);
    BTC_HIP(hipMemsetAsync(st.p, 0, 24, s));
    if (nc) BTC_LAUNCH_RED(stats_kernel, std::max<int64_t>(nc, m), s, g->d_rowptr.p, g->d_col.p, nc, m, st.p);
    unsigned long long h[3] = {0, 0, 0};
    BTC_HIP(hipMemcpyAsync(h, st.p, 24, hipMemcpyDeviceToHost, s));
    BTC_HIP(hipStreamSynchronize(s));
    g->max_row = (int32_t)h[0];
    g->wedges = (int64_t)h[1];
    g->merge_work = (int64_t)h[2];
    g->on_host = false;
}

namespace {

// Grid of a capacity-sized launch: its workgroups stride over the live
// count, so a few per CU suffice.
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

int32_t* canonicalise_capacity(const CanonCap& c, hipStream_t s) {
    PipeCtl* ctl = c.ctl;
    const int64_t E = c.nnz_cap, N = c.n_cap;
    const int b = id_bits(N);
    const uint64_t sentinel = (1ull << (2 * b)) - 1;
    const Cnt nnz = Cnt::dev(&ctl->nnz), n = Cnt::dev(&ctl->n), m = Cnt::dev(&ctl->m), nc = Cnt::dev(&ctl->nc);
    hipLaunchKernelGGL(pipe_reset_kernel, dim3(1), dim3(64), 0, s, ctl);
    BTC_CHECK_LAUNCH("pipe_reset_kernel");
    CAP_LAUNCH_RED(id_range_kernel, E, s, c.src, c.dst, nnz, &ctl->lo);

    // oriented keys, sorted, deduplicated
    CAP_LAUNCH(orient_checked_kernel, E, s, c.src, c.dst, ctl, c.n_cap, b, sentinel, c.ka);
    uint64_t* sorted = dsort_u64(c.ka, c.kb, (size_t)E, &ctl->nnz, 2 * b, c.ws, s);
    uint64_t* keys = sorted == c.ka ? c.kb : c.ka;
    uint64_t* spare = sorted;
    CAP_LAUNCH(first_flag_kernel, E, s, sorted, nnz, sentinel, c.flag);
    dscan_u32(c.flag, c.pos, (size_t)E, &ctl->nnz, &ctl->m, c.ws, s);
    CAP_LAUNCH(compact_kernel, E, s, sorted, c.flag, c.pos, nnz, keys);

    // order-preserving compaction of the used ids (ids keep b bits); the
    // graphs' ranges follow it
    const bool groups = c.max_graphs > 1;
    int32_t* labels = c.lab0;
    if (c.flags & BTC_CANON_COMPACT) {
        CAP_LAUNCH(fill_u32_kernel, N, s, c.used, n, 0u);
        CAP_LAUNCH(mark_used_kernel, E, s, keys, m, b, c.used);
        dscan_u32(c.used, c.newid, (size_t)N, &ctl->n, &ctl->nc, c.ws, s);
        CAP_LAUNCH(pipe_compact_ids_kernel, std::max(N, E), s, c.used, c.newid, ctl, labels, keys, b);
    } else {
        CAP_LAUNCH(pipe_identity_kernel, N, s, ctl, labels);
    }
    if (groups) {
        hipLaunchKernelGGL(pipe_csegs_kernel, dim3(1), dim3(64), 0, s, c.segs,
                           (c.flags & BTC_CANON_COMPACT) ? c.newid : nullptr, ctl, c.csegs);
        BTC_CHECK_LAUNCH("pipe_csegs_kernel");
    }

    if (c.flags & BTC_CANON_DEGREE_ORDER) {
        // keys N - 1 - degree < 2^b: descending degree, ties by id
        CAP_LAUNCH(fill_u32_kernel, N, s, c.deg, nc, 0u);
        hipLaunchKernelGGL(degree_kernel, dim3(std::min(cap_grid(E), kCapRedBlocks)), dim3(kT), 0, s, keys, m, b, c.deg);
        BTC_CHECK_LAUNCH("degree_kernel");
        int kb = b;  // key bits: degree rank, then (several graphs) the graph
        if (groups) {
            int gb = 1;
            while ((1 << gb) < c.max_graphs) ++gb;
            BTC_REQUIRE(b + gb <= 32, BTC_ERR_OVERFLOW, "%d ids and %d graphs exceed 32-bit relabel keys", c.n_cap,
                        c.max_graphs);
            CAP_LAUNCH(pipe_rank_keys_kernel, N, s, c.deg, ctl, c.csegs, (uint32_t)(N - 1), b, c.k0, c.v0);
            kb = b + gb;
        } else {
            CAP_LAUNCH(rank_keys_kernel, N, s, c.deg, nc, (uint32_t)(N - 1), c.k0, c.v0);
        }
        const bool alt = dsort_pairs_u32(c.k0, c.v0, c.k1, c.v1, (size_t)N, &ctl->nc, kb, c.ws, s);
        CAP_LAUNCH(rank_scatter_kernel, N, s, alt ? c.v1 : c.v0, nc, c.rank, labels, c.lab1);
        CAP_LAUNCH(relabel_kernel, E, s, keys, m, b, c.rank, b);
        keys = dsort_u64(keys, spare, (size_t)E, &ctl->m, 2 * b, c.ws, s);
        labels = c.lab1;
    }

    // CSR of L (with its 16-bit copy) and its statistics
    CAP_LAUNCH(csr_kernel, std::max(N + 1, E + (int64_t)kColPad), s, keys, m, b, nc, c.rowptr, c.col, c.col16);
    CAP_LAUNCH_RED(stats_kernel, std::max(N, E), s, c.rowptr, c.col, nc, m, ctl->rs);
    return labels;
}

void ensure_host(btc_graph_s* g) {
    if (g->on_host) return;
    std::vector<uint32_t> rp((size_t)g->n + 1);
    std::vector<int32_t> col((size_t)g->m), labels((size_t)g->n);
    BTC_HIP(hipMemcpy(rp.data(), g->d_rowptr.p, rp.size() * 4, hipMemcpyDeviceToHost));
    if (g->m) BTC_HIP(hipMemcpy(col.data(), g->d_col.p, (size_t)g->m * 4, hipMemcpyDeviceToHost));
    if (g->n) BTC_HIP(hipMemcpy(labels.data(), g->d_labels.p, (size_t)g->n * 4, hipMemcpyDeviceToHost));
    g->rowptr.assign(rp.begin(), rp.end());
    g->col.swap(col);
    g->labels.swap(labels);
    g->on_host = true;
}

}  // namespace btcx

extern "C" {

int btc_graph_from_coo_device(const int32_t* src, const int32_t* dst, int64_t nnz, int32_t n, uint32_t flags,
                              void* stream, btc_graph* out) {
    return btcx::guard([&] {
        BTC_REQUIRE(out, BTC_ERR_ARG, "out is NULL");
        *out = nullptr;
        BTC_REQUIRE(nnz >= 0, BTC_ERR_ARG, "nnz %lld < 0", (long long)nnz);
        BTC_REQUIRE(nnz == 0 || (src && dst), BTC_ERR_ARG, "src/dst is NULL");
        BTC_REQUIRE((flags & ~BTC_CANON_DEFAULT) == 0, BTC_ERR_ARG, "unknown flags 0x%x", flags);
        auto g = std::make_unique<btc_graph_s>();
        btcx::canonicalise_device(src, dst, nnz, n, flags, btcx::as_stream(stream), g.get(), nullptr, nullptr);
        *out = g.release();
    });
}

}  // extern "C"
