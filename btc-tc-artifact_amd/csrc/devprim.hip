// Sort and scan whose element count is read from device memory when the
// kernels run.  They enqueue kernels only -- no allocation, no memset node,
// no host synchronisation -- so a stream capture of them replays for any
// count up to the capacity their buffers were sized for (see prim.h).
//
// Scan: tile sums, one workgroup scanning the tile sums, then each tile
// applies its offset.  Sort: least-significant-digit radix sort, 8 bits a
// pass; each pass takes a digit histogram per tile (digit-major, so one
// flat scan gives every (digit, tile) its output offset), then ranks each
// tile stably in LDS (rocPRIM's block_radix_rank, match form: per-wave
// ballots of equal digits) and writes it out run by run.
#include <rocprim/block/block_radix_rank.hpp>

#include "common.h"
#include "device_util.h"
#include "prim.h"

namespace btcx {
namespace {

constexpr int kT = 256;
constexpr int kScanPer = 16;
constexpr uint32_t kScanTile = kT * kScanPer;
constexpr int kSortPer = 8;
constexpr uint32_t kSortTile = kT * kSortPer;
constexpr int kDigitBits = 8;
constexpr uint32_t kDigits = 1u << kDigitBits;
static_assert(kDigits == (uint32_t)kT, "one digit per thread");

size_t tiles(size_t cap, size_t tile) { return std::max<size_t>(1, (cap + tile - 1) / tile); }

unsigned grid_of(size_t cap, size_t tile) { return (unsigned)std::min<size_t>(tiles(cap, tile), 65535); }

// Exclusive scan of one value per thread of a 256-thread workgroup;
// *total gets the workgroup's sum.
template <class T>
__device__ T block_scan_ex(T v, T* lds, T* total) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    T x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const T y = __shfl_up(x, o);
        if (lane >= o) x += y;
    }
    if (lane == 63) lds[wid] = x;
    __syncthreads();
    T base = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < kT / 64; ++w) {
        if (w < wid) base += lds[w];
        tot += lds[w];
    }
    __syncthreads();
    *total = tot;
    return base + x - v;
}

template <class T>
__global__ __launch_bounds__(kT) void dscan_tiles(const T* __restrict__ in, const uint32_t* __restrict__ d_n,
                                                  T* __restrict__ sums) {
    __shared__ T lds[kT / 64];
    const uint32_t n = *d_n;
    const uint32_t nt = (n + kScanTile - 1) / kScanTile;
    for (uint32_t t = blockIdx.x; t < nt; t += gridDim.x) {
        const size_t base = (size_t)t * kScanTile + (size_t)threadIdx.x * kScanPer;
        T v = 0;
#pragma unroll
        for (int k = 0; k < kScanPer; ++k)
            if (base + k < n) v += in[base + k];
        T tot;
        block_scan_ex(v, lds, &tot);
        if (threadIdx.x == 0) sums[t] = tot;
    }
}

// One workgroup: tile sums -> tile offsets; out[n] and *d_total = the total.
template <class T>
__global__ __launch_bounds__(kT) void dscan_sums(T* __restrict__ sums, const uint32_t* __restrict__ d_n,
                                                 T* __restrict__ out, T* __restrict__ d_total) {
    __shared__ T lds[kT / 64];
    const uint32_t n = *d_n;
    const uint32_t nt = (n + kScanTile - 1) / kScanTile;
    T carry = 0;
    for (uint32_t b0 = 0; b0 < nt; b0 += kT) {
        const uint32_t b = b0 + threadIdx.x;
        const T v = b < nt ? sums[b] : T(0);
        T tot;
        const T ex = block_scan_ex(v, lds, &tot);
        if (b < nt) sums[b] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0) {
        out[n] = carry;
        if (d_total) *d_total = carry;
    }
}

template <class T>
__global__ __launch_bounds__(kT) void dscan_apply(const T* __restrict__ in, const uint32_t* __restrict__ d_n,
                                                  const T* __restrict__ sums, T* __restrict__ out) {
    __shared__ T lds[kT / 64];
    const uint32_t n = *d_n;
    const uint32_t nt = (n + kScanTile - 1) / kScanTile;
    for (uint32_t t = blockIdx.x; t < nt; t += gridDim.x) {
        const size_t base = (size_t)t * kScanTile + (size_t)threadIdx.x * kScanPer;
        T v[kScanPer];
        T sum = 0;
#pragma unroll
        for (int k = 0; k < kScanPer; ++k) {
            v[k] = base + k < n ? in[base + k] : T(0);
            sum += v[k];
        }
        T tot;
        T run = sums[t] + block_scan_ex(sum, lds, &tot);
#pragma unroll
        for (int k = 0; k < kScanPer; ++k) {
            if (base + k < n) out[base + k] = run;
            run += v[k];
        }
    }
}

template <class T>
void dscan_t(const T* in, T* out, size_t cap, const uint32_t* d_n, T* d_total, void* ws, hipStream_t s) {
    T* sums = static_cast<T*>(ws);
    const unsigned g = grid_of(cap, kScanTile);
    hipLaunchKernelGGL(dscan_tiles<T>, dim3(g), dim3(kT), 0, s, in, d_n, sums);
    BTC_CHECK_LAUNCH("dscan_tiles");
    hipLaunchKernelGGL(dscan_sums<T>, dim3(1), dim3(kT), 0, s, sums, d_n, out, d_total);
    BTC_CHECK_LAUNCH("dscan_sums");
    hipLaunchKernelGGL(dscan_apply<T>, dim3(g), dim3(kT), 0, s, in, d_n, sums, out);
    BTC_CHECK_LAUNCH("dscan_apply");
}

// ------------------------------------------------------------------ sort
// ctl[0] = sort tiles of the count, ctl[1] = digits x tiles (the length of
// the digit-count array a pass scans).
__global__ void dsort_setup(const uint32_t* __restrict__ d_n, uint32_t* __restrict__ ctl) {
    if (threadIdx.x == 0) {
        const uint32_t nt = (*d_n + kSortTile - 1) / kSortTile;
        ctl[0] = nt;
        ctl[1] = nt * kDigits;
    }
}

template <class K>
__device__ __forceinline__ uint32_t digit_of(K k, int shift) {
    return (uint32_t)(k >> shift) & (kDigits - 1);
}

// counts[d * nt + t] = keys of tile t with digit d.
template <class K>
__global__ __launch_bounds__(kT) void dsort_hist(const K* __restrict__ keys, const uint32_t* __restrict__ d_n,
                                                 const uint32_t* __restrict__ ctl, int shift,
                                                 uint32_t* __restrict__ counts) {
    __shared__ uint32_t h[kDigits];
    const uint32_t n = *d_n, nt = ctl[0];
    for (uint32_t t = blockIdx.x; t < nt; t += gridDim.x) {
        h[threadIdx.x] = 0;
        __syncthreads();
        const size_t base = (size_t)t * kSortTile;
#pragma unroll
        for (int k = 0; k < kSortPer; ++k) {
            const size_t e = base + (size_t)k * kT + threadIdx.x;
            if (e < n) atomicAdd(&h[digit_of(keys[e], shift)], 1u);
        }
        __syncthreads();
        counts[(size_t)threadIdx.x * nt + t] = h[threadIdx.x];
        __syncthreads();
    }
}

// Stable scatter of one pass.  Key i of a tile sits at index
// wave * 64 * kSortPer + item * 64 + lane, the order in which the match
// ranking breaks ties, so equal digits keep their input order; padding keys
// (all bits set) rank after every real key of the tile.
template <class K, bool V>
__global__ __launch_bounds__(kT) void dsort_scatter(const K* __restrict__ kin, const uint32_t* __restrict__ vin,
                                                    K* __restrict__ kout, uint32_t* __restrict__ vout,
                                                    const uint32_t* __restrict__ d_n,
                                                    const uint32_t* __restrict__ ctl, int shift,
                                                    const uint32_t* __restrict__ offs) {
    using Rank = rocprim::block_radix_rank<kT, kDigitBits, rocprim::block_radix_rank_algorithm::match>;
    static_assert(Rank::digits_per_thread == 1, "one digit per thread");
    __shared__ typename Rank::storage_type rank_lds;
    __shared__ K sk[kSortTile];
    __shared__ uint32_t sv[V ? kSortTile : 1];
    __shared__ uint32_t s_pre[kDigits], s_out[kDigits];
    const uint32_t n = *d_n, nt = ctl[0];
    const uint32_t lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    for (uint32_t t = blockIdx.x; t < nt; t += gridDim.x) {
        const size_t base = (size_t)t * kSortTile;
        const uint32_t valid = (uint32_t)std::min<size_t>(kSortTile, n - base);
        K k[kSortPer];
        uint32_t v[kSortPer], r[kSortPer];
#pragma unroll
        for (int i = 0; i < kSortPer; ++i) {
            const uint32_t idx = wid * (64 * kSortPer) + i * 64 + lane;
            k[i] = idx < valid ? kin[base + idx] : ~K(0);
            if constexpr (V) v[i] = idx < valid ? vin[base + idx] : 0u;
        }
        uint32_t pre[1], cnt[1];
        Rank().rank_keys(k, r, rank_lds, [shift](const K& x) { return digit_of(x, shift); }, pre, cnt);
#pragma unroll
        for (int i = 0; i < kSortPer; ++i) {
            sk[r[i]] = k[i];
            if constexpr (V) sv[r[i]] = v[i];
        }
        s_pre[threadIdx.x] = pre[0];
        s_out[threadIdx.x] = offs[(size_t)threadIdx.x * nt + t];
        __syncthreads();
        for (uint32_t idx = threadIdx.x; idx < valid; idx += kT) {
            const K x = sk[idx];
            const uint32_t d = digit_of(x, shift);
            const uint32_t pos = s_out[d] + idx - s_pre[d];
            if (pos >= n) continue;  // unreachable when the counts agree; keeps a bad count in bounds
            kout[pos] = x;
            if constexpr (V) vout[pos] = sv[idx];
        }
        __syncthreads();
    }
}

// Workspace: [0, 256) control words, then the digit counts and their
// offsets (digits x tiles + 1 each), then the scan's tile sums.
struct SortWs {
    uint32_t* ctl;
    uint32_t* counts;
    uint32_t* offs;
    void* scan;
};

SortWs sort_ws(void* ws, size_t cap) {
    const size_t nc = tiles(cap, kSortTile) * kDigits + 1;
    auto* b = static_cast<unsigned char*>(ws);
    SortWs w;
    w.ctl = reinterpret_cast<uint32_t*>(b);
    w.counts = reinterpret_cast<uint32_t*>(b + 256);
    w.offs = w.counts + nc;
    const size_t at = (256 + 2 * nc * 4 + 255) & ~size_t(255);
    w.scan = b + at;
    return w;
}

// Returns whether the result is in (kt, vt).
template <class K, bool V>
bool dsort_t(K* keys, uint32_t* vals, K* kt, uint32_t* vt, size_t cap, const uint32_t* d_n, int end_bit, void* ws,
             hipStream_t s) {
    const int passes = (std::max(end_bit, 0) + kDigitBits - 1) / kDigitBits;
    if (passes == 0) return false;
    const SortWs w = sort_ws(ws, cap);
    const size_t ncap = tiles(cap, kSortTile) * kDigits;
    const unsigned g = grid_of(cap, kSortTile);
    hipLaunchKernelGGL(dsort_setup, dim3(1), dim3(64), 0, s, d_n, w.ctl);
    BTC_CHECK_LAUNCH("dsort_setup");
    K* ki = keys;
    K* ko = kt;
    uint32_t* vi = vals;
    uint32_t* vo = vt;
    for (int p = 0; p < passes; ++p) {
        const int shift = p * kDigitBits;
        hipLaunchKernelGGL(dsort_hist<K>, dim3(g), dim3(kT), 0, s, ki, d_n, w.ctl, shift, w.counts);
        BTC_CHECK_LAUNCH("dsort_hist");
        dscan_t<uint32_t>(w.counts, w.offs, ncap, w.ctl + 1, nullptr, w.scan, s);
        hipLaunchKernelGGL((dsort_scatter<K, V>), dim3(g), dim3(kT), 0, s, ki, vi, ko, vo, d_n, w.ctl, shift,
                           w.offs);
        BTC_CHECK_LAUNCH("dsort_scatter");
        std::swap(ki, ko);
        std::swap(vi, vo);
    }
    return passes & 1;
}

}  // namespace

size_t dscan_ws_bytes(size_t cap) { return tiles(cap, kScanTile) * 8 + 256; }

size_t dsort_ws_bytes(size_t cap) {
    const size_t nc = tiles(cap, kSortTile) * kDigits + 1;
    return ((256 + 2 * nc * 4 + 255) & ~size_t(255)) + dscan_ws_bytes(nc);
}

void dscan_u32(const uint32_t* in, uint32_t* out, size_t cap, const uint32_t* d_n, uint32_t* d_total, void* ws,
               hipStream_t s) {
    dscan_t<uint32_t>(in, out, cap, d_n, d_total, ws, s);
}

void dscan_u64(const uint64_t* in, uint64_t* out, size_t cap, const uint32_t* d_n, uint64_t* d_total, void* ws,
               hipStream_t s) {
    dscan_t<uint64_t>(in, out, cap, d_n, d_total, ws, s);
}

uint64_t* dsort_u64(uint64_t* keys, uint64_t* tmp, size_t cap, const uint32_t* d_n, int end_bit, void* ws,
                    hipStream_t s) {
    return dsort_t<uint64_t, false>(keys, nullptr, tmp, nullptr, cap, d_n, end_bit, ws, s) ? tmp : keys;
}

bool dsort_pairs_u32(uint32_t* keys, uint32_t* vals, uint32_t* kt, uint32_t* vt, size_t cap, const uint32_t* d_n,
                     int end_bit, void* ws, hipStream_t s) {
    return dsort_t<uint32_t, true>(keys, vals, kt, vt, cap, d_n, end_bit, ws, s);
}

}  // namespace btcx
