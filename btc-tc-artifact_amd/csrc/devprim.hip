// Sort and scan whose element count is read from device memory when the
// kernels run.  They enqueue kernels only -- no allocation, no memset node,
// no host synchronisation -- so a stream capture of them replays for any
// count up to the capacity their buffers were sized for (see prim.h).
//
// Scan: tile sums, one workgroup scanning the tile sums, then each tile
// applies its offset.  Sort: least-significant-digit radix sort, 8 bits a
// pass, in one sweep per pass: one kernel takes the digit totals of every
// pass in a single read of the keys; each pass kernel then ranks a tile
// stably in LDS (rocPRIM's block_radix_rank, match form: per-wave ballots
// of equal digits) and finds the tile's offset for each digit by
// decoupled look-back over the tiles before it -- 2 + passes kernels a
// sort.
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

// A scan's tile sums start past the first 256 bytes of its workspace, which
// hold a sort's control words when the two share one workspace.
constexpr size_t kScanSumsAt = 256;

template <class T>
void dscan_t(const T* in, T* out, size_t cap, const uint32_t* d_n, T* d_total, void* ws, hipStream_t s) {
    T* sums = reinterpret_cast<T*>(static_cast<unsigned char*>(ws) + kScanSumsAt);
    const unsigned g = grid_of(cap, kScanTile);
    hipLaunchKernelGGL(dscan_tiles<T>, dim3(g), dim3(kT), 0, s, in, d_n, sums);
    BTC_CHECK_LAUNCH("dscan_tiles");
    hipLaunchKernelGGL(dscan_sums<T>, dim3(1), dim3(kT), 0, s, sums, d_n, out, d_total);
    BTC_CHECK_LAUNCH("dscan_sums");
    hipLaunchKernelGGL(dscan_apply<T>, dim3(g), dim3(kT), 0, s, in, d_n, sums, out);
    BTC_CHECK_LAUNCH("dscan_apply");
}

// ------------------------------------------------------------------ sort
// Workspace control words: [0] tiles of the count, [1 + p] pass p's tile
// ticket, [kMaxPasses + 1] set when a look-back ran past its spin limit.
constexpr int kMaxPasses = 8;
constexpr uint32_t kCtlWords = 64;  // 256 bytes
constexpr uint32_t kErrWord = kMaxPasses + 1;
static_assert(kCtlWords * 4 <= kScanSumsAt, "a scan sharing the workspace leaves the control words alone");
// Look-back status of (tile, digit): flag << 32 | count; flag 1 = the
// tile's own count, 2 = inclusive of every earlier tile.
constexpr uint64_t kAgg = 1ull << 32, kInc = 2ull << 32;
constexpr uint32_t kSpinMax = 1u << 22;

template <class K>
__device__ __forceinline__ uint32_t digit_of(K k, int shift) {
    return (uint32_t)(k >> shift) & (kDigits - 1);
}

// One workgroup: clears the digit totals and the tickets, counts tiles
// (the error word stays set once set).
__global__ __launch_bounds__(kT) void dsort_setup(const uint32_t* __restrict__ d_n, uint32_t* __restrict__ ctl,
                                                  uint32_t* __restrict__ totals, int passes) {
    for (uint32_t i = threadIdx.x; i < (uint32_t)passes * kDigits; i += kT) totals[i] = 0;
    if (threadIdx.x < kErrWord) ctl[threadIdx.x] = threadIdx.x == 0 ? (*d_n + kSortTile - 1) / kSortTile : 0u;
}

// Digit totals of every pass in one read of the keys (per tile in LDS,
// then one atomic per non-zero (pass, digit)), and the tile's look-back
// statuses of every pass cleared.
template <class K>
__global__ __launch_bounds__(kT) void dsort_hist(const K* __restrict__ keys, const uint32_t* __restrict__ d_n,
                                                 const uint32_t* __restrict__ ctl, int passes,
                                                 uint32_t* __restrict__ totals, uint64_t* __restrict__ status,
                                                 size_t status_pass) {
    __shared__ uint32_t h[kMaxPasses][kDigits];
    const uint32_t n = *d_n, nt = ctl[0];
    for (uint32_t t = blockIdx.x; t < nt; t += gridDim.x) {
        for (int p = 0; p < passes; ++p) h[p][threadIdx.x] = 0;
        __syncthreads();
        const size_t base = (size_t)t * kSortTile;
#pragma unroll
        for (int k = 0; k < kSortPer; ++k) {
            const size_t e = base + (size_t)k * kT + threadIdx.x;
            if (e < n) {
                const K x = keys[e];
                for (int p = 0; p < passes; ++p) atomicAdd(&h[p][digit_of(x, p * kDigitBits)], 1u);
            }
        }
        __syncthreads();
        for (int p = 0; p < passes; ++p) {
            const uint32_t c = h[p][threadIdx.x];
            if (c) atomicAdd(&totals[p * kDigits + threadIdx.x], c);
            status[(size_t)p * status_pass + (size_t)t * kDigits + threadIdx.x] = 0;
        }
        __syncthreads();
    }
}

__device__ __forceinline__ uint64_t load_status(const uint64_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void store_status(uint64_t* p, uint64_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// One stable pass.  Workgroups take tiles in ticket order; a tile ranks its
// keys in LDS (rocPRIM's block_radix_rank, match form), publishes its count
// of each digit and looks back over earlier tiles for the digit's offset
// (thread d owns digit d), then writes its keys out run by run.  Key i of a
// tile sits at index wave * 64 * kSortPer + item * 64 + lane, the order in
// which the match ranking breaks ties, so equal digits keep their input
// order; padding keys (all bits set) rank after every real key of the tile.
// Every look-back ends: a ticket is only taken by a running workgroup,
// which publishes before it waits on anything.
template <class K, bool V>
__global__ __launch_bounds__(kT) void dsort_onesweep(const K* __restrict__ kin, const uint32_t* __restrict__ vin,
                                                     K* __restrict__ kout, uint32_t* __restrict__ vout,
                                                     const uint32_t* __restrict__ d_n, uint32_t* __restrict__ ctl,
                                                     int pass, const uint32_t* __restrict__ totals,
                                                     uint64_t* __restrict__ status) {
    using Rank = rocprim::block_radix_rank<kT, kDigitBits, rocprim::block_radix_rank_algorithm::match>;
    static_assert(Rank::digits_per_thread == 1, "one digit per thread");
    __shared__ typename Rank::storage_type rank_lds;
    __shared__ K sk[kSortTile];
    __shared__ uint32_t sv[V ? kSortTile : 1];
    __shared__ uint32_t s_pre[kDigits], s_out[kDigits];
    __shared__ uint32_t s_scan[kT / 64];
    __shared__ uint32_t s_tile;
    const int shift = pass * kDigitBits;
    const uint32_t n = *d_n, nt = ctl[0];
    const uint32_t lane = threadIdx.x & 63, wid = threadIdx.x >> 6, d = threadIdx.x;
    uint32_t all;
    const uint32_t dstart = block_scan_ex(totals[pass * kDigits + d], s_scan, &all);
    for (;;) {
        if (threadIdx.x == 0) s_tile = atomicAdd(&ctl[1 + pass], 1u);
        __syncthreads();
        const uint32_t t = s_tile;
        if (t >= nt) break;
        const size_t base = (size_t)t * kSortTile;
        const uint32_t valid = (uint32_t)min((size_t)kSortTile, (size_t)n - base);
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
        uint64_t* mine = status + (size_t)t * kDigits + d;
        uint32_t excl = 0;
        if (t == 0) {
            store_status(mine, kInc | cnt[0]);
        } else {
            store_status(mine, kAgg | cnt[0]);
            for (int64_t j = (int64_t)t - 1; j >= 0; --j) {
                const uint64_t* at = status + (size_t)j * kDigits + d;
                uint64_t x = load_status(at);
                for (uint32_t spin = 0; (x >> 32) == 0 && spin < kSpinMax; ++spin) {
                    __builtin_amdgcn_s_sleep(1);
                    x = load_status(at);
                }
                if ((x >> 32) == 0) {  // not reached: flags the sort as failed rather than wait on
                    atomicOr(&ctl[kErrWord], 1u);
                    break;
                }
                excl += (uint32_t)x;
                if ((x >> 32) == 2) break;
            }
            store_status(mine, kInc | (uint32_t)(excl + cnt[0]));
        }
#pragma unroll
        for (int i = 0; i < kSortPer; ++i) {
            sk[r[i]] = k[i];
            if constexpr (V) sv[r[i]] = v[i];
        }
        s_pre[d] = pre[0];
        s_out[d] = dstart + excl;
        __syncthreads();
        for (uint32_t idx = threadIdx.x; idx < valid; idx += kT) {
            const K x = sk[idx];
            const uint32_t dx = digit_of(x, shift);
            const uint32_t pos = s_out[dx] + idx - s_pre[dx];
            if (pos >= n) continue;  // unreachable when the counts agree; keeps a bad count in bounds
            kout[pos] = x;
            if constexpr (V) vout[pos] = sv[idx];
        }
        __syncthreads();
    }
}

// Workspace: control words, digit totals (kMaxPasses x 256), then the
// look-back statuses (per pass: tiles x 256 words).
struct SortWs {
    uint32_t* ctl;
    uint32_t* totals;
    uint64_t* status;
    size_t status_pass;  // words per pass
};

SortWs sort_ws(void* ws, size_t cap) {
    auto* b = static_cast<unsigned char*>(ws);
    SortWs w;
    w.ctl = reinterpret_cast<uint32_t*>(b);
    w.totals = w.ctl + kCtlWords;
    w.status = reinterpret_cast<uint64_t*>(b + (kCtlWords + kMaxPasses * kDigits) * 4);
    w.status_pass = tiles(cap, kSortTile) * kDigits;
    return w;
}

// Returns whether the result is in (kt, vt).
template <class K, bool V>
bool dsort_t(K* keys, uint32_t* vals, K* kt, uint32_t* vt, size_t cap, const uint32_t* d_n, int end_bit, void* ws,
             hipStream_t s) {
    const int passes = (std::max(end_bit, 0) + kDigitBits - 1) / kDigitBits;
    if (passes == 0) return false;
    BTC_REQUIRE(passes <= kMaxPasses && end_bit <= (int)(8 * sizeof(K)), BTC_ERR_ARG, "sort on %d bits", end_bit);
    const SortWs w = sort_ws(ws, cap);
    const unsigned g = grid_of(cap, kSortTile);
    const unsigned gp = std::min<unsigned>(g, 2048);  // workgroups take tiles by ticket
    hipLaunchKernelGGL(dsort_setup, dim3(1), dim3(kT), 0, s, d_n, w.ctl, w.totals, passes);
    BTC_CHECK_LAUNCH("dsort_setup");
    hipLaunchKernelGGL(dsort_hist<K>, dim3(g), dim3(kT), 0, s, keys, d_n, w.ctl, passes, w.totals, w.status,
                       w.status_pass);
    BTC_CHECK_LAUNCH("dsort_hist");
    K* ki = keys;
    K* ko = kt;
    uint32_t* vi = vals;
    uint32_t* vo = vt;
    for (int p = 0; p < passes; ++p) {
        hipLaunchKernelGGL((dsort_onesweep<K, V>), dim3(gp), dim3(kT), 0, s, ki, vi, ko, vo, d_n, w.ctl, p,
                           w.totals, w.status + (size_t)p * w.status_pass);
        BTC_CHECK_LAUNCH("dsort_onesweep");
        std::swap(ki, ko);
        std::swap(vi, vo);
    }
    return passes & 1;
}

}  // namespace

size_t dscan_ws_bytes(size_t cap) { return kScanSumsAt + tiles(cap, kScanTile) * 8; }

const uint32_t* dsort_error_word(const void* ws) { return static_cast<const uint32_t*>(ws) + kErrWord; }

size_t dsort_ws_bytes(size_t cap) {
    return (kCtlWords + kMaxPasses * kDigits) * 4 + (size_t)kMaxPasses * tiles(cap, kSortTile) * kDigits * 8;
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
