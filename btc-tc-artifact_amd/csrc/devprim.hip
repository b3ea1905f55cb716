// Sort and scan whose element count is read from device memory when the
// kernels run.  They enqueue kernels only -- no allocation, no memset node,
// no host synchronisation -- so a stream capture of them replays for any
// count up to the capacity their buffers were sized for (see prim.h).
//
// Scan: one kernel; workgroups take 4096-element tiles in ticket order,
// publish each tile's sum and find its offset by decoupled look-back over
// the tiles before it.  Sort: least-significant-digit radix sort, 8 bits a
// pass, in one sweep per pass: one kernel takes the digit totals of every
// pass in a single read of the keys; each pass kernel then ranks a tile
// stably in LDS (rocPRIM's block_radix_rank, match form: per-wave ballots
// of equal digits) and finds the tile's offset for each digit by decoupled
// look-back -- 1 + passes kernels a sort.
//
// Every look-back waits only on tiles whose tickets were taken by running
// workgroups, which publish before they wait, so every wait ends (a spin
// limit sets an error word rather than waiting on).  The workspace starts
// zeroed and every scan and sort leaves its tickets, totals and statuses
// zeroed again: the last workgroup to finish a scan, or a sort's last pass,
// clears them, so no clearing kernel or memset is needed between uses.
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

// Workspace layout (256-byte aligned parts): control words, then the sort's
// digit totals (kMaxPasses x 256 u32), then the scan's tile statuses, then
// the sort's tile statuses (per pass: tiles x 256).
constexpr int kMaxPasses = 8;
constexpr uint32_t kCtlWords = 64;  // 256 bytes
// control words: [0] sort tiles, [1 + p] sort pass p's ticket, then:
constexpr uint32_t kErrWord = kMaxPasses + 1;     // a look-back ran past its spin limit (stays set)
constexpr uint32_t kSortDone = kMaxPasses + 2;    // workgroups of a sort's last pass that have finished
constexpr uint32_t kScanTicket = kMaxPasses + 3;
constexpr uint32_t kScanDone = kMaxPasses + 4;
static_assert(kScanDone < kCtlWords, "control words");
constexpr size_t kTotalsAt = kCtlWords * 4;
constexpr size_t kScanStatusAt = kTotalsAt + (size_t)kMaxPasses * kDigits * 4;
static_assert(kScanStatusAt % 256 == 0, "aligned parts");

// Look-back status of a tile (and digit): flag << 62 | value; flag 1 = the
// tile's own value, 2 = inclusive of every earlier tile.  Values stay below
// 2^62 (element counts are u32; scanned u64 values are check counts).
constexpr uint64_t kAgg = 1ull << 62, kInc = 2ull << 62, kValMask = (1ull << 62) - 1;
constexpr uint32_t kSpinMax = 1u << 22;

__device__ __forceinline__ uint64_t load_status(const uint64_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void store_status(uint64_t* p, uint64_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Sum of the values of tiles [0, t) for look-back cell j0 = t - 1 downwards
// (stride apart): waits on each flag, stops at an inclusive one.
__device__ __forceinline__ uint64_t look_back(const uint64_t* cells, int64_t j0, size_t stride, uint32_t* err) {
    uint64_t excl = 0;
    for (int64_t j = j0; j >= 0; --j) {
        const uint64_t* at = cells + (size_t)j * stride;
        uint64_t x = load_status(at);
        for (uint32_t spin = 0; (x >> 62) == 0 && spin < kSpinMax; ++spin) {
            __builtin_amdgcn_s_sleep(1);
            x = load_status(at);
        }
        if ((x >> 62) == 0) {  // not reached: flags the result as failed rather than wait on
            atomicOr(err, 1u);
            break;
        }
        excl += x & kValMask;
        if ((x >> 62) == 2) break;
    }
    return excl;
}

// Thread 0 of the workgroup that finishes last (every thread calls it, at
// the kernel's end); the others see false.
__device__ __forceinline__ bool last_to_finish(uint32_t* done) {
    __shared__ bool s_last;
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence();
        s_last = atomicAdd(done, 1u) == gridDim.x - 1;
    }
    __syncthreads();
    return s_last;
}

// Exclusive scan of in[0, n) into out, out[n] and *d_total = the total.
template <class T>
__global__ __launch_bounds__(kT) void dscan_kernel(const T* __restrict__ in, T* __restrict__ out,
                                                  const uint32_t* __restrict__ d_n, T* __restrict__ d_total,
                                                  uint32_t* __restrict__ ctl, uint64_t* __restrict__ status) {
    __shared__ T lds[kT / 64];
    __shared__ uint32_t s_tile;
    __shared__ T s_excl;
    const uint32_t n = *d_n;
    const uint32_t nt = (n + kScanTile - 1) / kScanTile;
    for (;;) {
        if (threadIdx.x == 0) s_tile = atomicAdd(&ctl[kScanTicket], 1u);
        __syncthreads();
        const uint32_t t = s_tile;
        if (t >= nt) break;
        const size_t base = (size_t)t * kScanTile + (size_t)threadIdx.x * kScanPer;
        T v[kScanPer];
        T sum = 0;
#pragma unroll
        for (int k = 0; k < kScanPer; ++k) {
            v[k] = base + k < n ? in[base + k] : T(0);
            sum += v[k];
        }
        T tot;
        const T ex = block_scan_ex(sum, lds, &tot);
        if (threadIdx.x == 0) {
            T excl = 0;
            if (t == 0) {
                store_status(status, kInc | (uint64_t)tot);
            } else {
                store_status(status + t, kAgg | (uint64_t)tot);
                excl = (T)look_back(status, (int64_t)t - 1, 1, &ctl[kErrWord]);
                store_status(status + t, kInc | (uint64_t)(excl + tot));
            }
            if (t == nt - 1) {
                out[n] = excl + tot;
                if (d_total) *d_total = excl + tot;
            }
            s_excl = excl;
        }
        __syncthreads();
        T run = s_excl + ex;
#pragma unroll
        for (int k = 0; k < kScanPer; ++k) {
            if (base + k < n) out[base + k] = run;
            run += v[k];
        }
        __syncthreads();
    }
    if (last_to_finish(&ctl[kScanDone])) {
        for (uint32_t i = threadIdx.x; i < nt; i += kT) status[i] = 0;
        if (threadIdx.x == 0) {
            ctl[kScanTicket] = 0;
            ctl[kScanDone] = 0;
            if (nt == 0) {
                out[0] = 0;
                if (d_total) *d_total = 0;
            }
        }
    }
}

template <class T>
void dscan_t(const T* in, T* out, size_t cap, const uint32_t* d_n, T* d_total, const DevWs& ws, hipStream_t s) {
    BTC_REQUIRE(cap <= ws.cap, BTC_ERR_ARG, "scan of %zu elements in a workspace for %zu", cap, ws.cap);
    auto* b = static_cast<unsigned char*>(ws.p);
    const unsigned g = std::min<unsigned>(grid_of(cap, kScanTile), 2048);  // tiles go by ticket
    hipLaunchKernelGGL(dscan_kernel<T>, dim3(g), dim3(kT), 0, s, in, out, d_n, d_total,
                       reinterpret_cast<uint32_t*>(b), reinterpret_cast<uint64_t*>(b + kScanStatusAt));
    BTC_CHECK_LAUNCH("dscan_kernel");
}

// ------------------------------------------------------------------ sort
template <class K>
__device__ __forceinline__ uint32_t digit_of(K k, int shift) {
    return (uint32_t)(k >> shift) & (kDigits - 1);
}

// Digit totals of every pass in one read of the keys (per tile in LDS,
// then one atomic per non-zero (pass, digit), into totals the previous
// sort left cleared), the tile's look-back statuses of every pass cleared,
// and the tile count recorded for the pass kernels.
template <class K>
__global__ __launch_bounds__(kT) void dsort_hist(const K* __restrict__ keys, const uint32_t* __restrict__ d_n,
                                                 uint32_t* __restrict__ ctl, int passes,
                                                 uint32_t* __restrict__ totals, uint64_t* __restrict__ status,
                                                 size_t status_pass) {
    __shared__ uint32_t h[kMaxPasses][kDigits];
    const uint32_t n = *d_n, nt = (n + kSortTile - 1) / kSortTile;
    if (blockIdx.x == 0 && threadIdx.x == 0) ctl[0] = nt;
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
                                                     int pass, int passes, uint32_t* __restrict__ totals,
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
            excl = (uint32_t)look_back(status + d, (int64_t)t - 1, kDigits, &ctl[kErrWord]);
            store_status(mine, kInc | (uint64_t)(excl + cnt[0]));
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
    // after the last pass the sort's totals and tickets go back to zero
    if (pass == passes - 1 && last_to_finish(&ctl[kSortDone])) {
        for (uint32_t i = threadIdx.x; i < (uint32_t)passes * kDigits; i += kT) totals[i] = 0;
        if (threadIdx.x <= (uint32_t)passes) ctl[threadIdx.x] = 0;  // tile count and tickets
        if (threadIdx.x == 0) ctl[kSortDone] = 0;
    }
}

struct SortWs {
    uint32_t* ctl;
    uint32_t* totals;
    uint64_t* status;
    size_t status_pass;  // words per pass
};

size_t scan_status_bytes(size_t cap) { return (tiles(cap, kScanTile) * 8 + 255) & ~size_t(255); }

SortWs sort_ws(const DevWs& ws) {
    auto* b = static_cast<unsigned char*>(ws.p);
    SortWs w;
    w.ctl = reinterpret_cast<uint32_t*>(b);
    w.totals = reinterpret_cast<uint32_t*>(b + kTotalsAt);
    w.status = reinterpret_cast<uint64_t*>(b + kScanStatusAt + scan_status_bytes(ws.cap));
    w.status_pass = tiles(ws.cap, kSortTile) * kDigits;
    return w;
}

// Returns whether the result is in (kt, vt).
template <class K, bool V>
bool dsort_t(K* keys, uint32_t* vals, K* kt, uint32_t* vt, size_t cap, const uint32_t* d_n, int end_bit,
             const DevWs& ws, hipStream_t s) {
    const int passes = (std::max(end_bit, 0) + kDigitBits - 1) / kDigitBits;
    if (passes == 0) return false;
    BTC_REQUIRE(passes <= kMaxPasses && end_bit <= (int)(8 * sizeof(K)), BTC_ERR_ARG, "sort on %d bits", end_bit);
    BTC_REQUIRE(cap <= ws.cap, BTC_ERR_ARG, "sort of %zu keys in a workspace for %zu", cap, ws.cap);
    const SortWs w = sort_ws(ws);
    const unsigned g = grid_of(cap, kSortTile);
    const unsigned gp = std::min<unsigned>(g, 2048);  // workgroups take tiles by ticket
    hipLaunchKernelGGL(dsort_hist<K>, dim3(g), dim3(kT), 0, s, keys, d_n, w.ctl, passes, w.totals, w.status,
                       w.status_pass);
    BTC_CHECK_LAUNCH("dsort_hist");
    K* ki = keys;
    K* ko = kt;
    uint32_t* vi = vals;
    uint32_t* vo = vt;
    for (int p = 0; p < passes; ++p) {
        hipLaunchKernelGGL((dsort_onesweep<K, V>), dim3(gp), dim3(kT), 0, s, ki, vi, ko, vo, d_n, w.ctl, p, passes,
                           w.totals, w.status + (size_t)p * w.status_pass);
        BTC_CHECK_LAUNCH("dsort_onesweep");
        std::swap(ki, ko);
        std::swap(vi, vo);
    }
    return passes & 1;
}

}  // namespace

size_t dprim_ws_bytes(size_t cap) {
    return kScanStatusAt + scan_status_bytes(cap) + (size_t)kMaxPasses * tiles(cap, kSortTile) * kDigits * 8;
}

const uint32_t* dsort_error_word(const DevWs& ws) { return static_cast<const uint32_t*>(ws.p) + kErrWord; }

void dscan_u32(const uint32_t* in, uint32_t* out, size_t cap, const uint32_t* d_n, uint32_t* d_total, const DevWs& ws,
               hipStream_t s) {
    dscan_t<uint32_t>(in, out, cap, d_n, d_total, ws, s);
}

void dscan_u64(const uint64_t* in, uint64_t* out, size_t cap, const uint32_t* d_n, uint64_t* d_total, const DevWs& ws,
               hipStream_t s) {
    dscan_t<uint64_t>(in, out, cap, d_n, d_total, ws, s);
}

uint64_t* dsort_u64(uint64_t* keys, uint64_t* tmp, size_t cap, const uint32_t* d_n, int end_bit, const DevWs& ws,
                    hipStream_t s) {
    return dsort_t<uint64_t, false>(keys, nullptr, tmp, nullptr, cap, d_n, end_bit, ws, s) ? tmp : keys;
}

bool dsort_pairs_u32(uint32_t* keys, uint32_t* vals, uint32_t* kt, uint32_t* vt, size_t cap, const uint32_t* d_n,
                     int end_bit, const DevWs& ws, hipStream_t s) {
    return dsort_t<uint32_t, true>(keys, vals, kt, vt, cap, d_n, end_bit, ws, s);
}

}  // namespace btcx
