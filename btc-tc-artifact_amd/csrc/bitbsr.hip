// Bit-BSR builder and the reference-semantics tile count ("btc-plain") for
// gfx950.
//
// Builder: the reference sorts 64-bit tile keys of every edge and then sets
// bits with one atomicOr per edge after a binary search
// (btc/operations/bcsr_16x128.h:137-241, bcsr_16x32.h:144-249).  Here L is
// already a row-sorted CSR, so a row block's tiles are the 16-way merge of
// its rows' column blocks: 16 lanes (one per row) step through their rows
// together, the group minimum of the next column block is the next tile, and
// each lane writes its row of that tile directly -- no sort, no atomics, no
// memset of the tile array.  A wave handles 4 row blocks (4 x 16 lanes).  A
// counting pass sizes the row blocks, a scan gives indptr, a fill pass
// writes indices, row_indices and tiles.
//
// Count: one wave64 per mask tile S = (I, J_L), as one warp per tile in
// tc_16x128_mma_v5.h:9-211 / tc_16x32_mma_v6.h:9-238: popcount(S) > tau
// picks the dense path (wave-uniform two-pointer merge of the tile columns of
// row blocks I and J_out, every matched pair contributing a full 16 x 16 block
// of AND+popcount dot products, masked by S afterwards) or the sparse path
// (per-lane merge, only the set mask bits of the lane are evaluated).  Lane
// (r, q) = (lane / 4, lane % 4) owns row r of the tile and the 4 columns
// 4q..4q+3 of each 16-column J_out slab, so every lane's partial sum belongs
// to a single vertex 16I + r.
#include <type_traits>

#include "device_util.h"
#include "prim.h"
#include "internal.h"

namespace btcx {

// ------------------------------------------------------------ graph upload
void ensure_device(btc_graph_s* g, hipStream_t s) {
    if (g->d_rowptr.p) return;
    BTC_REQUIRE(g->m < (int64_t)0xFFFFFFFFLL, BTC_ERR_OVERFLOW,
                "nnz(L) = %lld needs 64-bit row offsets (limit 2^32 - 1)", (long long)g->m);
    std::vector<uint32_t> rp32((size_t)g->n + 1);
    for (size_t i = 0; i < rp32.size(); ++i) rp32[i] = (uint32_t)g->rowptr[i];
    g->d_rowptr.alloc(rp32.size());
    g->d_col.alloc((size_t)g->m + kColPad);
    BTC_HIP(hipMemsetAsync(g->d_col.p + g->m, 0, kColPad * 4, s));
    BTC_HIP(hipMemcpyAsync(g->d_rowptr.p, rp32.data(), rp32.size() * 4, hipMemcpyHostToDevice, s));
    if (g->m)
        BTC_HIP(hipMemcpyAsync(g->d_col.p, g->col.data(), (size_t)g->m * 4, hipMemcpyHostToDevice, s));
    BTC_HIP(hipStreamSynchronize(s));
}

int device_cu_count() {
    int dev = 0;
    BTC_HIP(hipGetDevice(&dev));
    int cus = 0;
    BTC_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    return cus;
}

// ------------------------------------------------------------------- scan
namespace {

constexpr int kScanThreads = 256;
constexpr int kScanItems = 4;
constexpr int kScanTile = kScanThreads * kScanItems;

// Block-wide exclusive scan of one value per thread; returns the block total.
__device__ uint64_t block_exclusive_scan(uint64_t v, uint64_t* out, uint64_t* lds) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    uint64_t x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        uint64_t y = __shfl_up(x, o);
        if (lane >= o) x += y;
    }
    if (lane == 63) lds[wid] = x;
    __syncthreads();
    uint64_t base = 0, total = 0;
    for (int w = 0; w < kScanThreads / 64; ++w) {
        if (w < wid) base += lds[w];
        total += lds[w];
    }
    __syncthreads();
    *out = base + x - v;
    return total;
}

__global__ __launch_bounds__(kScanThreads) void scan_tile_sums(const uint32_t* __restrict__ in,
                                                                size_t n, uint64_t* __restrict__ sums) {
    __shared__ uint64_t lds[kScanThreads / 64];
    size_t base = (size_t)blockIdx.x * kScanTile + (size_t)threadIdx.x * kScanItems;
    uint64_t v = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k)
        if (base + k < n) v += in[base + k];
    uint64_t ex;
    uint64_t tot = block_exclusive_scan(v, &ex, lds);
    if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

__global__ __launch_bounds__(kScanThreads) void scan_sums(uint64_t* __restrict__ sums, size_t nb,
                                                           uint64_t* __restrict__ total_out) {
    __shared__ uint64_t lds[kScanThreads / 64];
    uint64_t carry = 0;
    for (size_t b0 = 0; b0 < nb; b0 += kScanThreads) {
        size_t b = b0 + threadIdx.x;
        uint64_t v = b < nb ? sums[b] : 0;
        uint64_t ex;
        uint64_t tot = block_exclusive_scan(v, &ex, lds);
        if (b < nb) sums[b] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0) *total_out = carry;
}

__global__ __launch_bounds__(kScanThreads) void scan_apply(const uint32_t* __restrict__ in, size_t n,
                                                            const uint64_t* __restrict__ sums,
                                                            uint64_t* __restrict__ out) {
    __shared__ uint64_t lds[kScanThreads / 64];
    size_t base = (size_t)blockIdx.x * kScanTile + (size_t)threadIdx.x * kScanItems;
    uint32_t v[kScanItems];
    uint64_t s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        v[k] = base + k < n ? in[base + k] : 0;
        s += v[k];
    }
    uint64_t ex;
    block_exclusive_scan(s, &ex, lds);
    uint64_t run = sums[blockIdx.x] + ex;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        if (base + k < n) out[base + k] = run;
        run += v[k];
    }
}

}  // namespace

void scan_u32_to_u64(const uint32_t* in, uint64_t* out, size_t n, hipStream_t s) {
    size_t nb = (n + kScanTile - 1) / kScanTile;
    if (nb == 0) {
        BTC_HIP(hipMemsetAsync(out, 0, sizeof(uint64_t), s));
        return;
    }
    TmpBuf<uint64_t> sums;
    sums.alloc(nb, s);
    hipLaunchKernelGGL(scan_tile_sums, dim3((unsigned)nb), dim3(kScanThreads), 0, s, in, n, sums.p);
    BTC_CHECK_LAUNCH("scan_tile_sums");
    hipLaunchKernelGGL(scan_sums, dim3(1), dim3(kScanThreads), 0, s, sums.p, nb, out + n);
    BTC_CHECK_LAUNCH("scan_sums");
    hipLaunchKernelGGL(scan_apply, dim3((unsigned)nb), dim3(kScanThreads), 0, s, in, n, sums.p, out);
    BTC_CHECK_LAUNCH("scan_apply");

}

// ---------------------------------------------------------------- builder
namespace {

// Bits of one 16 x K tile row held by a lane.
template <int K>
struct RowBits;

template <>
struct RowBits<32> {
    uint32_t w = 0;
    __device__ void set(int c) { w |= 1u << c; }
    __device__ static RowBits load(const uint64_t* T, size_t tile, int r) {
        RowBits x;
        x.w = reinterpret_cast<const uint32_t*>(T + tile * 8)[r];
        return x;
    }
    // the row at p (4-byte aligned)
    __device__ static RowBits at(const unsigned char* p) {
        RowBits x;
        x.w = *reinterpret_cast<const uint32_t*>(p);
        return x;
    }
    __device__ int and_popc(const RowBits& o) const { return __popc(w & o.w); }
    __device__ uint32_t bits(int c0) const { return w >> c0; }
};

template <>
struct RowBits<64> {
    uint64_t w = 0;
    __device__ void set(int c) { w |= 1ull << c; }
    __device__ static RowBits load(const uint64_t* T, size_t tile, int r) {
        RowBits x;
        x.w = T[tile * 16 + r];
        return x;
    }
    // the row at p (8-byte aligned)
    __device__ static RowBits at(const unsigned char* p) {
        RowBits x;
        x.w = *reinterpret_cast<const uint64_t*>(p);
        return x;
    }
    __device__ int and_popc(const RowBits& o) const { return __popcll(w & o.w); }
    __device__ uint32_t bits(int c0) const { return (uint32_t)(w >> c0); }
};

template <>
struct RowBits<128> {
    uint64_t w0 = 0, w1 = 0;
    __device__ void set(int c) {
        if (c < 64) w0 |= 1ull << c;
        else w1 |= 1ull << (c - 64);
    }
    __device__ static RowBits load(const uint64_t* T, size_t tile, int r) {
        RowBits x;
        const uint64_t* p = T + tile * 32 + 2 * r;
        x.w0 = p[0];
        x.w1 = p[1];
        return x;
    }
    // the row at p (16-byte aligned)
    __device__ static RowBits at(const unsigned char* p) {
        RowBits x;
        const uint4 v = *reinterpret_cast<const uint4*>(p);
        x.w0 = ((uint64_t)v.y << 32) | v.x;
        x.w1 = ((uint64_t)v.w << 32) | v.z;
        return x;
    }
    __device__ int and_popc(const RowBits& o) const {
        return __popcll(w0 & o.w0) + __popcll(w1 & o.w1);
    }
    __device__ uint32_t bits(int c0) const {
        return c0 < 64 ? (uint32_t)(w0 >> c0) : (uint32_t)(w1 >> (c0 - 64));
    }
};

// FILL = false: counts[I] = number of tiles of row block I.
// FILL = true: writes tile I's indices / row_indices / bits at offs[I].
template <int K, bool FILL>
__global__ __launch_bounds__(256) void build_kernel(int n, int n_rb, const uint32_t* __restrict__ rowptr,
                                                     const int32_t* __restrict__ col,
                                                     uint32_t* __restrict__ counts,
                                                     const uint64_t* __restrict__ offs,
                                                     uint32_t* __restrict__ indices,
                                                     uint32_t* __restrict__ row_indices,
                                                     uint64_t* __restrict__ tiles) {
    const int lane = threadIdx.x & 63;
    const int r = lane & 15;
    const int64_t I = ((int64_t)blockIdx.x * 4 + (threadIdx.x >> 6)) * 4 + (lane >> 4);
    const bool valid = I < n_rb;
    const int64_t row = I * 16 + r;
    uint32_t p = 0, e = 0;
    if (valid && row < n) {
        p = rowptr[row];
        e = rowptr[row + 1];
    }
    uint64_t t = (FILL && valid) ? offs[I] : 0;
    uint32_t cnt = 0;
    // the row's next four ids (kNoId past its end), refilled one load
    // ahead per consumed id so a run of ids in one tile waits on memory
    // once every four ids instead of once per id
    constexpr uint32_t kNoId = 0xFFFFFFFFu;
    auto at = [&](uint32_t q) { return q < e ? (uint32_t)col[q] : kNoId; };
    uint32_t b0 = at(p), b1 = at(p + 1), b2 = at(p + 2), b3 = at(p + 3);
    for (;;) {
        uint32_t nxt = b0 != kNoId ? b0 / K : kNoTile;
        uint32_t cur = group_min_u32(nxt, 16);
        if (__all(cur == kNoTile)) break;
        if (cur != kNoTile) {
            RowBits<K> bits;
            while (b0 != kNoId && b0 / K == cur) {
                bits.set((int)(b0 % K));
                ++p;
                b0 = b1;
                b1 = b2;
                b2 = b3;
                b3 = at(p + 3);
            }
            if (FILL) {
                if (r == 0) {
                    indices[t] = cur;
                    row_indices[t] = (uint32_t)I;
                }
                if constexpr (K == 32) {
                    uint32_t other = __shfl_xor(bits.w, 1, 16);
                    if ((r & 1) == 0) tiles[t * 8 + r / 2] = (uint64_t)bits.w | ((uint64_t)other << 32);
                } else if constexpr (K == 64) {
                    tiles[t * 16 + r] = bits.w;
                } else {
                    tiles[t * 32 + 2 * r] = bits.w0;
                    tiles[t * 32 + 2 * r + 1] = bits.w1;
                }
                ++t;
            }
            ++cnt;
        }
    }
    if (!FILL && valid && r == 0) counts[I] = cnt;
}

__global__ void narrow_u64_to_u32(const uint64_t* __restrict__ in, uint32_t* __restrict__ out, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = (uint32_t)in[i];
}

template <int K>
void build_impl(btc_graph_s* g, hipStream_t s, btc_bitbsr_s* b) {
    const int n = g->n;
    const int n_rb = (n + 15) / 16;
    b->n = n;
    b->n_rb = n_rb;
    b->K = K;
    b->wpt = K / 4;
    b->max_row = g->max_row;
    const unsigned grid = (unsigned)((n_rb + 15) / 16);
    TmpBuf<uint32_t> counts;
    TmpBuf<uint64_t> offs;
    counts.alloc((size_t)n_rb, s);
    offs.alloc((size_t)n_rb + 1, s);
    if (grid) {
        hipLaunchKernelGGL((build_kernel<K, false>), dim3(grid), dim3(256), 0, s, n, n_rb, g->d_rowptr.p,
                           g->d_col.p, counts.p, nullptr, nullptr, nullptr, nullptr);
        BTC_CHECK_LAUNCH("build_kernel<count>");
    }
    scan_u32_to_u64(counts.p, offs.p, (size_t)n_rb, s);
    uint64_t nblk = 0;
    BTC_HIP(hipMemcpyAsync(&nblk, offs.p + n_rb, 8, hipMemcpyDeviceToHost, s));
    BTC_HIP(hipStreamSynchronize(s));
    BTC_REQUIRE(nblk < 0xFFFFFFFFull, BTC_ERR_OVERFLOW, "%llu tiles exceed the u32 tile index",
                (unsigned long long)nblk);
    b->nblk = (uint32_t)nblk;
    b->indptr.alloc((size_t)n_rb + 1);
    b->indices.alloc(nblk);
    b->row_indices.alloc(nblk);
    b->tiles.alloc(nblk * (size_t)b->wpt);
    b->result.alloc(1);
    hipLaunchKernelGGL(narrow_u64_to_u32, dim3((unsigned)((n_rb + 1 + 255) / 256)), dim3(256), 0, s, offs.p,
                       b->indptr.p, (size_t)n_rb + 1);
    BTC_CHECK_LAUNCH("narrow_u64_to_u32");
    if (grid && nblk) {
        hipLaunchKernelGGL((build_kernel<K, true>), dim3(grid), dim3(256), 0, s, n, n_rb, g->d_rowptr.p,
                           g->d_col.p, nullptr, offs.p, b->indices.p, b->row_indices.p, b->tiles.p);
        BTC_CHECK_LAUNCH("build_kernel<fill>");
    }
}

// ------------------------------------------------------------------- count
// i8 MFMA form of the dense intersection.  A matched tile pair (pa of row
// block I, pb of row block Jo) contributes C = A B^T, A and B the 16 x K 0/1
// matrices of the two tiles, to the 16 x 16 block of (row of I, row of Jo)
// intersections.  v_mfma_i32_16x16x64_i8 takes A[i][k] with i = lane % 16
// and B[k][j] with j = lane % 16, 16 bytes per lane; k is a sum index, so
// any assignment of tile columns to a lane's bytes is exact as long as A and
// B use the same one: lane group g = lane / 16 takes columns 16g..16g+15 of
// each 64-column half (8g..8g+7 with the 16x16x32 form for K = 32).  D[r][c]
// sits in register v of lane l with r = 4 (l / 16) + v, c = l % 16.
typedef int v4i __attribute__((ext_vector_type(4)));

// 4 bits -> 4 bytes of value 0/1 (byte b = bit b); x < 16, no carries.
__device__ __forceinline__ int nib_bytes(uint32_t x) { return (int)((x * 0x00204081u) & 0x01010101u); }

__device__ __forceinline__ v4i bytes16(uint32_t b) {
    v4i v;
    v.x = nib_bytes(b & 15u);
    v.y = nib_bytes((b >> 4) & 15u);
    v.z = nib_bytes((b >> 8) & 15u);
    v.w = nib_bytes((b >> 12) & 15u);
    return v;
}

__device__ __forceinline__ long bytes8(uint32_t b) {
    return (long)((uint64_t)(uint32_t)nib_bytes(b & 15u) | ((uint64_t)(uint32_t)nib_bytes((b >> 4) & 15u) << 32));
}

// One matched tile's operand words for lane row i (K = 32: one 32-bit
// word; else K / 64 64-bit words), loaded apart from the MFMAs so that the
// loads of several matches can be in flight together.
template <int K>
struct MfmaOp {
    uint64_t v[K >= 64 ? K / 64 : 1];
};

template <int K>
__device__ __forceinline__ MfmaOp<K> mfma_load(const uint64_t* __restrict__ T, uint32_t p, int i) {
    MfmaOp<K> o;
    if constexpr (K == 32) {
        o.v[0] = reinterpret_cast<const uint32_t*>(T + (size_t)p * 8)[i];
    } else {
#pragma unroll
        for (int h = 0; h < K / 64; ++h) o.v[h] = T[(size_t)p * (K / 4) + (K / 64) * i + h];
    }
    return o;
}

template <int K>
__device__ __forceinline__ v4i mfma_mac(const MfmaOp<K>& a, const MfmaOp<K>& b, int g, v4i acc) {
    if constexpr (K == 32) {
        return __builtin_amdgcn_mfma_i32_16x16x32_i8(bytes8((uint32_t)(a.v[0] >> (8 * g)) & 0xFFu),
                                                     bytes8((uint32_t)(b.v[0] >> (8 * g)) & 0xFFu), acc, 0, 0, 0);
    } else {
#pragma unroll
        for (int h = 0; h < K / 64; ++h)
            acc = __builtin_amdgcn_mfma_i32_16x16x64_i8(bytes16((uint32_t)(a.v[h] >> (16 * g)) & 0xFFFFu),
                                                        bytes16((uint32_t)(b.v[h] >> (16 * g)) & 0xFFFFu), acc, 0, 0,
                                                        0);
        return acc;
    }
}

template <int K>
__device__ __forceinline__ v4i mfma_pair(const uint64_t* __restrict__ T, uint32_t pa, uint32_t pb, int i, int g,
                                         v4i acc) {
    if constexpr (K == 32) {
        const uint32_t a = reinterpret_cast<const uint32_t*>(T + (size_t)pa * 8)[i];
        const uint32_t b = reinterpret_cast<const uint32_t*>(T + (size_t)pb * 8)[i];
        return __builtin_amdgcn_mfma_i32_16x16x32_i8(bytes8((a >> (8 * g)) & 0xFFu), bytes8((b >> (8 * g)) & 0xFFu),
                                                     acc, 0, 0, 0);
    } else {
#pragma unroll
        for (int h = 0; h < K / 64; ++h) {
            const uint64_t a = T[(size_t)pa * (K / 4) + (K / 64) * i + h];
            const uint64_t b = T[(size_t)pb * (K / 4) + (K / 64) * i + h];
            acc = __builtin_amdgcn_mfma_i32_16x16x64_i8(bytes16((uint32_t)(a >> (16 * g)) & 0xFFFFu),
                                                        bytes16((uint32_t)(b >> (16 * g)) & 0xFFFFu), acc, 0, 0, 0);
        }
        return acc;
    }
}

template <int K, bool BINSEARCH, bool PV, bool MFMA>
__global__ __launch_bounds__(256) void tile_count_kernel(
    int n, int n_rb, const uint32_t* __restrict__ indptr, const uint32_t* __restrict__ indices,
    const uint32_t* __restrict__ row_indices, const uint64_t* __restrict__ tiles, uint32_t w0,
    uint32_t w_end, const uint32_t* __restrict__ tile_list, uint32_t lstride, int tau,
    unsigned long long* __restrict__ result, PvOut pv) {
    constexpr int WPT = K / 4;
    constexpr int SLABS = K / 16;
    const int lane = threadIdx.x & 63;
    const uint32_t w = uniform(w0 + blockIdx.x * 4 + (threadIdx.x >> 6));
    if (w >= w_end) return;
    const uint32_t t = tile_list ? uniform(tile_list[(size_t)w * lstride]) : w;
    uint32_t I;
    if (BINSEARCH) {  // row block of tile t by binary search over indptr (tc_16x128_mma_v4.h:29-38)
        uint32_t lo = 0, hi = (uint32_t)n_rb;
        while (lo < hi) {
            uint32_t mid = (lo + hi) >> 1;
            if (indptr[mid + 1] <= t) lo = mid + 1;
            else hi = mid;
        }
        I = uniform(lo);
    } else {
        I = uniform(row_indices[t]);
    }
    const uint32_t JL = uniform(indices[t]);
    const uint32_t a0 = uniform(indptr[I]), a1 = uniform(indptr[I + 1]);

    uint32_t c = lane < WPT ? (uint32_t)__popcll(tiles[(size_t)t * WPT + lane]) : 0u;
    const bool dense = (int)wave_sum_u32(c) > tau;

    const int r = lane >> 2, q = lane & 3;
    const RowBits<K> S = RowBits<K>::load(tiles, t, r);
    uint32_t sum = 0;
    if (MFMA && dense) {
        // lane owns rows 4g + v (v < 4) of the mask tile and column c of each slab
        const int g = lane >> 4, c = lane & 15;
        RowBits<K> M[4];
#pragma unroll
        for (int v = 0; v < 4; ++v) M[v] = RowBits<K>::load(tiles, t, 4 * g + v);
        uint32_t rs[4] = {0, 0, 0, 0};
        for (int it = 0; it < SLABS; ++it) {
            const uint32_t Jo = JL * SLABS + it;
            if (Jo >= (uint32_t)n_rb) break;
            const uint32_t b0 = uniform(indptr[Jo]), b1 = uniform(indptr[Jo + 1]);
            v4i acc = {0, 0, 0, 0};
            uint32_t pa = a0, pb = b0;
            while (pa < a1 && pb < b1) {
                const uint32_t ka = uniform(indices[pa]), kb = uniform(indices[pb]);
                if (ka < kb) {
                    ++pa;
                } else if (ka > kb) {
                    ++pb;
                } else {
                    acc = mfma_pair<K>(tiles, pa, pb, c, g, acc);
                    ++pa;
                    ++pb;
                }
            }
            rs[0] += (M[0].bits(it * 16 + c) & 1u) ? (uint32_t)acc.x : 0u;
            rs[1] += (M[1].bits(it * 16 + c) & 1u) ? (uint32_t)acc.y : 0u;
            rs[2] += (M[2].bits(it * 16 + c) & 1u) ? (uint32_t)acc.z : 0u;
            rs[3] += (M[3].bits(it * 16 + c) & 1u) ? (uint32_t)acc.w : 0u;
        }
        if (PV) {
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                uint32_t x = rs[v];
                x += __shfl_xor(x, 1);
                x += __shfl_xor(x, 2);
                x += __shfl_xor(x, 4);
                x += __shfl_xor(x, 8);
                const int64_t vtx = (int64_t)I * 16 + 4 * g + v;
                if (c == 0 && x && vtx < n) pv_add(pv, vtx, x);
            }
        }
        const uint64_t tot = wave_sum_u64((uint64_t)(rs[0] + rs[1] + rs[2] + rs[3]));
        if (lane == 0 && tot) atomicAdd(result, (unsigned long long)tot);
        return;
    }
    for (int it = 0; it < SLABS; ++it) {
        const uint32_t Jo = JL * SLABS + it;
        if (Jo >= (uint32_t)n_rb) break;
        const uint32_t b0 = uniform(indptr[Jo]), b1 = uniform(indptr[Jo + 1]);
        const uint32_t m4 = S.bits(it * 16 + 4 * q) & 0xFu;
        if (dense) {
            uint32_t acc0 = 0, acc1 = 0, acc2 = 0, acc3 = 0;
            uint32_t pa = a0, pb = b0;
            while (pa < a1 && pb < b1) {
                const uint32_t ka = uniform(indices[pa]), kb = uniform(indices[pb]);
                if (ka < kb) {
                    ++pa;
                } else if (ka > kb) {
                    ++pb;
                } else {
                    const RowBits<K> A = RowBits<K>::load(tiles, pa, r);
                    acc0 += A.and_popc(RowBits<K>::load(tiles, pb, 4 * q + 0));
                    acc1 += A.and_popc(RowBits<K>::load(tiles, pb, 4 * q + 1));
                    acc2 += A.and_popc(RowBits<K>::load(tiles, pb, 4 * q + 2));
                    acc3 += A.and_popc(RowBits<K>::load(tiles, pb, 4 * q + 3));
                    ++pa;
                    ++pb;
                }
            }
            sum += ((m4 & 1u) ? acc0 : 0u) + ((m4 & 2u) ? acc1 : 0u) + ((m4 & 4u) ? acc2 : 0u) +
                   ((m4 & 8u) ? acc3 : 0u);
        } else if (m4) {
            uint32_t pa = a0, pb = b0;
            while (pa < a1 && pb < b1) {
                const uint32_t ka = indices[pa], kb = indices[pb];
                if (ka < kb) {
                    ++pa;
                } else if (ka > kb) {
                    ++pb;
                } else {
                    const RowBits<K> A = RowBits<K>::load(tiles, pa, r);
                    uint32_t mm = m4;
                    while (mm) {
                        const int cc = __ffs(mm) - 1;
                        sum += A.and_popc(RowBits<K>::load(tiles, pb, 4 * q + cc));
                        mm &= mm - 1;
                    }
                    ++pa;
                    ++pb;
                }
            }
        }
    }
    if (PV) {
        uint32_t rs = sum + __shfl_xor(sum, 1);
        rs += __shfl_xor(rs, 2);
        const int64_t v = (int64_t)I * 16 + r;
        if (q == 0 && rs && v < n) pv_add(pv, v, rs);
    }
    const uint64_t tot = wave_sum_u64((uint64_t)sum);
    if (lane == 0 && tot) atomicAdd(result, (unsigned long long)tot);
}

// Calls f(pa, pb) for every tile column present in both sorted lists
// indices[a0, a1) and indices[b0, b1), in ascending order; wave-uniform.
// Each lane takes one element of the shorter list per round of 64 and
// binary-searches the longer one, so the matching costs log-depth load
// chains instead of a two-pointer walk of both lists.
template <class F>
__device__ __forceinline__ void for_each_match(const uint32_t* __restrict__ indices, uint32_t a0, uint32_t a1,
                                               uint32_t b0, uint32_t b1, int lane, F&& f) {
    const bool swap = (b1 - b0) < (a1 - a0);
    const uint32_t s0 = swap ? b0 : a0, s1 = swap ? b1 : a1;  // iterated (shorter)
    const uint32_t l0 = swap ? a0 : b0, l1 = swap ? a1 : b1;  // searched (longer)
    for (uint32_t cbase = s0; cbase < s1; cbase += 64) {
        const uint32_t ps = cbase + (uint32_t)lane;
        uint32_t hit = 0xFFFFFFFFu;
        if (ps < s1) {
            const uint32_t k = indices[ps];
            uint32_t lo = l0, hi = l1;
            while (lo < hi) {
                const uint32_t mid = (lo + hi) >> 1;
                if (indices[mid] < k) lo = mid + 1;
                else hi = mid;
            }
            if (lo < l1 && indices[lo] == k) hit = lo;
        }
        uint64_t m = __ballot(hit != 0xFFFFFFFFu);
        while (m) {
            const int l = __ffsll((long long)m) - 1;
            m &= m - 1;
            const uint32_t x = uniform(cbase + (uint32_t)l), y = uniform((uint32_t)__shfl((int)hit, l));
            if (swap) f(y, x);
            else f(x, y);
        }
    }
}

// Dense units per wave: lists of at least kDenseRunUnits units (a dense
// kernel beside the wedge kernel at R-MAT s22) give each wave
// kDenseRunLong consecutive units, so column sets are reused while I stays
// and the wedge kernel keeps most of the chip (s22 count with the 4-lane
// wedge kernel: 2.727 ms with 8, 2.740 with 16, 2.774 with 32, 2.743
// with 4; with the
// 8-lane one 16 had been best); shorter lists give each wave half a tile's slabs, at
// least 2 (planted 2^17 with AND + popcount, tiles in descending cost:
// 1 / 2 / 4 units per wave took 50 / 46 / 40 us at 16 x 128, 39 / 32 / 36 us
// at 16 x 64, 26 / 23 / 29 us at 16 x 32).
constexpr uint32_t kDenseRunUnits = 65536;
constexpr uint32_t kDenseRunLong = 8;

// Tile columns of one row block as an LDS bitmap over column blocks with
// the number of set bits before each word: find(c) is the position of
// column c in the row block's sorted list (a0 + its rank), or kNoTile.
// Graphs with at most 32 kColSetWords column blocks use it.
constexpr uint32_t kColSetWords = 1344;
struct ColumnSet {
    uint32_t* bits;
    uint16_t* rank;  // < 2^16: a row block has fewer tiles than 32 kColSetWords
    uint32_t a0;
    uint32_t I = kNoTile;  // row block whose columns the set holds (kNoTile: empty)
    __device__ __forceinline__ uint32_t find(uint32_t c) const {
        const uint32_t w = bits[c >> 5];
        if (!((w >> (c & 31u)) & 1u)) return kNoTile;
        return a0 + rank[c >> 5] + (uint32_t)__popc(w & ((1u << (c & 31u)) - 1u));
    }
    // Sets the bits of indices[a0, a1) (the words [0, nw) are zero on
    // entry) and the ranks; one wave.
    __device__ void build(const uint32_t* __restrict__ indices, uint32_t a1, uint32_t nw, int lane) {
        for (uint32_t p = a0 + (uint32_t)lane; p < a1; p += 64) {
            const uint32_t c = indices[p];
            atomicOr(&bits[c >> 5], 1u << (c & 31u));
        }
        wave_lds_sync();
        const uint32_t per = (nw + 63) / 64, wb = min(nw, (uint32_t)lane * per), we = min(nw, wb + per);
        uint32_t s = 0;
        for (uint32_t x = wb; x < we; ++x) s += (uint32_t)__popc(bits[x]);
        uint32_t inc = s;  // inclusive scan over the lanes
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t y = (uint32_t)__shfl_up((int)inc, o);
            if (lane >= o) inc += y;
        }
        uint32_t run = inc - s;
        for (uint32_t x = wb; x < we; ++x) {
            rank[x] = (uint16_t)run;
            run += (uint32_t)__popc(bits[x]);
        }
        wave_lds_sync();
    }
    __device__ void clear(uint32_t nw, int lane) {
        wave_lds_sync();
        for (uint32_t x = (uint32_t)lane; x < nw; x += 64) bits[x] = 0;
        wave_lds_sync();
    }
    __device__ static void wave_lds_sync() {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
};

// Streaming a list of lb tile columns through a column set takes about
// lb / 64 dependent rounds; for_each_match takes min / 64 rounds of
// log2(max)-deep searches.
__device__ __forceinline__ bool stream_cheaper(uint32_t la, uint32_t lb) {
    const uint32_t mn = min(la, lb), mx = max(la, lb);
    return (lb + 63) / 64 <= ((mn + 63) / 64) * (uint32_t)(32 - __clz(mx | 1u));
}

// Calls f(pa, pb) for every tile column of indices[b0, b1) that the column
// set holds (pa its position in the set's row block, pb = its position
// here); wave-uniform.  The list is read 64 entries at a time, coalesced.
template <class F>
__device__ __forceinline__ void for_each_match_set(const ColumnSet& cs, const uint32_t* __restrict__ indices,
                                                   uint32_t b0, uint32_t b1, int lane, F&& f) {
    for (uint32_t base = b0; base < b1; base += 64) {
        const uint32_t pb = base + (uint32_t)lane;
        uint32_t hit = kNoTile;
        if (pb < b1) hit = cs.find(indices[pb]);
        uint64_t m = __ballot(hit != kNoTile);
        while (m) {
            const int l = __ffsll((long long)m) - 1;
            m &= m - 1;
            f(uniform((uint32_t)__shfl((int)hit, l)), uniform(base + (uint32_t)l));
        }
    }
}

// ------------------------------------------------------------- dense units
// A dense unit is one slab of a dense mask tile: tile t of row block I
// (tile columns [a0, a1) of I) and the row block Jo = indices[t] * K/16 +
// slab under its columns (tile columns [b0, b1) of Jo).  A unit is listed
// only when Jo holds tile columns and the mask tile has bits in the slab's
// 16 columns (other slabs add nothing).  Two uint4 per unit: {t, I, a0, a1},
// {Jo, b0, b1, slab}; units of a tile are consecutive, tiles in list order.

// Number of tile columns present in both sorted lists indices[a0, a1) and
// indices[b0, b1) (one thread, two pointers; plan time only).
__device__ uint32_t count_common(const uint32_t* __restrict__ indices, uint32_t a0, uint32_t a1, uint32_t b0,
                                 uint32_t b1) {
    uint32_t c = 0;
    while (a0 < a1 && b0 < b1) {
        const uint32_t x = indices[a0], y = indices[b0];
        c += x == y;
        a0 += x <= y;
        b0 += y <= x;
    }
    return c;
}

// FILL = false: for tile list[w], counts[w] = its units and key[w] = ~cost,
// cost = the sum over its units of 1 + ceil(|Jo| / 64) + ceil(matches /
// kStagePairs) (the unit's dependent list-chunk and tile-stage steps), val[w]
// = w.  FILL = true: thread k writes the units of tile list[order[k]] from
// units[offs[k]] on.  One thread per listed tile.
constexpr uint32_t kUnitStagePairs = 8;  // = kStagePairs of dense_unit_kernel
template <int K, bool FILL>
__global__ __launch_bounds__(256) void dense_units_kernel(const uint32_t* __restrict__ list, uint32_t nt, int n_rb,
                                                          const uint32_t* __restrict__ indptr,
                                                          const uint32_t* __restrict__ indices,
                                                          const uint32_t* __restrict__ row_indices,
                                                          const uint64_t* __restrict__ tiles,
                                                          uint32_t* __restrict__ counts, uint32_t* __restrict__ key,
                                                          uint32_t* __restrict__ val, const uint32_t* __restrict__ order,
                                                          const uint64_t* __restrict__ offs, uint4* __restrict__ units) {
    constexpr int SLABS = K / 16;
    const uint32_t k = blockIdx.x * 256 + threadIdx.x;
    if (k >= nt) return;
    const uint32_t w = FILL ? order[k] : k;
    const uint32_t t = list[w];
    // the tile's 16-bit row chunks: row r's bits of slab s are chunk r * SLABS + s
    const uint16_t* h = reinterpret_cast<const uint16_t*>(tiles + (size_t)t * (K / 4));
    uint32_t occ = 0;  // bit s: slab s holds a mask bit
    for (int r = 0; r < 16; ++r)
        for (int sl = 0; sl < SLABS; ++sl)
            if (h[r * SLABS + sl]) occ |= 1u << sl;
    const uint32_t JL = indices[t];
    const uint32_t I = row_indices[t], a0 = indptr[I], a1 = indptr[I + 1];
    uint64_t o = FILL ? offs[k] : 0;
    uint32_t c = 0;
    uint64_t cost = 0;
    for (int sl = 0; sl < SLABS; ++sl) {
        const uint64_t Jo = (uint64_t)JL * SLABS + (uint64_t)sl;
        if (!((occ >> sl) & 1u) || Jo >= (uint64_t)n_rb) continue;
        const uint32_t b0 = indptr[Jo], b1 = indptr[Jo + 1];
        if (b1 == b0) continue;
        if (FILL) {
            units[2 * o] = make_uint4(t, I, a0, a1);
            units[2 * o + 1] = make_uint4((uint32_t)Jo, b0, b1, (uint32_t)sl);
            ++o;
        } else {
            const uint32_t m = count_common(indices, a0, a1, b0, b1);
            cost += 1 + (b1 - b0 + 63) / 64 + (m + kUnitStagePairs - 1) / kUnitStagePairs;
        }
        ++c;
    }
    if (!FILL) {
        counts[w] = c;
        key[w] = 0xFFFFFFFFu - (uint32_t)min<uint64_t>(cost, 0xFFFFFFFFull);
        val[w] = w;
    }
}

__global__ __launch_bounds__(256) void gather_u32_kernel(const uint32_t* __restrict__ in,
                                                         const uint32_t* __restrict__ order, uint32_t n,
                                                         uint32_t* __restrict__ out) {
    const uint32_t k = blockIdx.x * 256 + threadIdx.x;
    if (k < n) out[k] = in[order[k]];
}

// Matched tile pairs of a unit are collected in a per-wave LDS ring and
// their tiles staged into LDS kStagePairs pairs at a time.
constexpr int kStagePairs = (int)kUnitStagePairs;
constexpr uint32_t kMatchRing = 128;  // > kStagePairs - 1 + 64 (one list chunk's matches)
template <int K>
constexpr int kTileBytes = 2 * K;  // 16 rows of K bits
template <int K>
constexpr int kStageBytes = 2 * kStagePairs * kTileBytes<K>;
// AND + popcount: a unit's mask entries (one byte each, row << 4 | column;
// at most 16 x 16) and, for per-vertex sums, its 16 row totals
constexpr int kEntryBytes = 256;
constexpr int kRowSumBytes = 16 * 8;
// per wave: stage, the match ring (uint2 each), the entries, the row totals
template <int K>
constexpr int kWaveDenseBytes = kStageBytes<K> + (int)kMatchRing * 8 + kEntryBytes + kRowSumBytes;

typedef const void __attribute__((address_space(1)))* GlobalVoid;
typedef void __attribute__((address_space(3)))* LdsVoid;

// The engine's dense tiles as units.  Each wave takes `run` consecutive
// units (unit first + w * stride for w in its range; the next unit's
// descriptor is fetched ahead).  Per unit it matches the tile columns of I
// and Jo in the kernel: through an LDS ColumnSet of I's columns (built when
// I changes; the first 64 entries of Jo's list are loaded before the build)
// where streaming Jo's list is cheaper, else by search.  Matched pairs go to
// the wave's LDS ring (a chunk's hits in parallel, ranked by mbcnt); every
// kStagePairs of them have their A and B tiles copied into the wave's LDS
// stage by global_load_lds_dwordx4 (each lane 16 bytes, all loads in flight
// together, no registers), then each pair's 16 x 16 block A B^T is taken
// from LDS: i8 MFMA (MFMA) or 64-bit AND + v_bcnt popcount.  The mask
// tile's slab selects which block entries count: MFMA computes all 256 and
// masks the sums; AND + popcount at K = 128 first lists the unit's mask
// entries (a ballot prefix sum) and computes only those, up to 4 per lane.
template <int K, bool PV, bool MFMA>
__global__ __launch_bounds__(256) void dense_unit_kernel(int n, const uint32_t* __restrict__ indices,
                                                         const uint64_t* __restrict__ tiles,
                                                         const uint4* __restrict__ units, uint32_t first,
                                                         uint32_t count, uint32_t stride, uint32_t set_words,
                                                         uint32_t run, unsigned long long* __restrict__ slots,
                                                         PvOut pv) {
    constexpr int SLABS = K / 16;
    constexpr uint32_t TB = kTileBytes<K>;
    constexpr uint32_t WB = kWaveDenseBytes<K>;
    constexpr uint32_t RB = K / 8;  // bytes of a tile row
    // AND + popcount lists the mask entries at K = 128 only (planted 2^17: 39.4 -> 28.8 us there; at
    // K = 64 31.8 against 31.9 us, at K = 32 28.5 against 23.0 us)
    constexpr bool kCompact = !MFMA && K == 128;
    // dynamic LDS: 4 waves' stage + ring, then 4 column sets' bits, then their ranks
    extern __shared__ __align__(16) unsigned char dense_lds[];
    __shared__ unsigned long long wg_part[4];
    const int lane = threadIdx.x & 63;
    const uint32_t wv = uniform(threadIdx.x >> 6);
    const uint32_t w0 = min(count, (blockIdx.x * 4 + wv) * run);  // waves past the end reach the barrier too
    const uint32_t w1 = min(count, w0 + run);
    unsigned char* stage_b = dense_lds + wv * WB;
    const uint64_t* stage = reinterpret_cast<const uint64_t*>(stage_b);
    uint2* ring = reinterpret_cast<uint2*>(dense_lds + wv * WB + kStageBytes<K>);
    uint8_t* ents = dense_lds + wv * WB + kStageBytes<K> + kMatchRing * 8;
    unsigned long long* row_sums = reinterpret_cast<unsigned long long*>(ents + kEntryBytes);
    if (PV && lane < 16) row_sums[lane] = 0;
    uint32_t* sets = reinterpret_cast<uint32_t*>(dense_lds + 4 * WB);
    ColumnSet cs{sets + wv * set_words, reinterpret_cast<uint16_t*>(sets + 4 * set_words) + wv * set_words, 0};
    if (set_words)
        for (uint32_t x = (uint32_t)lane; x < set_words; x += 64) cs.bits[x] = 0;
    uint64_t tot = 0;
    // a unit's mask chunk (lane r < 16 holds row r's 16 bits of the slab) and
    // the first 64 entries of Jo's list; loaded for the next unit while the
    // current one stages its tiles
    auto head_loads = [&](const uint4& A, const uint4& B, uint32_t& mrow, uint32_t& jc) {
        const uint32_t t = uniform(A.x), a0 = uniform(A.z), a1 = uniform(A.w);
        const uint32_t b0 = uniform(B.y), b1 = uniform(B.z), slab = uniform(B.w);
        mrow = 0;
        if (lane < 16) mrow = reinterpret_cast<const uint16_t*>(tiles + (size_t)t * (K / 4))[lane * SLABS + slab];
        jc = 0;
        if (set_words && stream_cheaper(a1 - a0, b1 - b0) && b0 + (uint32_t)lane < b1) jc = indices[b0 + lane];
    };
    uint4 nA = make_uint4(0, 0, 0, 0), nB = nA;
    uint32_t n_mrow = 0, n_jc = 0;
    if (w0 < w1) {
        const size_t u = (size_t)first + (size_t)w0 * stride;
        nA = units[2 * u];
        nB = units[2 * u + 1];
        head_loads(nA, nB, n_mrow, n_jc);
    }
    for (uint32_t w = w0; w < w1; ++w) {
        const uint32_t I = uniform(nA.y), a0 = uniform(nA.z), a1 = uniform(nA.w);
        const uint32_t b0 = uniform(nB.y), b1 = uniform(nB.z);
        const uint32_t mrow = n_mrow, jc = n_jc;
        const bool use_set = set_words && stream_cheaper(a1 - a0, b1 - b0);
        const bool more = w + 1 < w1;
        if (more) {  // the next unit's descriptor, fetched ahead
            const size_t u = (size_t)first + (size_t)(w + 1) * stride;
            nA = units[2 * u];
            nB = units[2 * u + 1];
        }
        uint32_t head = 0, tail = 0;  // ring positions (wave-uniform)
        v4i acc = {0, 0, 0, 0};                 // MFMA: D[4 (l / 16) + v][l % 16]
        // AND at K = 128: entry j * 64 + lane of the unit's mask entries
        // (valid below `listed`) has its A and B rows at byte offsets
        // oab[j] & 0xFFFF and oab[j] >> 16 of a staged pair; rounds = the
        // wave's entry rounds (1 to 4).  AND at K = 32 and 64: row l / 4,
        // columns 4 (l % 4) + v, every entry.
        uint32_t oab[4] = {0, 0, 0, 0}, sj[4] = {0, 0, 0, 0};
        uint32_t rounds = 0, listed = 0;
        if constexpr (kCompact) {
            const int r = lane >> 2, q = lane & 3;
            const uint32_t m4 = (__shfl((int)mrow, r) >> (4 * q)) & 0xFu;
            const uint32_t c4 = (uint32_t)__popc(m4);
            uint32_t pre = 0, total = 0;  // exclusive prefix of c4 (< 8) over the lanes, bit by bit
#pragma unroll
            for (int bt = 0; bt < 3; ++bt) {
                const uint64_t bal = __ballot((c4 >> bt) & 1u);
                pre += __builtin_amdgcn_mbcnt_hi((uint32_t)(bal >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)bal, 0u))
                       << bt;
                total += (uint32_t)__popcll(bal) << bt;
            }
#pragma unroll
            for (int v = 0; v < 4; ++v)
                if ((m4 >> v) & 1u) ents[pre++] = (uint8_t)((r << 4) | (4 * q + v));
            ColumnSet::wave_lds_sync();
            listed = uniform(total);
            rounds = (listed + 63) / 64;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const uint32_t x = (uint32_t)j * 64 + (uint32_t)lane;
                const uint32_t e = x < listed ? ents[x] : 0u;
                oab[j] = (e >> 4) * RB | (TB + (e & 15u) * RB) << 16;
            }
            ColumnSet::wave_lds_sync();  // the next unit rewrites the entries
        }
        // the listed entries of staged pairs [0, nb) at K = 128: R rounds,
        // full batches unrolled
        auto and_listed = [&](auto rounds_c, uint32_t nb) __attribute__((always_inline)) {
            constexpr int R = decltype(rounds_c)::value;
            if (nb == (uint32_t)kStagePairs) {
#pragma unroll
                for (int m = 0; m < kStagePairs; ++m)
#pragma unroll
                    for (int j = 0; j < R; ++j)
                        sj[j] += RowBits<K>::at(stage_b + 2 * m * TB + (oab[j] & 0xFFFFu))
                                     .and_popc(RowBits<K>::at(stage_b + 2 * m * TB + (oab[j] >> 16)));
            } else {
#pragma unroll 1
                for (uint32_t m = 0; m < nb; ++m)
#pragma unroll
                    for (int j = 0; j < R; ++j)
                        sj[j] += RowBits<K>::at(stage_b + 2 * m * TB + (oab[j] & 0xFFFFu))
                                     .and_popc(RowBits<K>::at(stage_b + 2 * m * TB + (oab[j] >> 16)));
            }
        };
        uint32_t s0 = 0, s1 = 0, s2 = 0, s3 = 0;  // AND at K = 32 and 64
        // stages the tiles of ring pairs [head, head + nb) and adds their blocks
        auto drain = [&](uint32_t nb) __attribute__((always_inline)) {
            const uint32_t bytes = 2 * nb * TB;
            for (uint32_t k = 0; k * 1024u < bytes; ++k) {
                const uint32_t o = k * 1024u + (uint32_t)lane * 16u;
                if (o < bytes) {
                    const uint32_t sl = o / TB;
                    const uint2 mp = ring[(head + (sl >> 1)) & (kMatchRing - 1)];
                    const uint32_t p = (sl & 1u) ? mp.y : mp.x;
                    const char* src = reinterpret_cast<const char*>(tiles) + (size_t)p * TB + (o % TB);
                    __builtin_amdgcn_global_load_lds((GlobalVoid)src, (LdsVoid)(stage_b + k * 1024u), 16, 0, 0);
                }
            }
            __builtin_amdgcn_s_waitcnt(0);  // the copies into LDS have landed
            ColumnSet::wave_lds_sync();
            if constexpr (MFMA) {
                const int g = lane >> 4, c = lane & 15;
                for (uint32_t m = 0; m < nb; ++m)
                    acc = mfma_mac<K>(mfma_load<K>(stage, 2 * m, c), mfma_load<K>(stage, 2 * m + 1, c), g, acc);
            } else if constexpr (kCompact) {
                switch (rounds) {
                    case 1: and_listed(std::integral_constant<int, 1>{}, nb); break;
                    case 2: and_listed(std::integral_constant<int, 2>{}, nb); break;
                    case 3: and_listed(std::integral_constant<int, 3>{}, nb); break;
                    default: and_listed(std::integral_constant<int, 4>{}, nb); break;
                }
            } else {
                const int r = lane >> 2, q = lane & 3;
                for (uint32_t m = 0; m < nb; ++m) {
                    const RowBits<K> A = RowBits<K>::load(stage, 2 * m, r);
                    s0 += A.and_popc(RowBits<K>::load(stage, 2 * m + 1, 4 * q + 0));
                    s1 += A.and_popc(RowBits<K>::load(stage, 2 * m + 1, 4 * q + 1));
                    s2 += A.and_popc(RowBits<K>::load(stage, 2 * m + 1, 4 * q + 2));
                    s3 += A.and_popc(RowBits<K>::load(stage, 2 * m + 1, 4 * q + 3));
                }
            }
            ColumnSet::wave_lds_sync();  // the stage is refilled by the next batch
            head += nb;
        };
        if (use_set) {
            if (cs.I != I) {
                if (cs.I != kNoTile) cs.clear(set_words, lane);
                cs.a0 = a0;
                cs.build(indices, a1, set_words, lane);
                cs.I = I;
            }
            for (uint32_t base = b0; base < b1; base += 64) {
                const uint32_t pb = base + (uint32_t)lane;
                uint32_t hit = kNoTile;
                if (pb < b1) hit = cs.find(base == b0 ? jc : indices[pb]);
                const uint64_t m = __ballot(hit != kNoTile);
                if (hit != kNoTile) {
                    const uint32_t rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                                                    __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
                    ring[(tail + rank) & (kMatchRing - 1)] = make_uint2(hit, pb);
                }
                tail += (uint32_t)__popcll(m);
                ColumnSet::wave_lds_sync();
                while (tail - head >= (uint32_t)kStagePairs) drain((uint32_t)kStagePairs);
            }
        } else {
            for_each_match(indices, a0, a1, b0, b1, lane, [&](uint32_t pa, uint32_t pb) {
                if (lane == 0) ring[tail & (kMatchRing - 1)] = make_uint2(pa, pb);
                ++tail;
                if (tail - head == (uint32_t)kStagePairs) {
                    ColumnSet::wave_lds_sync();
                    drain((uint32_t)kStagePairs);
                }
            });
            ColumnSet::wave_lds_sync();
        }
        if (more) head_loads(nA, nB, n_mrow, n_jc);  // in flight with the last stage's copies
        if (tail > head) drain(tail - head);
        if constexpr (MFMA) {
            const int g = lane >> 4, c = lane & 15;
            uint32_t rv[4];
            rv[0] = ((__shfl((int)mrow, 4 * g + 0) >> c) & 1) ? (uint32_t)acc.x : 0u;
            rv[1] = ((__shfl((int)mrow, 4 * g + 1) >> c) & 1) ? (uint32_t)acc.y : 0u;
            rv[2] = ((__shfl((int)mrow, 4 * g + 2) >> c) & 1) ? (uint32_t)acc.z : 0u;
            rv[3] = ((__shfl((int)mrow, 4 * g + 3) >> c) & 1) ? (uint32_t)acc.w : 0u;
            tot += (uint64_t)rv[0] + rv[1] + rv[2] + rv[3];
            if (PV) {
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    uint64_t x = rv[v];  // 16 lanes' entries of one row: 64-bit sum
                    x += __shfl_xor(x, 1);
                    x += __shfl_xor(x, 2);
                    x += __shfl_xor(x, 4);
                    x += __shfl_xor(x, 8);
                    const int64_t vtx = (int64_t)I * 16 + 4 * g + v;
                    if (c == 0 && x && vtx < n) pv_add(pv, vtx, x);
                }
            }
        } else if constexpr (kCompact) {
            uint64_t sum = 0;
#pragma unroll
            for (int j = 0; j < 4; ++j)
                if ((uint32_t)j * 64 + (uint32_t)lane < listed) sum += sj[j];
            tot += sum;
            if (PV) {
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    if ((uint32_t)j * 64 + (uint32_t)lane < listed && sj[j])
                        atomicAdd(&row_sums[(oab[j] & 0xFFFFu) / RB], (unsigned long long)sj[j]);
                ColumnSet::wave_lds_sync();
                if (lane < 16) {
                    const unsigned long long x = row_sums[lane];
                    row_sums[lane] = 0;
                    const int64_t v = (int64_t)I * 16 + lane;
                    if (x && v < n) pv_add(pv, v, x);
                }
                ColumnSet::wave_lds_sync();
            }
        } else {
            const int r = lane >> 2, q = lane & 3;
            const uint32_t m4 = (__shfl((int)mrow, r) >> (4 * q)) & 0xFu;
            const uint64_t sum = (uint64_t)((m4 & 1u) ? s0 : 0u) + ((m4 & 2u) ? s1 : 0u) + ((m4 & 4u) ? s2 : 0u) +
                                 ((m4 & 8u) ? s3 : 0u);
            tot += sum;
            if (PV) {
                uint64_t x = sum + __shfl_xor(sum, 1);
                x += __shfl_xor(x, 2);
                const int64_t v = (int64_t)I * 16 + r;
                if (q == 0 && x && v < n) pv_add(pv, v, x);
            }
        }
    }
    // one atomic per workgroup, spread over the slots
    tot = wave_sum_u64(tot);
    if (lane == 0) wg_part[wv] = tot;
    __syncthreads();
    if (threadIdx.x == 0) {
        const unsigned long long x = wg_part[0] + wg_part[1] + wg_part[2] + wg_part[3];
        if (x) atomicAdd(slots + kSlotStride * (blockIdx.x % kSlots), x);
    }
}

// pop[t] = set bits of tile t: one 64-bit word per lane, 64 / (K / 4)
// tiles per wave, summed over each tile's lanes; waves stride over the
// words (a wave's base is uniform, so every lane of a tile is present).
template <int K>
__global__ __launch_bounds__(256) void tile_popcount_kernel(const uint64_t* __restrict__ tiles, uint32_t nblk,
                                                             uint32_t* __restrict__ pop) {
    constexpr int WPT = K / 4;
    static_assert(64 % WPT == 0, "whole tiles per wave");
    const int lane = threadIdx.x & 63;
    const uint64_t nw = (uint64_t)nblk * WPT;
    for (uint64_t base = ((uint64_t)blockIdx.x * 256 + (threadIdx.x & ~63u)); base < nw;
         base += (uint64_t)gridDim.x * 256) {
        const uint64_t i = base + (uint64_t)lane;
        uint32_t c = i < nw ? (uint32_t)__popcll(tiles[i]) : 0u;
#pragma unroll
        for (int o = WPT / 2; o > 0; o >>= 1) c += __shfl_xor(c, o, WPT);
        if (lane % WPT == 0 && i < nw) pop[i / WPT] = c;
    }
}

// A dispatch holds at most 2^32 work-items, so long tile lists are split
// into launches of kMaxTilesPerLaunch waves.
constexpr uint32_t kMaxTilesPerLaunch = 1u << 24;

template <int K, bool BS, bool MF>
void launch_k(const btc_bitbsr_s* b, const uint32_t* list, uint32_t lstride, uint32_t nt, int tau,
              unsigned long long* res, PvOut pv, hipStream_t s) {
    for (uint32_t w0 = 0; w0 < nt; w0 += kMaxTilesPerLaunch) {
        const uint32_t chunk = std::min(kMaxTilesPerLaunch, nt - w0);
        const unsigned grid = (chunk + 3) / 4;
        const uint32_t end = w0 + chunk;
        if (pv.on())
            hipLaunchKernelGGL((tile_count_kernel<K, BS, true, MF>), dim3(grid), dim3(256), 0, s, b->n, b->n_rb,
                               b->indptr.p, b->indices.p, b->row_indices.p, b->tiles.p, w0, end, list, lstride, tau,
                               res, pv);
        else
            hipLaunchKernelGGL((tile_count_kernel<K, BS, false, MF>), dim3(grid), dim3(256), 0, s, b->n, b->n_rb,
                               b->indptr.p, b->indices.p, b->row_indices.p, b->tiles.p, w0, end, list, lstride, tau,
                               res, pv);
        BTC_CHECK_LAUNCH("tile_count_kernel");
    }
}

template <int K>
void launch_kk(const btc_bitbsr_s* b, const uint32_t* list, uint32_t lstride, uint32_t nt, bool bs, bool mfma,
               int tau, unsigned long long* res, PvOut pv, hipStream_t s) {
    if (bs) {
        mfma ? launch_k<K, true, true>(b, list, lstride, nt, tau, res, pv, s)
             : launch_k<K, true, false>(b, list, lstride, nt, tau, res, pv, s);
    } else {
        mfma ? launch_k<K, false, true>(b, list, lstride, nt, tau, res, pv, s)
             : launch_k<K, false, false>(b, list, lstride, nt, tau, res, pv, s);
    }
}

}  // namespace

void bitbsr_build(btc_graph_s* g, int K, hipStream_t s, btc_bitbsr_s* b) {
    ensure_device(g, s);
    switch (K) {
        case 32: build_impl<32>(g, s, b); break;
        case 64: build_impl<64>(g, s, b); break;
        case 128: build_impl<128>(g, s, b); break;
        default: fail(BTC_ERR_ARG, "tile width K = %d (expected 32, 64 or 128)", K);
    }
}

void launch_tile_count(const btc_bitbsr_s* b, const uint32_t* tile_list, uint32_t ntiles, bool binsearch,
                       bool mfma, int tau, unsigned long long* d_result, PvOut d_pv, hipStream_t s,
                       uint32_t lstride) {
    if (ntiles == 0) return;
    BTC_REQUIRE(lane_sums_fit(b->max_row, b->K), BTC_ERR_OVERFLOW,
                "longest row %d: a lane's 32-bit sum over %d tile columns could overflow", b->max_row, b->K / 4);
    switch (b->K) {
        case 32: launch_kk<32>(b, tile_list, lstride, ntiles, binsearch, mfma, tau, d_result, d_pv, s); break;
        case 64: launch_kk<64>(b, tile_list, lstride, ntiles, binsearch, mfma, tau, d_result, d_pv, s); break;
        default: launch_kk<128>(b, tile_list, lstride, ntiles, binsearch, mfma, tau, d_result, d_pv, s); break;
    }
}

uint32_t build_dense_units(const btc_bitbsr_s* b, const uint32_t* list, uint32_t nt, DevBuf<uint4>& units,
                           hipStream_t s) {
    units.release();
    if (nt == 0) return 0;
    TmpBuf<uint32_t> counts, key, val, kt, vt, sorted_counts;
    TmpBuf<uint64_t> offs;
    counts.alloc(nt, s);
    key.alloc(nt, s);
    val.alloc(nt, s);
    kt.alloc(nt, s);
    vt.alloc(nt, s);
    sorted_counts.alloc(nt, s);
    offs.alloc((size_t)nt + 1, s);
    const unsigned grid = (nt + 255) / 256;
    auto launch = [&](auto kern, const uint32_t* order, uint4* out) {
        hipLaunchKernelGGL(kern, dim3(grid), dim3(256), 0, s, list, nt, b->n_rb, b->indptr.p, b->indices.p,
                           b->row_indices.p, b->tiles.p, counts.p, key.p, val.p, order, offs.p, out);
        BTC_CHECK_LAUNCH("dense_units_kernel");
    };
    switch (b->K) {
        case 32: launch(dense_units_kernel<32, false>, nullptr, nullptr); break;
        case 64: launch(dense_units_kernel<64, false>, nullptr, nullptr); break;
        default: launch(dense_units_kernel<128, false>, nullptr, nullptr); break;
    }
    // tiles in descending cost (stable: list order among equals), so the
    // longest units start first and a tile's units stay consecutive
    const bool alt = sort_pairs_u32(key.p, val.p, kt.p, vt.p, nt, 32, s);
    const uint32_t* order = alt ? vt.p : val.p;
    hipLaunchKernelGGL(gather_u32_kernel, dim3(grid), dim3(256), 0, s, counts.p, order, nt, sorted_counts.p);
    BTC_CHECK_LAUNCH("gather_u32_kernel");
    scan_u32_to_u64(sorted_counts.p, offs.p, nt, s);
    uint64_t nu = 0;
    BTC_HIP(hipMemcpyAsync(&nu, offs.p + nt, 8, hipMemcpyDeviceToHost, s));
    BTC_HIP(hipStreamSynchronize(s));
    BTC_REQUIRE(nu < 0xFFFFFFFFull, BTC_ERR_OVERFLOW, "%llu dense units exceed a u32 index", (unsigned long long)nu);
    if (nu == 0) return 0;
    units.alloc(2 * nu);
    switch (b->K) {
        case 32: launch(dense_units_kernel<32, true>, order, units.p); break;
        case 64: launch(dense_units_kernel<64, true>, order, units.p); break;
        default: launch(dense_units_kernel<128, true>, order, units.p); break;
    }
    return (uint32_t)nu;
}

namespace {
template <int K, bool MF>
void launch_units_k(const btc_bitbsr_s* b, const uint4* units, uint32_t first, uint32_t count, uint32_t stride,
                    unsigned long long* res, PvOut pv, hipStream_t s) {
    // column blocks of the graph: ids of tile columns are below this
    const uint64_t ncb = ((uint64_t)b->n + K - 1) / K;
    const uint32_t set_words = ncb <= 32ull * kColSetWords ? (uint32_t)((ncb + 31) / 32) : 0u;
    // 4 waves x (stage + ring, set bit words, 16-bit ranks)
    const size_t lds = 4 * (size_t)kWaveDenseBytes<K> + 4 * (size_t)set_words * 4 + 4 * (size_t)set_words * 2;
    const uint32_t run = count >= kDenseRunUnits ? kDenseRunLong : (uint32_t)std::max(2, K / 32);
    for (uint32_t w0 = 0; w0 < count; w0 += kMaxTilesPerLaunch) {
        const uint32_t chunk = std::min(kMaxTilesPerLaunch, count - w0);
        const unsigned grid = (chunk + 4 * run - 1) / (4 * run);
        const uint32_t f0 = first + w0 * stride;
        if (pv.on())
            hipLaunchKernelGGL((dense_unit_kernel<K, true, MF>), dim3(grid), dim3(256), lds, s, b->n, b->indices.p,
                               b->tiles.p, units, f0, chunk, stride, set_words, run, res, pv);
        else
            hipLaunchKernelGGL((dense_unit_kernel<K, false, MF>), dim3(grid), dim3(256), lds, s, b->n, b->indices.p,
                               b->tiles.p, units, f0, chunk, stride, set_words, run, res, pv);
        BTC_CHECK_LAUNCH("dense_unit_kernel");
    }
}
}  // namespace

void launch_dense_units(const btc_bitbsr_s* b, const uint4* units, uint32_t first, uint32_t count, uint32_t stride,
                        bool mfma, unsigned long long* d_slots, PvOut d_pv, hipStream_t s) {
    if (count == 0) return;
    BTC_REQUIRE(lane_sums_fit(b->max_row, b->K), BTC_ERR_OVERFLOW,
                "longest row %d: a lane's 32-bit sum over %d tile columns could overflow", b->max_row, b->K / 4);
    switch (b->K) {
        case 32: mfma ? launch_units_k<32, true>(b, units, first, count, stride, d_slots, d_pv, s)
                      : launch_units_k<32, false>(b, units, first, count, stride, d_slots, d_pv, s); break;
        case 64: mfma ? launch_units_k<64, true>(b, units, first, count, stride, d_slots, d_pv, s)
                      : launch_units_k<64, false>(b, units, first, count, stride, d_slots, d_pv, s); break;
        default: mfma ? launch_units_k<128, true>(b, units, first, count, stride, d_slots, d_pv, s)
                      : launch_units_k<128, false>(b, units, first, count, stride, d_slots, d_pv, s); break;
    }
}

namespace {
__global__ void zero_u32_kernel(uint32_t* __restrict__ p, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = 0;
}
}  // namespace

namespace {
__global__ __launch_bounds__(64) void slot_sum_kernel(unsigned long long* __restrict__ slots,
                                                      unsigned long long* __restrict__ result) {
    const int lane = threadIdx.x;
    unsigned long long v = slots[kSlotStride * lane];
    slots[kSlotStride * lane] = 0;
    v = wave_sum_u64(v);
    if (lane == 0 && v) atomicAdd(result, v);
}
}  // namespace

void launch_slot_sum(unsigned long long* d_slots, unsigned long long* d_result, hipStream_t s) {
    static_assert(kSlots == 64, "one lane per slot");
    hipLaunchKernelGGL(slot_sum_kernel, dim3(1), dim3(64), 0, s, d_slots, d_result);
    BTC_CHECK_LAUNCH("slot_sum_kernel");
}

void launch_zero_u32(uint32_t* p, size_t n, hipStream_t s) {
    if (n == 0) return;
    const unsigned grid = (unsigned)std::min<size_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(zero_u32_kernel, dim3(grid), dim3(256), 0, s, p, n);
    BTC_CHECK_LAUNCH("zero_u32_kernel");
}

void launch_tile_popcount(const btc_bitbsr_s* b, uint32_t* d_pop, hipStream_t s) {
    if (b->nblk == 0) return;
    const uint64_t words = (uint64_t)b->nblk * (uint64_t)(b->K / 4);
    const unsigned grid = (unsigned)std::min<uint64_t>((words + 255) / 256, 1u << 16);
    switch (b->K) {
        case 32: hipLaunchKernelGGL(tile_popcount_kernel<32>, dim3(grid), dim3(256), 0, s, b->tiles.p, b->nblk, d_pop); break;
        case 64: hipLaunchKernelGGL(tile_popcount_kernel<64>, dim3(grid), dim3(256), 0, s, b->tiles.p, b->nblk, d_pop); break;
        default: hipLaunchKernelGGL(tile_popcount_kernel<128>, dim3(grid), dim3(256), 0, s, b->tiles.p, b->nblk, d_pop); break;
    }
    BTC_CHECK_LAUNCH("tile_popcount_kernel");
}

}  // namespace btcx

// ------------------------------------------------------------------ C ABI
extern "C" {

int btc_bitbsr_build(btc_graph g, int32_t tile_k, void* stream, btc_bitbsr* out) {
    return btcx::guard([&] {
        BTC_REQUIRE(g && out, BTC_ERR_ARG, "NULL argument");
        *out = nullptr;
        BTC_REQUIRE(tile_k == 32 || tile_k == 64 || tile_k == 128, BTC_ERR_ARG,
                    "tile width K = %d (expected 32, 64 or 128)", tile_k);
        auto b = std::make_unique<btc_bitbsr_s>();
        btcx::bitbsr_build(g, tile_k, btcx::as_stream(stream), b.get());
        BTC_HIP(hipStreamSynchronize(btcx::as_stream(stream)));
        *out = b.release();
    });
}

int btc_bitbsr_get_view(btc_bitbsr b, btc_bitbsr_view* v) {
    return btcx::guard([&] {
        BTC_REQUIRE(b && v, BTC_ERR_ARG, "NULL argument");
        v->indptr = b->indptr.p;
        v->indices = b->indices.p;
        v->row_indices = b->row_indices.p;
        v->tiles = b->tiles.p;
        v->n = b->n;
        v->n_row_blocks = b->n_rb;
        v->num_blocks = b->nblk;
        v->tile_k = b->K;
        v->words_per_tile = b->wpt;
    });
}

int btc_bitbsr_copy_to_host(btc_bitbsr b, uint32_t* indptr, uint32_t* indices, uint32_t* row_indices,
                            uint64_t* tiles) {
    return btcx::guard([&] {
        BTC_REQUIRE(b, BTC_ERR_ARG, "NULL handle");
        if (indptr) BTC_HIP(hipMemcpy(indptr, b->indptr.p, ((size_t)b->n_rb + 1) * 4, hipMemcpyDeviceToHost));
        if (b->nblk == 0) return;
        if (indices) BTC_HIP(hipMemcpy(indices, b->indices.p, (size_t)b->nblk * 4, hipMemcpyDeviceToHost));
        if (row_indices)
            BTC_HIP(hipMemcpy(row_indices, b->row_indices.p, (size_t)b->nblk * 4, hipMemcpyDeviceToHost));
        if (tiles)
            BTC_HIP(hipMemcpy(tiles, b->tiles.p, (size_t)b->nblk * b->wpt * 8, hipMemcpyDeviceToHost));
    });
}

void btc_bitbsr_free(btc_bitbsr b) { delete b; }

}  // extern "C"

namespace {

// btc_bitbsr_count with per-vertex output to int32 (pv32) or int64 (pv64).
void bitbsr_count(btc_bitbsr b, int32_t variant, int32_t threshold_blocks, int32_t tau, void* stream,
                  uint64_t* total, int32_t* pv32, int64_t* pv64, float* kernel_ms) {
    BTC_REQUIRE(b && total, BTC_ERR_ARG, "NULL argument");
    const bool mfma = (variant & BTC_DENSE_MFMA) != 0;
    variant &= ~BTC_DENSE_MFMA;
    BTC_REQUIRE(variant >= BTC_VARIANT_AUTO && variant <= BTC_VARIANT_DIRECT, BTC_ERR_ARG,
                "unknown variant %d", variant);
    hipStream_t s = btcx::as_stream(stream);
    if (tau < 0) tau = btc_default_tau(b->K);
    if (threshold_blocks <= 0) threshold_blocks = b->K == 128 ? 2048 : 40000;
    const bool bs = variant == BTC_VARIANT_BINSEARCH ||
                    (variant == BTC_VARIANT_AUTO && b->nblk < (uint32_t)threshold_blocks);
    btcx::PvBuf pv;
    if (pv32 || pv64) {
        pv.alloc((size_t)b->n, b->max_row);
        BTC_HIP(hipMemsetAsync(pv.data(), 0, pv.bytes(), s));
    }
    BTC_HIP(hipMemsetAsync(b->result.p, 0, 8, s));
    btcx::EventTimer tm(kernel_ms != nullptr);
    tm.start(s);
    btcx::launch_tile_count(b, nullptr, b->nblk, bs, mfma, tau, b->result.p, pv.out(), s);
    tm.stop(s);
    unsigned long long h = 0;
    BTC_HIP(hipMemcpyAsync(&h, b->result.p, 8, hipMemcpyDeviceToHost, s));
    pv.copy_out(pv32, pv64, s);  // synchronises s
    if (kernel_ms) *kernel_ms = tm.ms();
    *total = h;
}

}  // namespace

extern "C" {

int btc_bitbsr_count(btc_bitbsr b, int32_t variant, int32_t threshold_blocks, int32_t tau, void* stream,
                     uint64_t* total, int32_t* per_vertex, float* kernel_ms) {
    return btcx::guard(
        [&] { bitbsr_count(b, variant, threshold_blocks, tau, stream, total, per_vertex, nullptr, kernel_ms); });
}

int btc_bitbsr_count64(btc_bitbsr b, int32_t variant, int32_t threshold_blocks, int32_t tau, void* stream,
                       uint64_t* total, int64_t* per_vertex, float* kernel_ms) {
    return btcx::guard(
        [&] { bitbsr_count(b, variant, threshold_blocks, tau, stream, total, nullptr, per_vertex, kernel_ms); });
}

}  // extern "C"
