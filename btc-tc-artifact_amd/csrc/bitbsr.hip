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
#include "device_util.h"
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
    for (;;) {
        uint32_t nxt = p < e ? (uint32_t)col[p] / K : kNoTile;
        uint32_t cur = group_min_u32(nxt, 16);
        if (__all(cur == kNoTile)) break;
        if (cur != kNoTile) {
            RowBits<K> bits;
            while (p < e && (uint32_t)col[p] / K == cur) {
                bits.set((int)((uint32_t)col[p] % K));
                ++p;
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
                if (c == 0 && x && vtx < n) pv.add(vtx, x);
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
        if (q == 0 && rs && v < n) pv.add(v, rs);
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

// Tile columns of one row block as an LDS bitmap over column blocks with
// the number of set bits before each word: find(c) is the position of
// column c in the row block's sorted list (a0 + its rank), or kNoTile.
// Graphs with at most 32 kColSetWords column blocks use it.
constexpr uint32_t kColSetWords = 1344;
// Lists of at least this many dense tiles give each wave 4 consecutive
// tiles (the column set is reused while I stays; fewer waves leave more of
// the chip to the wedge kernel running beside): 2.86 vs 3.06 ms per count
// at R-MAT s22 (24 K tiles).  Shorter lists take one tile per wave, which
// keeps the chip busy (planted 2^17, 4 K tiles at K = 128: 4 per wave took
// 0.32 ms, one 0.09 ms).
constexpr uint32_t kDenseRunTiles = 16384;

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

// One dense mask tile t of row block I (tile columns [a0, a1)): returns this
// lane's part of the tile's count and adds per-vertex sums into pv.  The
// tile columns of I and of each row block J under the tile are matched
// through the column set cs (use_set) or by for_each_match.
template <int K, bool PV, bool MFMA>
__device__ __forceinline__ uint64_t dense_tile(int n, int n_rb, const uint32_t* __restrict__ indptr,
                                               const uint32_t* __restrict__ indices,
                                               const uint64_t* __restrict__ tiles, uint32_t t, uint32_t I, uint32_t a0,
                                               uint32_t a1, uint32_t set_words, ColumnSet& cs, int lane,
                                               const PvOut& pv) {
    constexpr int SLABS = K / 16;
    const uint32_t JL = uniform(indices[t]);
    // matches of I's and J's tile columns: through the column set (built
    // for I on first use) when streaming J's list is cheaper, else by search
    auto match = [&](uint32_t b0, uint32_t b1, auto&& f) {
        if (set_words && stream_cheaper(a1 - a0, b1 - b0)) {
            if (cs.I != I) {
                if (cs.I != kNoTile) cs.clear(set_words, lane);
                cs.a0 = a0;
                cs.build(indices, a1, set_words, lane);
                cs.I = I;
            }
            for_each_match_set(cs, indices, b0, b1, lane, f);
        } else {
            for_each_match(indices, a0, a1, b0, b1, lane, f);
        }
    };
    if constexpr (MFMA) {
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
            // matches are taken 4 at a time: their operand loads are issued
            // together, then the MFMAs (q0 is the newest of nq queued)
            uint32_t qa0 = 0, qb0 = 0, qa1 = 0, qb1 = 0, qa2 = 0, qb2 = 0, qa3 = 0, qb3 = 0;
            int nq = 0;
            auto drain = [&]() {
                const MfmaOp<K> a0 = mfma_load<K>(tiles, qa0, c), b0p = mfma_load<K>(tiles, qb0, c);
                MfmaOp<K> a1{}, b1p{}, a2{}, b2p{}, a3{}, b3p{};
                if (nq > 1) {
                    a1 = mfma_load<K>(tiles, qa1, c);
                    b1p = mfma_load<K>(tiles, qb1, c);
                }
                if (nq > 2) {
                    a2 = mfma_load<K>(tiles, qa2, c);
                    b2p = mfma_load<K>(tiles, qb2, c);
                }
                if (nq > 3) {
                    a3 = mfma_load<K>(tiles, qa3, c);
                    b3p = mfma_load<K>(tiles, qb3, c);
                }
                acc = mfma_mac<K>(a0, b0p, g, acc);
                if (nq > 1) acc = mfma_mac<K>(a1, b1p, g, acc);
                if (nq > 2) acc = mfma_mac<K>(a2, b2p, g, acc);
                if (nq > 3) acc = mfma_mac<K>(a3, b3p, g, acc);
                nq = 0;
            };
            auto f = [&](uint32_t pa, uint32_t pb) {
                qa3 = qa2;
                qb3 = qb2;
                qa2 = qa1;
                qb2 = qb1;
                qa1 = qa0;
                qb1 = qb0;
                qa0 = pa;
                qb0 = pb;
                if (++nq == 4) drain();
            };
            match(b0, b1, f);
            if (nq) drain();
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
                if (c == 0 && x && vtx < n) pv.add(vtx, x);
            }
        }
        return (uint64_t)(rs[0] + rs[1] + rs[2] + rs[3]);
    } else {
        const int r = lane >> 2, q = lane & 3;
        const RowBits<K> S = RowBits<K>::load(tiles, t, r);
        uint32_t sum = 0;
        for (int it = 0; it < SLABS; ++it) {
            const uint32_t Jo = JL * SLABS + it;
            if (Jo >= (uint32_t)n_rb) break;
            const uint32_t b0 = uniform(indptr[Jo]), b1 = uniform(indptr[Jo + 1]);
            const uint32_t m4 = S.bits(it * 16 + 4 * q) & 0xFu;
            uint32_t acc0 = 0, acc1 = 0, acc2 = 0, acc3 = 0;
            auto f = [&](uint32_t pa, uint32_t pb) {
                const RowBits<K> A = RowBits<K>::load(tiles, pa, r);
                acc0 += A.and_popc(RowBits<K>::load(tiles, pb, 4 * q + 0));
                acc1 += A.and_popc(RowBits<K>::load(tiles, pb, 4 * q + 1));
                acc2 += A.and_popc(RowBits<K>::load(tiles, pb, 4 * q + 2));
                acc3 += A.and_popc(RowBits<K>::load(tiles, pb, 4 * q + 3));
            };
            match(b0, b1, f);
            sum += ((m4 & 1u) ? acc0 : 0u) + ((m4 & 2u) ? acc1 : 0u) + ((m4 & 4u) ? acc2 : 0u) +
                   ((m4 & 8u) ? acc3 : 0u);
        }
        if (PV) {
            uint32_t rsum = sum + __shfl_xor(sum, 1);
            rsum += __shfl_xor(rsum, 2);
            const int64_t v = (int64_t)I * 16 + r;
            if (q == 0 && rsum && v < n) pv.add(v, rsum);
        }
        return (uint64_t)sum;
    }
}

// The engine's dense mask tiles (every listed tile on the dense path), as
// in tile_count_kernel's dense path.  Each wave takes tpw consecutive list
// entries (tile order, so mostly one row block I).  With
// set_words != 0 (column blocks <= 32 kColSetWords) I's tile columns go into
// an LDS ColumnSet, built when first needed for I, and the tile columns of
// a row block J stream through it where that is cheaper than searching;
// otherwise the two lists are matched by for_each_match.
template <int K, bool PV, bool MFMA>
__global__ __launch_bounds__(256) void dense_list_kernel(
    int n, int n_rb, const uint32_t* __restrict__ indptr, const uint32_t* __restrict__ indices,
    const uint32_t* __restrict__ row_indices, const uint64_t* __restrict__ tiles, uint32_t w0, uint32_t w_end,
    const uint32_t* __restrict__ tile_list, uint32_t lstride, uint32_t set_words, uint32_t tpw,
    unsigned long long* __restrict__ result, PvOut pv) {
    // dynamic LDS sized to the graph (dense_set_bytes): 4 waves' bits, then
    // their ranks
    extern __shared__ uint32_t dyn_lds[];
    const int lane = threadIdx.x & 63;
    const uint32_t wv = uniform(threadIdx.x >> 6);
    const uint32_t first = w0 + (blockIdx.x * 4 + wv) * tpw;
    if (first >= w_end) return;
    const uint32_t last = min(w_end, first + tpw);
    ColumnSet cs{dyn_lds + wv * set_words, reinterpret_cast<uint16_t*>(dyn_lds + 4 * set_words) + wv * set_words, 0};
    if (set_words)
        for (uint32_t x = (uint32_t)lane; x < set_words; x += 64) cs.bits[x] = 0;
    uint64_t tot = 0;
    for (uint32_t w = first; w < last; ++w) {
        const uint32_t t = uniform(tile_list[(size_t)w * lstride]);
        const uint32_t I = uniform(row_indices[t]);
        const uint32_t a0 = uniform(indptr[I]), a1 = uniform(indptr[I + 1]);
        tot += dense_tile<K, PV, MFMA>(n, n_rb, indptr, indices, tiles, t, I, a0, a1, set_words, cs, lane, pv);
    }
    tot = wave_sum_u64(tot);
    if (lane == 0 && tot) atomicAdd(result, (unsigned long long)tot);
}

template <int K>
__global__ __launch_bounds__(256) void tile_popcount_kernel(const uint64_t* __restrict__ tiles, uint32_t nblk,
                                                             uint32_t* __restrict__ pop) {
    constexpr int WPT = K / 4;
    const int lane = threadIdx.x & 63;
    const uint32_t w = uniform(blockIdx.x * 4 + (threadIdx.x >> 6));
    if (w >= nblk) return;
    uint32_t c = lane < WPT ? (uint32_t)__popcll(tiles[(size_t)w * WPT + lane]) : 0u;
    c = wave_sum_u32(c);
    if (lane == 0) pop[w] = c;
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

namespace {
template <int K, bool MF>
void launch_dense_k(const btc_bitbsr_s* b, const uint32_t* list, uint32_t lstride, uint32_t nt,
                    unsigned long long* res, PvOut pv, hipStream_t s) {
    // column blocks of the graph: ids of tile columns are below this
    const uint64_t ncb = ((uint64_t)b->n + K - 1) / K;
    const uint32_t set_words = ncb <= 32ull * kColSetWords ? (uint32_t)((ncb + 31) / 32) : 0u;
    // 4 waves x (set_words bit words + set_words 16-bit ranks), whole words
    const size_t lds = 4 * (size_t)set_words * 4 + (4 * (size_t)set_words * 2 + 3) / 4 * 4;
    const uint32_t tpw = nt >= kDenseRunTiles ? 4u : 1u;
    for (uint32_t w0 = 0; w0 < nt; w0 += kMaxTilesPerLaunch) {
        const uint32_t chunk = std::min(kMaxTilesPerLaunch, nt - w0);
        const unsigned grid = (chunk + 4 * tpw - 1) / (4 * tpw);
        if (pv.on())
            hipLaunchKernelGGL((dense_list_kernel<K, true, MF>), dim3(grid), dim3(256), lds, s, b->n, b->n_rb,
                               b->indptr.p, b->indices.p, b->row_indices.p, b->tiles.p, w0, w0 + chunk, list, lstride,
                               set_words, tpw, res, pv);
        else
            hipLaunchKernelGGL((dense_list_kernel<K, false, MF>), dim3(grid), dim3(256), lds, s, b->n, b->n_rb,
                               b->indptr.p, b->indices.p, b->row_indices.p, b->tiles.p, w0, w0 + chunk, list, lstride,
                               set_words, tpw, res, pv);
        BTC_CHECK_LAUNCH("dense_list_kernel");
    }
}
}  // namespace

void launch_dense_list(const btc_bitbsr_s* b, const uint32_t* tile_list, uint32_t ntiles, bool mfma,
                       unsigned long long* d_result, PvOut d_pv, hipStream_t s, uint32_t lstride) {
    if (ntiles == 0) return;
    BTC_REQUIRE(lane_sums_fit(b->max_row, b->K), BTC_ERR_OVERFLOW,
                "longest row %d: a lane's 32-bit sum over %d tile columns could overflow", b->max_row, b->K / 4);
    switch (b->K) {
        case 32: mfma ? launch_dense_k<32, true>(b, tile_list, lstride, ntiles, d_result, d_pv, s)
                      : launch_dense_k<32, false>(b, tile_list, lstride, ntiles, d_result, d_pv, s); break;
        case 64: mfma ? launch_dense_k<64, true>(b, tile_list, lstride, ntiles, d_result, d_pv, s)
                      : launch_dense_k<64, false>(b, tile_list, lstride, ntiles, d_result, d_pv, s); break;
        default: mfma ? launch_dense_k<128, true>(b, tile_list, lstride, ntiles, d_result, d_pv, s)
                      : launch_dense_k<128, false>(b, tile_list, lstride, ntiles, d_result, d_pv, s); break;
    }
}

namespace {
__global__ void zero_u32_kernel(uint32_t* __restrict__ p, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = 0;
}
}  // namespace

void launch_zero_u32(uint32_t* p, size_t n, hipStream_t s) {
    if (n == 0) return;
    const unsigned grid = (unsigned)std::min<size_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(zero_u32_kernel, dim3(grid), dim3(256), 0, s, p, n);
    BTC_CHECK_LAUNCH("zero_u32_kernel");
}

void launch_tile_popcount(const btc_bitbsr_s* b, uint32_t* d_pop, hipStream_t s) {
    if (b->nblk == 0) return;
    const unsigned grid = (b->nblk + 3) / 4;
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
