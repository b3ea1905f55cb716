/*
 * oracle.c -- CPU restatement of the reference's triangle-counting path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the
 * checker.  No product code links or calls it.
 *
 * Paths are relative to the reference repository (fanna1234/btc-tc-artifact).
 * Every function is written for clarity, not speed: sorting is qsort, tiles
 * are found by binary search, the brute force is a triple loop.
 *
 *   orc_canon         convert.h:281-324 (+ CPU variant :335-380), :387-425,
 *                     reorder.h:39-46; plus the degree relabel of this engine
 *   orc_tc_merge      triangle_count_cpu.h:60-101, with per-row sums
 *   orc_tc_brute      definition: #{i > j > k : ij, ik, jk edges}
 *   orc_bitbsr_build  bcsr_16x128.h:56-131,137-241 / bcsr_16x32.h:25-101,144-249
 *                     (K = 64 is the same procedure at width 64)
 *   orc_bitbsr_count  tc_16x128_mma_v5.h:9-211, tc_16x128_mma_v4.h:29-38,
 *                     tc_16x32_mma_v6.h:9-238, tc_16x32_mma_v5.h:27-36
 *   orc_lite_tile_k   apps/btc_tc_lite.cu:45-65
 *   orc_variant       tc_16x128_adaptive.h:12-29, tc_16x32_adaptive.h:14-31
 *
 * How each is pinned (tests/test_oracle.py): karate = 45 from the LAGraph
 * test suite (baselines/LAGraph/src/test/test_TriangleCount.c:38-51 with
 * graph_zachary_karate.h), closed forms (K_n, 2-D 9-point and 3-D
 * 27-point stencils, trees and bipartite graphs = 0), brute force against
 * merge on random graphs, the reference's own CPU verifier compiled from its
 * header (oracle/_ref), tau invariance of the tile count
 * (apps/tau_sweep.cu:529-537), tile-width and variant invariance, decode of
 * the Bit-BSR back to L, and set-based canonicalisation on tiny inputs.
 */
#include "oracle.h"

#include <stdlib.h>
#include <string.h>

static int cmp_u64(const void* a, const void* b) {
    uint64_t x = *(const uint64_t*)a, y = *(const uint64_t*)b;
    return x < y ? -1 : x > y;
}

static int cmp_i64(const void* a, const void* b) {
    int64_t x = *(const int64_t*)a, y = *(const int64_t*)b;
    return x < y ? -1 : x > y;
}

/* Sorted unique (hi << 32 | lo) keys -> CSR over n rows. */
static void keys_to_csr(const uint64_t* key, int64_t m, int32_t n, int64_t* rowptr, int32_t* col) {
    memset(rowptr, 0, sizeof(int64_t) * ((size_t)n + 1));
    for (int64_t e = 0; e < m; ++e) {
        rowptr[(key[e] >> 32) + 1]++;
        col[e] = (int32_t)(key[e] & 0xFFFFFFFFu);
    }
    for (int32_t i = 0; i < n; ++i) rowptr[i + 1] += rowptr[i];
}

int64_t orc_canon(const int32_t* src, const int32_t* dst, int64_t nnz, int32_t n, uint32_t flags,
                  int64_t** rowptr_out, int32_t** col_out, int32_t** labels_out, int32_t* n_out) {
    /* ids must lie in [0, n); n <= 0 means max id + 1 */
    int32_t mx = -1;
    for (int64_t e = 0; e < nnz; ++e) {
        if (src[e] < 0 || dst[e] < 0) return -1;
        if (src[e] > mx) mx = src[e];
        if (dst[e] > mx) mx = dst[e];
    }
    if (n <= 0) n = mx + 1;
    if (mx >= n) return -1;
    /* orient (u, v) -> (max, min), drop self-loops (reorder.h:39-46, convert.h:294-300) */
    uint64_t* key = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)(nnz ? nnz : 1));
    int64_t m = 0;
    for (int64_t e = 0; e < nnz; ++e) {
        uint32_t u = (uint32_t)src[e], v = (uint32_t)dst[e];
        if (u == v) continue;
        uint32_t hi = u > v ? u : v, lo = u > v ? v : u;
        key[m++] = ((uint64_t)hi << 32) | lo;
    }
    /* sort + unique (convert.h:312-314) */
    qsort(key, (size_t)m, sizeof(uint64_t), cmp_u64);
    int64_t k = 0;
    for (int64_t e = 0; e < m; ++e)
        if (k == 0 || key[e] != key[k - 1]) key[k++] = key[e];
    m = k;
    /* order-preserving compaction (convert.h:387-425); skipped for an empty graph (:392) */
    int32_t* newid = (int32_t*)malloc(sizeof(int32_t) * ((size_t)n + 1));
    int32_t nc = n;
    if ((flags & 1u) && m > 0) {
        char* used = (char*)calloc((size_t)n + 1, 1);
        for (int64_t e = 0; e < m; ++e) {
            used[key[e] >> 32] = 1;
            used[key[e] & 0xFFFFFFFFu] = 1;
        }
        nc = 0;
        for (int32_t v = 0; v < n; ++v) newid[v] = used[v] ? nc++ : -1;
        free(used);
    } else {
        for (int32_t v = 0; v < n; ++v) newid[v] = v;
    }
    int32_t* labels = (int32_t*)malloc(sizeof(int32_t) * ((size_t)nc + 1));
    for (int32_t v = 0; v < n; ++v)
        if (newid[v] >= 0) labels[newid[v]] = v;
    for (int64_t e = 0; e < m; ++e) {
        uint32_t hi = (uint32_t)newid[key[e] >> 32], lo = (uint32_t)newid[key[e] & 0xFFFFFFFFu];
        key[e] = ((uint64_t)hi << 32) | lo;
    }
    if (flags & 2u) {
        /* relabel by (-degree, id): rank[v] = position of v in that order */
        int64_t* deg = (int64_t*)calloc((size_t)nc + 1, sizeof(int64_t));
        for (int64_t e = 0; e < m; ++e) {
            deg[key[e] >> 32]++;
            deg[key[e] & 0xFFFFFFFFu]++;
        }
        /* sort key (-deg, id) packed as (maxdeg - deg) << 32 | id */
        int64_t* ord = (int64_t*)malloc(sizeof(int64_t) * ((size_t)nc + 1));
        for (int32_t v = 0; v < nc; ++v) ord[v] = ((int64_t)(0x7FFFFFFF - deg[v]) << 32) | v;
        qsort(ord, (size_t)nc, sizeof(int64_t), cmp_i64);
        int32_t* rank = (int32_t*)malloc(sizeof(int32_t) * ((size_t)nc + 1));
        int32_t* lab2 = (int32_t*)malloc(sizeof(int32_t) * ((size_t)nc + 1));
        for (int32_t r = 0; r < nc; ++r) {
            int32_t v = (int32_t)(ord[r] & 0xFFFFFFFF);
            rank[v] = r;
            lab2[r] = labels[v];
        }
        for (int64_t e = 0; e < m; ++e) {
            uint32_t a = (uint32_t)rank[key[e] >> 32], b = (uint32_t)rank[key[e] & 0xFFFFFFFFu];
            uint32_t hi = a > b ? a : b, lo = a > b ? b : a;
            key[e] = ((uint64_t)hi << 32) | lo;
        }
        qsort(key, (size_t)m, sizeof(uint64_t), cmp_u64);
        free(labels);
        labels = lab2;
        free(deg);
        free(ord);
        free(rank);
    }
    int64_t* rowptr = (int64_t*)malloc(sizeof(int64_t) * ((size_t)nc + 1));
    int32_t* col = (int32_t*)malloc(sizeof(int32_t) * (size_t)(m ? m : 1));
    keys_to_csr(key, m, nc, rowptr, col);
    free(key);
    free(newid);
    *rowptr_out = rowptr;
    *col_out = col;
    *labels_out = labels;
    *n_out = nc;
    return m;
}

void orc_free(void* p) { free(p); }

/* triangle_count_cpu.h:60-101: for each (r, u) in L merge N_L(r) and N_L(u). */
uint64_t orc_tc_merge(const int64_t* rowptr, const int32_t* col, int32_t n, int32_t threads,
                      int64_t* per_row) {
    uint64_t total = 0;
#pragma omp parallel for num_threads(threads > 0 ? threads : 1) reduction(+ : total) schedule(dynamic, 256)
    for (int32_t r = 0; r < n; ++r) {
        int64_t row = 0;
        for (int64_t i = rowptr[r]; i < rowptr[r + 1]; ++i) {
            int32_t u = col[i];
            int64_t pr = rowptr[r], pu = rowptr[u];
            while (pr < rowptr[r + 1] && pu < rowptr[u + 1]) {
                if (col[pr] < col[pu]) ++pr;
                else if (col[pr] > col[pu]) ++pu;
                else { ++row; ++pr; ++pu; }
            }
        }
        if (per_row) per_row[r] = row;
        total += (uint64_t)row;
    }
    return total;
}

/* #{i > j > k : (i,j), (i,k), (j,k) in L} by a dense adjacency matrix (n <= 4096). */
uint64_t orc_tc_brute(const int64_t* rowptr, const int32_t* col, int32_t n, int64_t* per_row) {
    if (n > 4096) return (uint64_t)-1;
    char* adj = (char*)calloc((size_t)n * (size_t)n + 1, 1);
    for (int32_t i = 0; i < n; ++i)
        for (int64_t p = rowptr[i]; p < rowptr[i + 1]; ++p) {
            adj[(size_t)i * n + col[p]] = 1;
            adj[(size_t)col[p] * n + i] = 1;
        }
    uint64_t total = 0;
    for (int32_t i = 0; i < n; ++i) {
        int64_t row = 0;
        for (int32_t j = 0; j < i; ++j) {
            if (!adj[(size_t)i * n + j]) continue;
            for (int32_t k = 0; k < j; ++k)
                if (adj[(size_t)i * n + k] && adj[(size_t)j * n + k]) ++row;
        }
        if (per_row) per_row[i] = row;
        total += (uint64_t)row;
    }
    free(adj);
    return total;
}

/* ------------------------------------------------------------ Bit-BSR */

/* Phase 1 (bcsr_16x128.h:150-220): i64 tile key (r/16) * ncb + c/K per edge,
 * sort, unique; indptr from the row-block counts.  Phase 2 (:225-240): every
 * edge binary-searches its tile in its row block and ORs bit
 * p = (r % 16) * K + c % K into word p / 64.  Returns the tile count; with
 * indptr == NULL only counts. */
int64_t orc_bitbsr_build(const int64_t* rowptr, const int32_t* col, int32_t n, int32_t K, uint32_t* indptr,
                         uint32_t* indices, uint32_t* row_indices, uint64_t* tiles) {
    if (K != 32 && K != 64 && K != 128) return -1;
    int64_t m = rowptr[n];
    int64_t ncb = ((int64_t)n + K - 1) / K;
    int64_t n_rb = ((int64_t)n + 15) / 16;
    int64_t* key = (int64_t*)malloc(sizeof(int64_t) * (size_t)(m ? m : 1));
    for (int32_t r = 0; r < n; ++r)
        for (int64_t p = rowptr[r]; p < rowptr[r + 1]; ++p) key[p] = (int64_t)(r / 16) * ncb + col[p] / K;
    qsort(key, (size_t)m, sizeof(int64_t), cmp_i64);
    int64_t nb = 0;
    for (int64_t e = 0; e < m; ++e)
        if (nb == 0 || key[e] != key[nb - 1]) key[nb++] = key[e];
    if (!indptr) {
        free(key);
        return nb;
    }
    memset(indptr, 0, sizeof(uint32_t) * (size_t)(n_rb + 1));
    for (int64_t t = 0; t < nb; ++t) {
        indices[t] = (uint32_t)(key[t] % ncb);
        row_indices[t] = (uint32_t)(key[t] / ncb);
        indptr[row_indices[t] + 1]++;
    }
    for (int64_t I = 0; I < n_rb; ++I) indptr[I + 1] += indptr[I];
    const int wpt = K / 4;
    memset(tiles, 0, sizeof(uint64_t) * (size_t)(nb * wpt));
    for (int32_t r = 0; r < n; ++r)
        for (int64_t p = rowptr[r]; p < rowptr[r + 1]; ++p) {
            int32_t c = col[p];
            uint32_t rb = (uint32_t)(r / 16), cb = (uint32_t)(c / K);
            uint32_t lo = indptr[rb], hi = indptr[rb + 1];
            while (lo < hi) {
                uint32_t mid = lo + (hi - lo) / 2;
                if (indices[mid] < cb) lo = mid + 1;
                else hi = mid;
            }
            uint32_t pos = (uint32_t)(r % 16) * (uint32_t)K + (uint32_t)(c % K);
            tiles[(size_t)lo * wpt + pos / 64] |= 1ull << (pos % 64);
        }
    free(key);
    return nb;
}

static int tile_bit(const uint64_t* tiles, int K, uint32_t t, int r, int c) {
    uint32_t pos = (uint32_t)r * (uint32_t)K + (uint32_t)c;
    return (int)((tiles[(size_t)t * (K / 4) + pos / 64] >> (pos % 64)) & 1u);
}

static int popc64(uint64_t x) {
    int c = 0;
    while (x) { x &= x - 1; ++c; }
    return c;
}

/* AND + popcount of row ra of tile a with row rb of tile b (K bits). */
static int row_dot(const uint64_t* tiles, int K, uint32_t a, int ra, uint32_t b, int rb) {
    int s = 0;
    for (int c = 0; c < K; ++c) s += tile_bit(tiles, K, a, ra, c) & tile_bit(tiles, K, b, rb, c);
    return s;
}

/* Count over all mask tiles with the reference kernels' control flow.  Each
 * mask tile S = (I, J_L) (I from row_indices, or by binary search over indptr
 * for variant 1 as in tc_16x128_mma_v4.h:29-38) counts its bits; more than
 * tau routes it to the dense path (tc_16x128_mma_v5.h:55-135: for each of the
 * K/16 row blocks J_out = (K/16) J_L + iter, two-pointer merge of the tile
 * columns of I and J_out, full 16 x 16 AND+popcount blocks per match, then
 * the mask), else to the sparse path (:136-201: per row, per set mask bit,
 * the same merge and popcounts).  per_row (length 16 * n_rb, may be NULL)
 * receives the sum of each row of L. */
uint64_t orc_bitbsr_count(int32_t K, int32_t n, const uint32_t* indptr, const uint32_t* indices,
                          const uint32_t* row_indices, const uint64_t* tiles, uint32_t nblk, int32_t tau,
                          int32_t variant, int64_t* per_row, int64_t* dense_tiles) {
    const int32_t n_rb = (n + 15) / 16;
    uint64_t total = 0;
    int64_t ndense = 0;
    if (per_row) memset(per_row, 0, sizeof(int64_t) * (size_t)n_rb * 16);
    for (uint32_t t = 0; t < nblk; ++t) {
        uint32_t I;
        if (variant == 1) {
            uint32_t lo = 0, hi = (uint32_t)n_rb;
            while (lo < hi) {
                uint32_t mid = (lo + hi) / 2;
                if (indptr[mid + 1] <= t) lo = mid + 1;
                else hi = mid;
            }
            I = lo;
        } else {
            I = row_indices[t];
        }
        uint32_t JL = indices[t];
        int nnz = 0;
        for (int w = 0; w < K / 4; ++w) nnz += popc64(tiles[(size_t)t * (K / 4) + w]);
        int dense = nnz > tau;
        ndense += dense;
        for (int iter = 0; iter < K / 16; ++iter) {
            int64_t Jo = (int64_t)JL * (K / 16) + iter;
            if (Jo >= n_rb) continue;
            int64_t acc[16][16];
            memset(acc, 0, sizeof(acc));
            uint32_t pa = indptr[I], pb = indptr[Jo];
            while (pa < indptr[I + 1] && pb < indptr[Jo + 1]) {
                if (indices[pa] < indices[pb]) ++pa;
                else if (indices[pa] > indices[pb]) ++pb;
                else {
                    for (int r = 0; r < 16; ++r)
                        for (int c = 0; c < 16; ++c) {
                            /* the sparse path evaluates only set mask bits; the dense
                             * path evaluates all and masks afterwards: same sum */
                            if (!dense && !tile_bit(tiles, K, t, r, iter * 16 + c)) continue;
                            acc[r][c] += row_dot(tiles, K, pa, r, pb, c);
                        }
                    ++pa;
                    ++pb;
                }
            }
            for (int r = 0; r < 16; ++r) {
                if ((int64_t)I * 16 + r >= n) continue;
                for (int c = 0; c < 16; ++c)
                    if (tile_bit(tiles, K, t, r, iter * 16 + c)) {
                        total += (uint64_t)acc[r][c];
                        if (per_row) per_row[(size_t)I * 16 + r] += acc[r][c];
                    }
            }
        }
    }
    if (dense_tiles) *dense_tiles = ndense;
    return total;
}

/* apps/btc_tc_lite.cu:45-65 */
int32_t orc_lite_tile_k(int32_t n, int64_t m, int32_t mode) {
    if (mode == 1) return 32;
    if (mode == 2) return 128;
    float avg = (float)m / (float)n;
    if (avg < 15.0f || n < 50000) return 32;
    return 128;
}

/* tc_16x128_adaptive.h:12-29 (2048), tc_16x32_adaptive.h:14-31 (40000):
 * 1 = binary-search variant, 2 = direct row lookup. */
int32_t orc_variant(int32_t K, uint32_t nblk, int32_t threshold_blocks) {
    if (threshold_blocks <= 0) threshold_blocks = K == 128 ? 2048 : 40000;
    return nblk < (uint32_t)threshold_blocks ? 1 : 2;
}
