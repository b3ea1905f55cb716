/*
 * btc_amd.h -- C ABI of the MI355X (gfx950) exact triangle-counting engine.
 *
 * Plain C: pointers, sizes and integer status codes; no C++ or torch types.
 * Every function returns BTC_OK (0) or a BTC_ERR_* code and never exits or
 * aborts the process; btc_last_error() gives the message of the last failing
 * call on the calling thread.
 *
 * Memory: inputs and outputs named "host" are caller-owned host buffers.
 * Device memory is owned by the library and lives behind the opaque handles
 * (btc_graph, btc_bitbsr, btc_plan, btc_batch); free each handle exactly once.
 * A handle may be used by one host thread at a time.
 *
 * Streams: every entry point that launches GPU work takes a `stream`
 * argument (a hipStream_t passed as void*, NULL = the default stream) and is
 * synchronous with respect to the caller: it returns after the results are
 * in the host buffers, like the reference's count calls
 * (btc/operations/tc_16x128_mma_v5.h:213-252).
 *
 * Reference interfaces replaced (paths relative to the reference repository):
 *   preprocess_for_triangle_counting      btc/operations/convert.h:601-692
 *   convert_coo_to_bcsr_16x128_gpu        btc/operations/bcsr_16x128.h:137-241
 *   convert_coo_to_bcsr_16x32_gpu         btc/operations/bcsr_16x32.h:144-249
 *   BCSR_16x128_Device / BCSR_16x32_Device btc/operations/bcsr_16x128.h:24-51,
 *                                         btc/operations/bcsr_16x32.h:109-138
 *   count_triangles_16x128_{v4,v5}        btc/operations/tc_16x128_mma_v4.h:222-260,
 *                                         btc/operations/tc_16x128_mma_v5.h:213-252
 *   count_triangles_16x32_{v5,v6}         btc/operations/tc_16x32_mma_v5.h:246-284,
 *                                         btc/operations/tc_16x32_mma_v6.h:240-279
 *   count_triangles_16x{128,32}_adaptive  btc/operations/tc_16x128_adaptive.h:12-29,
 *                                         btc/operations/tc_16x32_adaptive.h:14-31
 *   Lite tile policy                      apps/btc_tc_lite.cu:45-65
 *   cpu_tc_intersection_lower_triangular  btc/operations/triangle_count_cpu.h:60-101
 */
#ifndef BTC_AMD_H
#define BTC_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ status */
#define BTC_OK 0
#define BTC_ERR_ARG 1      /* invalid argument (null pointer, bad size, id out of range) */
#define BTC_ERR_HIP 2      /* HIP runtime error (launch, copy, no device) */
#define BTC_ERR_NOMEM 3    /* device or host allocation failed / over the HBM budget */
#define BTC_ERR_OVERFLOW 4 /* a 32-bit index or counter limit would be exceeded */
#define BTC_ERR_STATE 5    /* handle used in a state that does not allow the call */

/* Message of the last failing call on this thread ("" if none). */
const char* btc_last_error(void);
/* Library version string, "MAJOR.MINOR.PATCH-gfx950"; the reference is 0.1.0 (btc/btc.h:40-42). */
const char* btc_version(void);
/* Select the HIP device for calls from this thread; returns BTC_ERR_HIP if absent. */
int btc_device_init(int32_t device);

/* --------------------------------------------------------- synthetic inputs
 * Seeded generators (host, OpenMP).  Every random draw is
 * splitmix64(seed + (counter + 1) * 0x9E3779B97F4A7C15) for a counter fixed
 * by the generator, so outputs are identical for any thread count.  Each
 * returns the number of (src, dst) pairs; with src == dst == NULL only the
 * count is computed (two-call pattern).  Pairs are undirected and may hold
 * duplicates and self-loops; canonicalisation removes them.  Return -1 on a
 * bad argument (see btc_last_error). */
int64_t btc_gen_rmat(int32_t scale, int32_t edge_factor, double a, double b, double c,
                     uint64_t seed, int32_t* src, int32_t* dst);
int64_t btc_gen_er(int32_t n, double avg_degree, uint64_t seed, int32_t* src, int32_t* dst);
int64_t btc_gen_stencil2d(int32_t nx, int32_t ny, int32_t* src, int32_t* dst); /* 9-point */
int64_t btc_gen_stencil3d(int32_t nx, int32_t ny, int32_t nz, int32_t* src, int32_t* dst); /* 27-point */
/* n vertices in communities of `community` vertices with intra-community
 * edge probability p_in, plus G(n, noise_degree/(n-1)) noise. */
int64_t btc_gen_planted(int32_t n, int32_t community, double p_in, double noise_degree,
                        uint64_t seed, int32_t* src, int32_t* dst);

/* ------------------------------------------------------ canonical graph L
 * btc_graph_from_coo replaces preprocess_for_triangle_counting
 * (btc/operations/convert.h:601-692): orient every pair to (max, min), drop
 * self-loops, sort and deduplicate (convert.h:281-324), compact vertex ids
 * in order (convert.h:387-425, flag BTC_CANON_COMPACT), and optionally relabel
 * by descending degree with ties broken by id (BTC_CANON_DEGREE_ORDER, not in
 * the reference; the triangle total is invariant under it).  The result is
 * the strictly lower triangle L in CSR with ascending columns, resident in
 * HBM.  Ids must lie in [0, n); n <= 0 means n = max id + 1.
 * BTC_CANON_DEVICE does the same on the GPU (pairs staged in HBM, radix
 * sort of packed (max, min) keys); the result is identical bit for bit and
 * the host copy of L is made only when a host consumer needs it. */
#define BTC_CANON_COMPACT 1u
#define BTC_CANON_DEGREE_ORDER 2u
#define BTC_CANON_DEFAULT (BTC_CANON_COMPACT | BTC_CANON_DEGREE_ORDER)
#define BTC_CANON_DEVICE 4u

typedef struct btc_graph_s* btc_graph;

typedef struct {
    int32_t n;          /* vertices of L (after compaction) */
    int32_t n_input;    /* vertex id space of the input */
    int64_t m;          /* nnz(L) = undirected edges */
    int32_t max_row;    /* max_i d_L(i) */
    uint32_t flags;     /* canonicalisation flags used */
    int64_t wedges;     /* sum_i d_L(i)(d_L(i)-1)/2 */
    int64_t merge_work; /* W = sum over (i,j) in L of d_L(i) + d_L(j) */
} btc_graph_info;

int btc_graph_from_coo(const int32_t* src, const int32_t* dst, int64_t nnz, int32_t n,
                       uint32_t flags, btc_graph* out);
/* The same from pairs already in HBM (device pointers), canonicalised on the
 * GPU on `stream` — the reference's preprocess_for_triangle_counting on a
 * device COO (convert.h:601-692). */
int btc_graph_from_coo_device(const int32_t* src, const int32_t* dst, int64_t nnz, int32_t n,
                              uint32_t flags, void* stream, btc_graph* out);
/* Adopt an already canonical L (strictly lower, ascending, duplicate-free). */
int btc_graph_from_csr(const int64_t* rowptr, const int32_t* col, int32_t n, btc_graph* out);
int btc_graph_get_info(btc_graph g, btc_graph_info* info);
/* Copy L to host: rowptr[n+1], col[m]. */
int btc_graph_copy_csr(btc_graph g, int64_t* rowptr, int32_t* col);
/* canon_to_input[v] = input id of canonical vertex v (n entries). */
int btc_graph_copy_labels(btc_graph g, int32_t* canon_to_input);
void btc_graph_free(btc_graph g);

/* ------------------------------------------------------------- Bit-BSR
 * 16 x K bit tiles of L, K in {32, 64, 128}; replaces
 * convert_coo_to_bcsr_16x{128,32}_gpu (bcsr_16x128.h:137-241,
 * bcsr_16x32.h:144-249).  Layout as in the reference: tile (I, J) holds rows
 * 16I..16I+15 and columns K*J..K*J+K-1; bit p = r*K + c of a tile sits in
 * 64-bit word p/64, bit p%64 (little-endian, so byte-identical to the
 * reference's 32-bit words p/32, bit p%32).  Tile columns ascend within each
 * row block; indptr[n_row_blocks] = num_blocks. */
typedef struct btc_bitbsr_s* btc_bitbsr;

typedef struct {
    const uint32_t* indptr;      /* device, n_row_blocks + 1 */
    const uint32_t* indices;     /* device, num_blocks: tile column J */
    const uint32_t* row_indices; /* device, num_blocks: tile row block I */
    const uint64_t* tiles;       /* device, num_blocks * words_per_tile */
    int32_t n;
    int32_t n_row_blocks;
    uint32_t num_blocks;
    int32_t tile_k;
    int32_t words_per_tile; /* 16*K/64 */
} btc_bitbsr_view;

int btc_bitbsr_build(btc_graph g, int32_t tile_k, void* stream, btc_bitbsr* out);
int btc_bitbsr_get_view(btc_bitbsr b, btc_bitbsr_view* v);
int btc_bitbsr_copy_to_host(btc_bitbsr b, uint32_t* indptr, uint32_t* indices,
                            uint32_t* row_indices, uint64_t* tiles);
void btc_bitbsr_free(btc_bitbsr b);

/* Reference-semantics hybrid count over a Bit-BSR ("btc-plain"): one wave per
 * mask tile, in-kernel two-pointer matching of tile columns, popcount(tile) >
 * tau routes the tile to the dense AND+popcount path, else the sparse per-bit
 * path (tc_16x128_mma_v5.h:9-211, tc_16x32_mma_v6.h:9-238).
 *   variant: BTC_VARIANT_AUTO picks BINSEARCH when num_blocks <
 *            threshold_blocks, else DIRECT (tc_16x128_adaptive.h:12-29,
 *            tc_16x32_adaptive.h:14-31); threshold_blocks <= 0 takes the
 *            reference default for K (2048 for K=128, 40000 for K=32 and 64).
 *   tau:     < 0 takes the default for K (512 for 128, 64 for 32, 128 for 64).
 *   per_vertex: host int32[n] row sums of L .* (L L^T), or NULL.
 *   kernel_ms: event time of the count kernel(s) on `stream`, or NULL. */
#define BTC_VARIANT_AUTO 0
#define BTC_VARIANT_BINSEARCH 1
#define BTC_VARIANT_DIRECT 2
/* OR'ed into variant: dense tiles take the i8 MFMA form (0/1-unpacked tile
 * rows, v_mfma_i32_16x16x64_i8, exact in int32) instead of AND + popcount. */
#define BTC_DENSE_MFMA 16
int btc_bitbsr_count(btc_bitbsr b, int32_t variant, int32_t threshold_blocks, int32_t tau,
                     void* stream, uint64_t* total, int32_t* per_vertex, float* kernel_ms);
/* The same with int64 per-vertex counts.  Per-vertex sums are kept in int32
 * on the device while the longest row of L is at most 65536 (a vertex's
 * count is at most C(d, 2) < 2^31) and in 64 bits beyond; with int32 output
 * btc_bitbsr_count returns BTC_ERR_OVERFLOW only when a count does not fit,
 * this variant never.  Totals are u64 for any graph whose longest row is
 * below 2^32 / (K / 4) (a lane sums its row over K / 4 tile columns in 32
 * bits, like the reference's per-warp u32 reduction, tc_16x128_mma_v5.h:203-210). */
int btc_bitbsr_count64(btc_bitbsr b, int32_t variant, int32_t threshold_blocks, int32_t tau,
                       void* stream, uint64_t* total, int64_t* per_vertex, float* kernel_ms);

/* Lite tile policy (apps/btc_tc_lite.cu:45-65): returns 32 if m/n < 15 or
 * n < 50000, else 128; mode 1 forces 32, mode 2 forces 128 (BTC_MODE). */
int32_t btc_lite_tile_k(int32_t n, int64_t m, int32_t mode);
/* Default tau for a tile width (512, 128, 64 for K = 128, 64, 32). */
int32_t btc_default_tau(int32_t tile_k);

/* ------------------------------------------------------ hybrid count engine
 * The MI355X path for L .* (L L^T).  Edges of L in dense 16xK mask tiles
 * (popcount > tau) are counted over the Bit-BSR by units of one tile slab
 * (the 16 columns over one row block Jo): a wave matches the tile columns
 * of I and Jo through an LDS bitmap of I's columns (or a wave-parallel
 * binary search), copies the matched tiles into LDS (global_load_lds) and
 * adds each pair's A.B^T by v_mfma_i32_16x16x64_i8 on 0/1-unpacked tile rows
 * or by 64-bit AND + popcount (dense_kernel).  Every other edge (i, j) is
 * counted as the wedges (i; j, k), k < j, k in N_L(i): a persistent grid
 * takes work items (a pivot j and a range of rows holding it) in a static
 * stride over work-sorted items; a workgroup or a single wave loads N_L(j)
 * into LDS as a bitmap window (else a hash table, else it searches N_L(j)
 * in HBM) and streams the rows' prefixes through it, ANDing hub ids as
 * bitmaps.  Same result as btc_bitbsr_count and the CPU merge, bit for bit. */
typedef struct {
    int32_t tile_k;       /* dense tile width (32/64/128); 0 = auto */
    int32_t tau;          /* dense threshold; < 0 = default for tile_k; >= 16*K disables dense */
    int32_t chunk_work;   /* wedge checks per work item; 0 = auto */
    int32_t waves_per_eu; /* resident workgroups (4 waves) per CU of the persistent grid; 0 = max */
    int32_t lookup;       /* 0 = auto, 1 = never the LDS bitmap, 2 = never the LDS hash */
    int32_t group_lanes;  /* lanes streaming one prefix list: 4, 8, 16 or 32; 0 = auto (4) */
    int32_t unroll;       /* 16-byte column loads in flight per lane: 1, 2 or 4; 0 = auto */
    int32_t schedule;     /* 0 = work stealing on one counter, 1 = static stride over sorted items */
    int32_t dense_kernel; /* dense tiles: 0 = auto (AND + popcount, the faster form at every K measured),
                             1 = 64-bit AND + v_bcnt popcount, 2 = i8 MFMA (v_mfma_i32_16x16x64_i8) on
                             0/1-unpacked tile rows */
    int32_t planner;      /* work lists built by: 0 = auto (the GPU for graphs canonicalised there or with
                             more than 2^20 edges), 1 = host threads, 2 = GPU kernels */
    int32_t core_ids;     /* hub-core id range C of the rows' core bitmaps: a multiple of 128 up to
                             32768; 0 = 2048, < 0 = no core bitmaps */
    int32_t tiny_items;   /* pivots with few lists, at most 64 ids and ids spread wider than a wave's
                             bitmap window counted by groups of 8 lanes in their own kernel: 0 = auto
                             (graphs with at least 2^30 wedges; for batches 2), 1 = always, 2 = every
                             pivot with few lists and at most 64 ids, whatever their spread, -1 =
                             never (they stay wave items) */
    int32_t wide_offsets; /* 1 = the wedge kernel addresses L with 64-bit byte offsets (in its default
                             shape); 0 = only when L holds 2^30 ids or more, where 32-bit offsets end */
} btc_count_opts;

typedef struct btc_plan_s* btc_plan;

typedef struct {
    int64_t sparse_edges;      /* edges counted by the wedge path */
    int64_t dense_edges;       /* edges inside dense mask tiles */
    int64_t dense_tiles;       /* dense mask tiles */
    int64_t wedge_checks;      /* hash lookups the wedge path performs */
    int64_t work_items;        /* pivot chunks */
    int64_t device_bytes;      /* HBM held by the plan */
    int64_t sparse_merge_work; /* sum over sparse edges (i,j) of d_L(i) + d_L(j) */
    int64_t stream_bytes;      /* bytes the wedge kernel reads: prefixes, pivot rows, lists, items */
    int64_t core_lists;        /* lists whose hub-core ids are counted by bitmap AND + popcount */
    int64_t core_rows;         /* rows holding a hub-core bitmap */
    int64_t core_ids;          /* hub-core id range C of the plan (0 = no core bitmaps) */
    int32_t tile_k;
    int32_t tau;
    int64_t tiny_items;        /* work items of tiny pivots: few lists, a small row and ids spread wider
                                  than a wave's bitmap window; one group of 8 lanes counts each */
} btc_plan_info;

void btc_count_opts_default(btc_count_opts* opts);
int btc_plan_create(btc_graph g, const btc_count_opts* opts, void* stream, btc_plan* out);
/* The same on a Bit-BSR b already built from g (btc_bitbsr_build): the dense
 * tiles are taken from b instead of a Bit-BSR of the plan's own, and the
 * plan's tile width is b's (opts->tile_k must be 0 or b's).  b must outlive
 * the plan.  This is how the reference's count_triangles_16x{128,32}_adaptive
 * (tc_16x128_adaptive.h:12-29, tc_16x32_adaptive.h:14-31), which count a
 * BCSR_16x{128,32}_Device, reach the engine. */
int btc_plan_create_on_bitbsr(btc_graph g, btc_bitbsr b, const btc_count_opts* opts, void* stream,
                              btc_plan* out);
int btc_plan_get_info(btc_plan p, btc_plan_info* info);
/* Changes the launch-only options of a plan (waves_per_eu, lookup, dense_kernel,
 * group_lanes, unroll, schedule); tile_k, tau and chunk_work need a new plan. */
int btc_plan_tune(btc_plan p, const btc_count_opts* opts);
int btc_plan_count(btc_plan p, void* stream, uint64_t* total, int32_t* per_vertex, float* kernel_ms);
/* The same with int64 per-vertex counts (see btc_bitbsr_count64 for when
 * int32 output overflows). */
int btc_plan_count64(btc_plan p, void* stream, uint64_t* total, int64_t* per_vertex, float* kernel_ms);
/* btc_plan_count that allocates nothing.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory (code under stream
 * capture, tight serving loops).  The total, the work-item counters, the
 * partial-sum slots and the per-vertex sums live in a static device array of
 * the library (16 MiB + 8.25 KiB, held on each device it runs on until the
 * process ends), so
 * per_vertex (host int32[n], or NULL) is limited to graphs of at most 2^22
 * vertices (BTC_ERR_ARG beyond) whose longest L-row is at most 65536
 * (BTC_ERR_OVERFLOW beyond; btc_plan_count64 gives those); totals have no
 * such limit.  Calls that overlap on one device share the array and race.
 * No kernel_ms (no events are created); the dense-tile and tiny-pivot
 * kernels run beside the wedge kernel only once an earlier btc_plan_count
 * of the plan has created its side streams.  Synchronous on `stream`. */
int btc_plan_count_noalloc(btc_plan p, void* stream, uint64_t* total, int32_t* per_vertex);
/* Event times of the last btc_plan_count with kernel_ms != NULL: wedge_ms
 * = the wedge kernel, dense_ms = the rest of the count (dense-tile and
 * tiny-pivot kernels, beside or around the wedge kernel; 0 when neither is
 * launched; both -1 after a replayed capture). */
int btc_plan_last_timing(btc_plan p, float* dense_ms, float* wedge_ms);
/* Counts share `share` of a `shares`-way split of the plan's work: the dense
 * units (tile slabs, in tile order) and the work items (of each class, sorted
 * by work) whose index is share modulo shares — what one GPU of a split
 * would run.  The totals of
 * all shares add up to the plan's total.  No reference counterpart (the
 * reference counts on one GPU). */
int btc_plan_count_share(btc_plan p, int32_t shares, int32_t share, void* stream, uint64_t* total,
                         float* kernel_ms);
/* Captures the totals-only count of p (zero fill, dense-tile, wedge and
 * tiny-pivot kernels) into a HIP graph; later btc_plan_count calls without per_vertex
 * replay it with one hipGraphLaunch on their stream (kernel_ms then times
 * the replay and btc_plan_last_timing reports -1 for the split).
 * btc_plan_tune drops the graph.  No reference counterpart (the reference
 * launches every kernel from the host each time). */
int btc_plan_capture(btc_plan p, void* stream);
void btc_plan_free(btc_plan p);
/* One-shot: plan + count + free. */
int btc_count(btc_graph g, const btc_count_opts* opts, void* stream, uint64_t* total,
              int32_t* per_vertex, float* kernel_ms);

/* ------------------------------------------------------------ autotuner
 * Times candidate plans (K in {32, 64, 128} selected by tile_ks_mask bits
 * 1/2/4, tau in {default, 2 x default, 16K = no dense tiles}), launch shapes
 * (lanes per list, loads in flight, resident workgroups per CU, static
 * stride or work stealing), the dense-tile kernel form and the hub-core
 * size (1024, 2048, 4096, none) on
 * the device within budget_ms and returns the fastest options.  The winner
 * is recorded in a process-wide table keyed by log2 of n, m, wedges and the
 * longest row; btc_plan_create fills options left on auto (tile_k 0, tau
 * < 0, group_lanes 0, unroll 0, waves_per_eu 0, dense_kernel 0, core_ids 0)
 * from that table (the schedule is returned in `best` but not tabled).
 * Replaces the reference's fixed Lite rule and tau literals. */
typedef struct {
    int32_t budget_ms;    /* default 4000 */
    int32_t reps;         /* timed counts per candidate, best kept; default 3 */
    int32_t tile_ks_mask; /* default 7 */
    int32_t reserved;
} btc_tune_opts;

void btc_tune_opts_default(btc_tune_opts* t);
int btc_autotune(btc_graph g, const btc_tune_opts* t, void* stream, btc_count_opts* best, float* best_ms);
/* BTC_OK and the tuned options when the table holds g's feature class. */
int btc_tune_lookup(btc_graph g, btc_count_opts* opts);
int btc_tune_table_save(const char* path);
int btc_tune_table_load(const char* path);
void btc_tune_table_clear(void);

/* ------------------------------------------------------------- batching
 * Many graphs counted by one persistent launch: their L's are placed on the
 * diagonal of one block-diagonal L (each graph's vertices start at a
 * multiple of 128 so no tile spans two graphs), planned once and counted
 * by one enqueue of the engine's kernels; per-graph totals are segment sums
 * of per-vertex counts.  The input graphs may be freed after creation.  No
 * counterpart in the reference, which counts one graph per process. */
typedef struct btc_batch_s* btc_batch;

typedef struct {
    int32_t graphs;
    int32_t n;  /* vertices of the block-diagonal L, padding included */
    int64_t m;
    btc_plan_info plan;
} btc_batch_info;

int btc_batch_create(const btc_graph* graphs, int32_t count, const btc_count_opts* opts, void* stream,
                     btc_batch* out);
/* A batch straight from the graphs' pairs (host buffers): graph k's pairs are
 * src/dst[edge_offsets[k], edge_offsets[k+1]) with ids in [0, n_per_graph[k])
 * (n <= 0: max id + 1).  All graphs are canonicalised in one GPU pass as
 * separate vertex ranges (each compacted and relabelled by degree within
 * itself, per `flags`) and planned as one block-diagonal graph. */
int btc_batch_from_coo(const int32_t* src, const int32_t* dst, const int64_t* edge_offsets,
                       const int32_t* n_per_graph, int32_t count, uint32_t flags, const btc_count_opts* opts,
                       void* stream, btc_batch* out);
int btc_batch_get_info(btc_batch b, btc_batch_info* info);
int btc_batch_tune(btc_batch b, const btc_count_opts* opts);
/* per_graph: host uint64[count] or NULL; total: sum over the batch. */
int btc_batch_count(btc_batch b, void* stream, uint64_t* per_graph, uint64_t* total, float* kernel_ms);
/* Runs btc_autotune on the batch's block-diagonal graph and re-plans the
 * batch with the winning options (returned in best), unless the batch's
 * current plan counts the batch faster (timed batch counts of both); best
 * then returns the current plan's options. */
int btc_batch_autotune(btc_batch b, const btc_tune_opts* topts, void* stream, btc_count_opts* best,
                       float* best_ms);
/* Captures a batch count (zero fill, kernels, per-graph sums) into a HIP
 * graph that later btc_batch_count calls replay; btc_batch_tune drops it. */
int btc_batch_capture(btc_batch b, void* stream);
void btc_batch_free(btc_batch b);

/* ------------------------------------------------------------- pipeline
 * The whole path from COO pairs to the triangle total -- the device
 * canonicalisation (as BTC_CANON_DEVICE), the wedge-kernel plan and the
 * count -- captured once into a HIP graph for inputs of up to n_cap ids and
 * nnz_cap pairs, then replayed for each input with one hipGraphLaunch:
 * every kernel reads the live sizes (pairs, nnz(L), vertices, lists, work
 * items) from device memory, so a run makes no allocation and no host
 * round trip between its steps.  Buffers (about 180 bytes per pair and
 * 60 per id of capacity) are allocated once, at creation.  The plan has no
 * dense tiles and no tiny items (the wedge kernel counts every edge), cut
 * into about 64 work items per CU; of opts, group_lanes, unroll, lookup,
 * schedule, waves_per_eu, core_ids and wide_offsets apply.  Counts equal
 * btc_count on btc_graph_from_coo_device's graph of the same pairs.  No
 * reference counterpart: the reference runs preprocess_for_triangle_counting
 * (convert.h:601-692), the Bit-BSR build (bcsr_16x128.h:137-241) and the
 * count from the host for every graph. */
typedef struct btc_pipeline_s* btc_pipeline;

typedef struct {
    int32_t n_cap;
    int32_t graph_nodes;  /* nodes of the captured HIP graph */
    int64_t nnz_cap;
    int64_t device_bytes; /* buffers held for the capacity */
    uint32_t flags;       /* BTC_CANON_* of the canonicalisation */
    /* the last accepted run */
    int32_t n;            /* vertices of L */
    int64_t m;            /* nnz(L) */
    int64_t wedges;       /* sum_i d_L(i)(d_L(i)-1)/2 */
    int64_t wedge_checks; /* ids looked up + core words ANDed by the plan */
    int32_t work_items;
    int32_t max_row;
    int32_t max_graphs;   /* graphs one run may hold */
    int32_t reserved;
} btc_pipeline_info;

int btc_pipeline_create(int32_t n_cap, int64_t nnz_cap, uint32_t flags, const btc_count_opts* opts, void* stream,
                        btc_pipeline* out);
/* A pipeline whose runs hold up to max_graphs (<= 64) graphs with n_cap ids
 * and nnz_cap pairs in all: canonicalised as separate vertex ranges (each
 * compacted and relabelled by degree within itself, as btc_batch_from_coo),
 * planned as one block-diagonal L and counted with per-vertex sums summed
 * per graph.  With BTC_CANON_DEGREE_ORDER, bits(n_cap) + bits(max_graphs)
 * must not exceed 32. */
int btc_pipeline_create_batch(int32_t n_cap, int64_t nnz_cap, int32_t max_graphs, uint32_t flags,
                              const btc_count_opts* opts, void* stream, btc_pipeline* out);
/* Triangles of the graph of nnz pairs (src[e], dst[e]) over ids [0, n)
 * (n <= 0: the largest id + 1): host arrays, or HBM when device_pairs is 1.
 * An id out of range, nnz > nnz_cap or more than n_cap ids returns
 * BTC_ERR_ARG, as btc_graph_from_coo_device would.  Runs on one pipeline
 * are serialised; kernel_ms (may be NULL) times the replay. */
int btc_pipeline_count(btc_pipeline p, const int32_t* src, const int32_t* dst, int64_t nnz, int32_t n,
                       int32_t device_pairs, void* stream, uint64_t* total, float* kernel_ms);
/* count graphs in one run (host pairs, the layout of btc_batch_from_coo:
 * graph k's pairs are src/dst[edge_offsets[k], edge_offsets[k+1]) over ids
 * [0, n_per_graph[k]), n <= 0: its largest id + 1; ids are checked on the
 * host); per_graph[k] (may be NULL) = graph k's triangles, *total = all. */
int btc_pipeline_count_batch(btc_pipeline p, const int32_t* src, const int32_t* dst, const int64_t* edge_offsets,
                             const int32_t* n_per_graph, int32_t count, void* stream, uint64_t* per_graph,
                             uint64_t* total, float* kernel_ms);
/* A copy of the last accepted run's canonical graph, in HBM: the same L and
 * labels as btc_graph_from_coo_device of the same pairs. */
int btc_pipeline_graph(btc_pipeline p, void* stream, btc_graph* out);
/* The last accepted run's L and plan, copied into buffers of their own:
 * *graph is btc_pipeline_graph's graph, *plan a plan of it (no dense tiles,
 * no tiny items) that btc_plan_count and the others count like any plan,
 * unaffected by later runs.  Free the plan before the graph. */
int btc_pipeline_export(btc_pipeline p, void* stream, btc_graph* graph, btc_plan* plan);
/* The last run of a pipeline of several graphs as a batch (btc_batch_count
 * gives its per-graph counts again), independent of later runs. */
int btc_pipeline_export_batch(btc_pipeline p, void* stream, btc_batch* out);
int btc_pipeline_get_info(btc_pipeline p, btc_pipeline_info* info);
void btc_pipeline_free(btc_pipeline p);


/* -------------------------------------------------------------- serving
 * A request loop with continuous batching.  btc_server_submit queues an
 * exact-count query for a graph identified by the caller's key; a graph not
 * yet resident is copied, canonicalised and planned by a pool of host prep
 * threads on their own HIP streams, admitted only when the resident set
 * stays under hbm_budget_bytes (idle graphs are evicted least recently used
 * first); later queries for a resident key skip all of that.  A GPU worker
 * counts every request whose plan is resident in one persistent launch over
 * all their plans, while the prep threads keep preparing new arrivals.  Each
 * result is exact and the same as the graph's count alone.  Times are
 * seconds on the server's clock.  No counterpart in the reference. */
typedef struct btc_server_s* btc_server;

typedef struct {
    int64_t hbm_budget_bytes; /* resident graphs + plans; default 64 GiB */
    int32_t max_batch;        /* requests per launch; default 512 */
    int32_t prep_threads;     /* host threads that canonicalise and plan; default 4 */
    int32_t canon_flags;      /* default BTC_CANON_DEFAULT | BTC_CANON_DEVICE (canonicalise and plan on the
                                 GPU, on each preparation thread's stream) */
    int32_t group_graphs;     /* with GPU preparation (and, when captured_prep is 1, a backlog), a prep
                                 thread takes up to this many queued new graphs and prepares them as one
                                 unit; default 16, 1 = one at a time */
    int32_t captured_prep;    /* 1 (default, with GPU preparation): while at most eight new graphs per
                                 prep thread are queued, a prep thread builds and counts each new graph of up to
                                 2^23 pairs with one replay of a btc_pipeline of its size class (ids and
                                 pairs rounded up to powers of two), answers the requests waiting for it at
                                 once, then keeps a copy of its L and plan resident for later requests
                                 (btc_pipeline_export); a longer queue is prepared in groups of up to
                                 group_graphs graphs, each group by one run of a pipeline for several graphs
                                 (btc_pipeline_count_batch, then btc_pipeline_export_batch).
                                 0: eager preparation (grouped), counts by the GPU worker */
} btc_server_opts;

typedef struct {
    uint64_t request_id;
    uint64_t graph_key;
    uint64_t triangles;
    int32_t status;       /* BTC_OK or the error of canonicalisation / planning / counting */
    int32_t resident_hit; /* 1: the graph was resident or already being prepared */
    int32_t batch_size;   /* requests in the launch that counted it */
    int32_t reserved;
    double t_submit, t_start, t_done;
} btc_result;

typedef struct {
    int64_t submitted, completed, failed, resident_hits, prepared, evictions, launches, launched_requests;
    int64_t resident_bytes, resident_graphs;
    double clock;
} btc_server_stats;

void btc_server_opts_default(btc_server_opts* opts);
int btc_server_create(const btc_server_opts* opts, btc_server* out);
/* src/dst/n describe the graph (copied); they may be NULL/0 when graph_key
 * was submitted before. */
int btc_server_submit(btc_server s, uint64_t graph_key, const int32_t* src, const int32_t* dst, int64_t nnz,
                      int32_t n, uint64_t* request_id);
/* Moves up to `capacity` finished results into out, waiting up to wait_ms
 * for the first one; *count = results returned. */
int btc_server_poll(btc_server s, btc_result* out, int32_t capacity, int32_t wait_ms, int32_t* count);
int btc_server_get_stats(btc_server s, btc_server_stats* stats);
void btc_server_free(btc_server s);

/* --------------------------------------------------------- CPU verify path
 * Host OpenMP merge over L (the reference's -v 1 path,
 * triangle_count_cpu.h:60-101); per_vertex may be NULL.  Not used by any GPU
 * entry point. */
int btc_count_cpu(btc_graph g, int32_t threads, uint64_t* total, int32_t* per_vertex);

#ifdef __cplusplus
}
#endif

#endif /* BTC_AMD_H */
