// Declarations shared between the library's device-side translation units.
#pragma once

#include <memory>

#include "common.h"
#include "prim.h"

// Bit-BSR of L with 16 x K tiles (K in {32, 64, 128}); see btc_amd.h.
struct btc_bitbsr_s {
    int32_t n = 0;
    int32_t n_rb = 0;
    int32_t K = 0;
    int32_t wpt = 0;  // 64-bit words per tile = K / 4
    uint32_t nblk = 0;
    int32_t max_row = 0;
    btcx::DevBuf<uint32_t> indptr;       // n_rb + 1
    btcx::DevBuf<uint32_t> indices;      // nblk
    btcx::DevBuf<uint32_t> row_indices;  // nblk
    btcx::DevBuf<uint64_t> tiles;        // nblk * wpt
    btcx::DevBuf<unsigned long long> result;
};

struct btc_pipeline_s;

struct btc_plan_s {
    btc_graph_s* g = nullptr;
    int K = 0;
    int tau = 0;
    uint32_t core_ids = 0;  // hub-core id range C of the row bitmaps (0 = none)
    bool use_tiny = false;  // tiny pivots become tiny items (btc_count_opts.tiny_items)
    bool tiny_all = false;  // tiny_items 2: every small pivot of at most 64 ids, whatever its spread
    btc_count_opts launch{};  // launch-only options (btc_plan_tune)
    // Bit-BSR of width K read by the dense-tile kernel (NULL when no tile is
    // dense): the plan's own, or a caller's Bit-BSR of g that must outlive
    // the plan (btc_plan_create_on_bitbsr)
    std::unique_ptr<btc_bitbsr_s> own_bsr;
    const btc_bitbsr_s* ext_bsr = nullptr;
    const btc_bitbsr_s* bsr = nullptr;
    btcx::DevBuf<uint32_t> dense_list;
    uint32_t n_dense = 0;
    btcx::DevBuf<uint4> dense_units;  // 2 x uint4 per (dense tile, slab) unit (build_dense_units)
    uint32_t n_units = 0;
    btcx::DevBuf<unsigned char> lists;  // items, list records, row core bitmaps, list rows
    btcx::DevBuf<uint4> items_dev;      // work items of a device-built plan (kept apart from lists)
    const uint4* items_p = nullptr;
    // a batch's per-graph sums: the wedge and tiny kernels add each item's
    // count at its graph index (the item's B.w; see batch.hip)
    unsigned long long* item_graph_sums = nullptr;
    const uint4* list_p = nullptr;
    const uint4* rowbits_p = nullptr;
    const int32_t* urow_p = nullptr;
    btcx::DevBuf<uint16_t> col16;  // low 16 bits of every id of L (m + padding)
    uint32_t nitems = 0;
    uint32_t n_block_items = 0;
    uint32_t n_tiny_items = 0;  // the last items: tiny pivots (tiny_pivot)
    btcx::DevBuf<unsigned long long> scratch;  // kWorkWords: [0] total, [1] work-item counters, slots
    btc_plan_info info{};
    float last_dense_ms = 0.f;
    float last_wedge_ms = 0.f;
    btcx::CapturedGraph count_graph;  // totals-only count, replayed when captured
    // second and third streams on which the dense-tile and tiny-pivot
    // kernels run beside the wedge kernel
    hipStream_t side = nullptr, side2 = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr, ev_join2 = nullptr;
    btc_plan_s() = default;
    btc_plan_s(const btc_plan_s&) = delete;
    btc_plan_s& operator=(const btc_plan_s&) = delete;
    ~btc_plan_s() {
        count_graph.reset();
        if (ev_fork) (void)hipEventDestroy(ev_fork);
        if (ev_join) (void)hipEventDestroy(ev_join);
        if (ev_join2) (void)hipEventDestroy(ev_join2);
        if (side) (void)hipStreamDestroy(side);
        if (side2) (void)hipStreamDestroy(side2);
    }
};

namespace btcx {

// Adds a partial per-vertex sum x to vertex v of pv (which is on).  The
// pointers are kernel arguments or plan fields, so the branch is uniform;
// int32 slots are used only where every count fits them.
__device__ __forceinline__ void pv_add(const PvOut& pv, int64_t v, uint64_t x) {
    if (pv.p64) atomicAdd(pv.p64 + v, (unsigned long long)x);
    else atomicAdd(pv.p32 + v, (int32_t)x);
}

constexpr uint32_t kNoTile = 0xFFFFFFFFu;

// Wedge-kernel plan layout shared by the host and device planners.
constexpr int kLdsWords = 4096;   // 16 KiB of LDS per workgroup
constexpr int kWaveWords = 1024;  // per-wave slice of it for wave items
// Wave items: pivots with at most this many prefix lists.
constexpr uint32_t kWaveItemLists = 64;
// Hub core: ids below C (a plan's core_ids, a multiple of 128, default
// kCoreIdsDefault).  Rows with at least kCoreRowMin core ids keep a C-bit
// bitmap of them (C / 32 words; 256 bytes for C = 2048).  2048 measured
// fastest on R-MAT s22 among 512..32768.
constexpr uint32_t kCoreIdsDefault = 2048;
constexpr uint32_t kCoreIdsMax = 32768;  // = the id span of a wave's LDS bitmap window
constexpr uint32_t kCoreRowMin = 32;
constexpr uint32_t kNoBitmap = 0xFFFFFFFFu;
// Pivots whose ids are all below this read their lists' prefixes from the
// 16-bit copy of L's columns (every prefix id is at most the pivot's last
// id), at half the bytes of the 32-bit ids.
constexpr uint32_t kNarrowIds = 1u << 16;

// Words of the core AND for a pivot whose ids span [lo, hi] (lo < C): from
// the 128-aligned window base to the 128-id block holding min(hi, C - 1),
// as the wedge kernel's loop runs.
__host__ __device__ inline uint32_t core_words(uint32_t lo, uint32_t hi, uint32_t C) {
    const uint32_t h = hi < C - 1 ? hi : C - 1;
    return ((h >> 7) + 1) * 4 - ((lo & ~127u) >> 5);
}

// A list whose row has a core bitmap ANDs the bitmap (`words` words of it)
// instead of looking up its `ids` core ids when 2 ids > words + 128.  On
// R-MAT s22 this rule gave the fastest wedge kernel among factors 1..4 and
// constants 0..256 (3.15 ms against 3.3..3.8 ms at C = 2048).
__host__ __device__ inline bool take_core_bitmap(uint32_t ids, uint32_t words) { return 2 * ids > words + 128; }

// A pivot's lists form one wave item when there are few of them and N_L(j)
// fits a wave's LDS slice as a bitmap window or a hash table.
// An LDS hash table in W words has 2^hash_log_max(W) slots and takes sets of
// up to half that many ids (load <= 1/2).
__host__ __device__ constexpr int hash_log_max(int W) {
    return W >= 4096 ? 12 : W >= 2048 ? 11 : W >= 1024 ? 10 : W >= 512 ? 9 : W >= 256 ? 8 : 7;
}
__host__ __device__ constexpr uint32_t hash_capacity(int W) { return (1u << hash_log_max(W)) / 2; }

__host__ __device__ inline bool wave_pivot(uint32_t lists, uint32_t lo, uint32_t hi, uint32_t d) {
    return lists <= kWaveItemLists && (hi - lo < 32u * kWaveWords || d <= hash_capacity(kWaveWords));
}

// Tiny pivots: wave pivots whose ids span too wide a range for a wave's
// bitmap window and that fit an LDS hash table of kTinyWords slots.  Their
// items are the plan's last; a single-plan count runs them in tiny_kernel,
// where each group of kTinyLanes lanes takes one pivot (8 per wave) so that
// the HBM round trips of many pivots overlap.  Multi-plan launches count
// them as wave items.
constexpr int kTinyWords = 128;
constexpr int kTinyLanes = 8;
// By default only graphs with at least this many wedges get tiny items
// (btc_count_opts.tiny_items overrides): the groups of
// a wave drift onto different paths and largely serialise, so the kernel
// pays only beside a long wedge kernel (R-MAT s22, 7.2e9 wedges: count
// 3.03 vs 3.07 ms); on planted 2^17 (9.2e7 wedges) it added 0.36 ms to a
// 0.14 ms wedge kernel.
constexpr int64_t kTinyMinWedges = int64_t(1) << 30;
// all: every such pivot, whatever the spread of its ids (batches of small
// graphs, whose wave items hold a few short lists each).
__host__ __device__ inline bool tiny_pivot(uint32_t lists, uint32_t lo, uint32_t hi, uint32_t d, bool all = false) {
    return lists <= kWaveItemLists && (all || hi - (lo & ~127u) >= 32u * kWaveWords - 1u) &&
           d <= hash_capacity(kTinyWords);
}

// A batch over an already block-diagonal canonical graph whose graphs
// occupy the vertex ranges [voff[k], voff[k + 1]) (takes g over).
btc_batch_s* batch_from_graph(btc_graph_s&& g, const std::vector<int64_t>& voff, const btc_count_opts* opts,
                              hipStream_t s);
// A batch over g (graphs on the vertex ranges voff) with an existing plan of
// g, which the batch takes over (plan->g is pointed at the batch's copy).
btc_batch_s* batch_from_plan(btc_graph_s&& g, btc_plan_s* plan, const std::vector<int64_t>& voff, hipStream_t s);
// Whether counts of batch b may fork their dense-tile and tiny-pivot
// kernels onto side streams of its plan (created on first use; default on).
void batch_set_side_streams(btc_batch_s* b, bool on);

// Graphs with more edges than this are planned on the GPU unless the
// options ask for the host planner.
constexpr int64_t kHostPlanEdges = 1 << 20;

// The edges of L a plan counts with the wedge kernel are cut into work
// items of about chunk checks.  Device planner for a graph resident in HBM
// (same plan contents as the host planner up to the order of lists within
// a pivot and the cut points of items).
void plan_device(btc_plan_s* p, int64_t chunk, hipStream_t s);
// The column array of L is padded so that 16-byte loads that start inside
// a row may run up to 3 ids past the last entry.
constexpr size_t kColPad = 16;

// ---------------------------------------------------------------- pipeline
// Device words of one capturable build + count (btc_pipeline): the host
// writes nnz and n_in before a replay and reads the block back after it.
struct PipeCtl {
    uint32_t nnz;     // input pairs
    int32_t n_in;     // vertices as given (<= 0: largest id + 1)
    uint32_t graphs;  // graphs of the run (vertex ranges segs[0, graphs]; 1 for a single graph)
    uint32_t host_reserved;
    int32_t lo, hi;  // smallest and largest id of the input
    uint32_t n;      // vertices of the input
    uint32_t status; // kPipe* bits: why the input was refused
    uint32_t m;      // nnz(L)
    uint32_t nc;     // vertices of L (after compaction)
    uint32_t ne;     // prefix lists
    uint32_t nbm;    // rows with a hub-core bitmap
    uint32_t nbm_words;
    uint32_t nitems;           // work items
    uint32_t item_counts[2];   // {workgroup items, workgroup + wave items}
    unsigned long long rs[3];  // longest row, sum C(d, 2), merge work
    unsigned long long st[8];  // planner statistics (as plan_device's)
    unsigned long long chunk;  // wedge checks per work item
    unsigned long long checks; // wedge checks in all
};
constexpr uint32_t kPipeNegativeId = 1, kPipeIdRange = 2, kPipeTooManyIds = 4;
constexpr uint32_t kPipeHostWords = 4;  // words of PipeCtl the host writes before a run
static_assert(sizeof(PipeCtl) % 8 == 0, "u64 fields aligned");

// Buffers of a capturable canonicalisation of up to nnz_cap pairs over up
// to n_cap ids (element counts in brackets).
struct CanonCap {
    const int32_t* src;
    const int32_t* dst;  // [nnz_cap] each
    int64_t nnz_cap;
    int32_t n_cap;
    uint32_t flags;      // BTC_CANON_*
    PipeCtl* ctl;
    uint64_t* ka;
    uint64_t* kb;        // [nnz_cap] each
    uint32_t* flag;      // [nnz_cap]
    uint32_t* pos;       // [nnz_cap + 1]
    uint32_t* used;      // [n_cap]
    uint32_t* newid;     // [n_cap + 1]
    uint32_t* deg;
    uint32_t* k0;
    uint32_t* v0;
    uint32_t* k1;
    uint32_t* v1;
    uint32_t* rank;      // [n_cap] each
    int32_t* lab0;
    int32_t* lab1;       // [n_cap] each
    uint32_t* rowptr;    // [n_cap + 1]
    int32_t* col;        // [nnz_cap + kColPad]
    uint16_t* col16;     // [nnz_cap + kColPad]: low 16 bits of col
    DevWs ws;            // dsort / dscan workspace for max(nnz_cap, n_cap)
    // several graphs (max_graphs > 1): graph g's ids are segs[g] ..
    // segs[g + 1] - 1 of the input (host-written), csegs[g] its first vertex
    // of L; the degree relabel keeps every graph a contiguous range
    int32_t max_graphs = 1;
    const uint32_t* segs = nullptr;  // [max_graphs + 1]
    uint32_t* csegs = nullptr;       // [max_graphs + 1]
};

// Enqueues the canonicalisation of ctl->nnz pairs (the same CSR as
// canonicalise_device makes) with kernels only; counts land in *ctl.
// Returns the buffer that receives the labels (lab0 or lab1).
int32_t* canonicalise_capacity(const CanonCap& c, hipStream_t s);

// Buffers of a capturable wedge-kernel plan of a canonical L of up to
// m_cap entries over up to n_cap rows (dense tiles and tiny items off).
struct PlanCap {
    const uint32_t* rowptr;  // [n_cap + 1]
    const int32_t* col;      // [m_cap + kColPad]
    int64_t m_cap;
    int32_t n_cap;
    uint32_t core_ids;       // C (0 = no hub-core bitmaps)
    int cus;                 // compute units (items of about 64 per CU)
    PipeCtl* ctl;
    uint32_t* erow;
    uint32_t* lflag;         // [m_cap] each
    uint32_t* lpos;          // [m_cap + 1]
    uint32_t* k0;
    uint32_t* v0;
    uint32_t* k1;
    uint32_t* v1;            // [m_cap] each
    uint32_t* off;           // [n_cap + 1]
    uint32_t* cend;
    uint32_t* has;           // [n_cap] each
    uint32_t* rbi;           // [n_cap + 1]
    uint4* lists;            // [m_cap]
    uint32_t* rowbits;       // [nbm_cap * C / 32]
    int32_t* ur;             // [m_cap]
    uint64_t* cost;          // [m_cap]
    uint64_t* P;             // [m_cap + 1]
    uint32_t* start;         // [m_cap]
    uint32_t* spos;          // [m_cap + 1]
    size_t items_cap;        // m_cap (every work item opens at a distinct list)
    uint32_t* first;         // [items_cap]
    uint4* tmp_items;        // [2 * items_cap]
    uint32_t* sk0;
    uint32_t* sv0;
    uint32_t* sk1;
    uint32_t* sv1;           // [items_cap] each
    uint4* items;            // [2 * items_cap]
    unsigned long long* work;  // the count's words (kWorkWords): total and cursors cleared with the items
    DevWs ws;                // dsort / dscan workspace for max(m_cap, items_cap)
};

// Rows that can hold a hub-core bitmap: at least kCoreRowMin entries each.
inline size_t pipe_bitmap_rows(int64_t m_cap, int32_t n_cap) {
    return (size_t)std::min<int64_t>(n_cap, m_cap / kCoreRowMin + 1);
}

// Enqueues the plan (as plan_device with tau above every tile and no tiny
// items; the item cut is about 64 items per CU with no upper bound on an
// item's checks) with kernels only; counts land in *ctl.
void plan_capacity(const PlanCap& c, hipStream_t s);

// btc_pipeline_count_batch for graphs whose ids the caller has checked
// (graph g: sizes[g] pairs at src[g] / dst[g], ids in [0, nv[g])), staged
// straight from the callers' arrays.
void pipeline_count_checked(btc_pipeline_s* p, int32_t count, const int32_t* const* src, const int32_t* const* dst,
                            const int64_t* sizes, const int64_t* nv, hipStream_t s, uint64_t* per_graph,
                            uint64_t* total);

// Checks the launch-shape fields of o (BTC_ERR_ARG on a bad value).
void check_launch_opts(const btc_count_opts& o);

// Enqueues the count of a plan_capacity plan (whose last kernel cleared the
// total and cursors at c.work; the slots are left cleared by every count):
// the wedge kernel on the item counts in ctl, then the slot sum into
// c.work[0].  col16 is L's 16-bit copy.  With pv set (several graphs),
// per_graph[g] (g < max_graphs, cleared first) gets graph g's count: each
// work item records its graph (the graph whose vertices [csegs[g],
// csegs[g + 1]) hold its pivot) and the wedge kernel adds the item's count
// there; with 64-bit offsets the per-vertex sums go to pv (cleared over
// ctl->nc vertices) and are summed per graph instead.
void pipeline_count_enqueue(const PlanCap& c, const uint16_t* col16, const btc_count_opts& o, hipStream_t s,
                            unsigned long long* pv = nullptr, const uint32_t* csegs = nullptr,
                            unsigned long long* per_graph = nullptr, int max_graphs = 1);

// Uploads L (rowptr narrowed to u32) to HBM on first use.
void ensure_device(btc_graph_s* g, hipStream_t s);

// out[0..n) = exclusive prefix sum of in[0..n), out[n] = total (u64).
void scan_u32_to_u64(const uint32_t* in, uint64_t* out, size_t n, hipStream_t s);

// Builds the 16 x K Bit-BSR of g on `s` into b (b's buffers are replaced).
void bitbsr_build(btc_graph_s* g, int K, hipStream_t s, btc_bitbsr_s* b);

// Sets p->bsr to the plan's Bit-BSR of width K: the caller's when the plan
// was made on one, else one built on s.  drop_plan_bsr releases it.
inline const btc_bitbsr_s* plan_bsr(btc_plan_s* p, hipStream_t s) {
    if (p->ext_bsr) {
        p->bsr = p->ext_bsr;
    } else {
        p->own_bsr = std::make_unique<btc_bitbsr_s>();
        bitbsr_build(p->g, p->K, s, p->own_bsr.get());
        p->bsr = p->own_bsr.get();
    }
    return p->bsr;
}
inline void drop_plan_bsr(btc_plan_s* p) {
    p->own_bsr.reset();
    p->bsr = nullptr;
}

// Count over mask tiles of b.  tile_list (device) selects ntiles mask tiles,
// or all tiles when NULL (list entries lstride apart: tile_list[w * lstride]).
// tau < 0 sends every tile down the dense path; mfma selects the i8 MFMA
// form of the dense path over AND + popcount.  Adds the total to *d_result
// and per-row sums to d_pv (when it is on).
void launch_tile_count(const btc_bitbsr_s* b, const uint32_t* tile_list, uint32_t ntiles,
                       bool binsearch, bool mfma, int tau, unsigned long long* d_result, PvOut d_pv,
                       hipStream_t s, uint32_t lstride = 1);

// Units of the dense tiles list[0, nt) of Bit-BSR b (one per tile slab with
// mask bits over a row block that has tile columns) into `units`; returns
// their number.  Synchronous on s.
uint32_t build_dense_units(const btc_bitbsr_s* b, const uint32_t* list, uint32_t nt, DevBuf<uint4>& units,
                           hipStream_t s);

// Counts dense units first + w * stride, w in [0, count), of Bit-BSR b: the
// engine's dense-tile path (i8 MFMA or AND + popcount), adding into the
// kSlots slots at d_slots and per-vertex sums into d_pv when it is on.
void launch_dense_units(const btc_bitbsr_s* b, const uint4* units, uint32_t first, uint32_t count, uint32_t stride,
                        bool mfma, unsigned long long* d_slots, PvOut d_pv, hipStream_t s);

// *d_result += sum of the kSlots slots at d_slots, which are cleared.
void launch_slot_sum(unsigned long long* d_slots, unsigned long long* d_result, hipStream_t s);

// p[0, n) = 0 by a kernel (safe inside captured graphs).
void launch_zero_u32(uint32_t* p, size_t n, hipStream_t s);

// Per-tile popcounts (device, nblk entries).
void launch_tile_popcount(const btc_bitbsr_s* b, uint32_t* d_pop, hipStream_t s);

// Number of compute units of the current device.
int device_cu_count();

// Device buffers reused across multi-plan launches.
struct MultiWorkspace {
    DevBuf<unsigned char> plans;  // device array of plan descriptors
    DevBuf<uint2> map;            // launch item -> {plan, item of the plan}
    DevBuf<unsigned long long> totals;
    DevBuf<uint32_t> counters;
    std::vector<uint2> h_map;
};

// Counts k plans with one persistent wedge-kernel launch (plus a dense-tile
// launch per plan that has dense tiles) on stream s; totals[j] = triangles
// of plans[j].  Synchronous.  Launch options are those of plans[0].
void plans_count(btc_plan_s* const* plans, int k, hipStream_t s, uint64_t* totals, MultiWorkspace& ws);

// Enqueues one exact count of plan p on stream s: zeroes the total and the
// work-item counters (16 bytes at d_work, default p->scratch.p), runs the
// dense-tile, wedge and tiny-pivot kernels; the u64 total lands in
// d_work[0] and per-row sums are added into d_pv when it is on.  tm times
// all kernels, tw the wedge kernel (either may be NULL).  With make_streams
// false the dense-tile and tiny-pivot kernels run beside the wedge kernel
// only when the plan's side streams exist already (no stream is created).
void plan_enqueue(btc_plan_s* p, hipStream_t s, PvOut d_pv, EventTimer* tm, EventTimer* tw,
                  unsigned long long* d_work = nullptr, bool make_streams = true);

}  // namespace btcx
