// Hybrid count engine for L .* (L L^T) on gfx950.
//
// sum over (i, j) in L of |N_L(i) & N_L(j)| is split by edge:
//
//  * edges inside dense 16 x K mask tiles (popcount > tau) go through the
//    Bit-BSR tile path (launch_tile_count with a list of dense tiles, every
//    tile on the dense AND+popcount path);
//
//  * every other edge (i, j) is evaluated as the wedges (i; j, k) with k a
//    neighbour of i below j: k is in N_L(j) exactly when (i, j, k) is a
//    triangle.  Because N_L(i) is sorted, those k are the prefix of row i
//    in front of j, clipped to [min N_L(j), max N_L(j)].  Edges are grouped
//    by pivot j; a workgroup (or one wave, for pivots with few prefixes)
//    loads N_L(j) once into LDS as a bitmap window (else a hash table, else
//    it stays in HBM for binary search) and streams the prefixes of all
//    rows i that hold j through it with 16-byte loads.  Hub ids (below
//    the plan's core size C) of rows with many of them are counted by ANDing the row's
//    core bitmap with the window.  With descending-degree labels this checks
//    sum_i C(d_L(i), 2) wedges at most, far below the sum of d_L(i) +
//    d_L(j) a merge would read.
//
// Work items are (pivot, range of prefixes) chunks of bounded size, taken
// by a persistent grid in a static stride over work-sorted items (or from
// one atomic counter), heaviest first.  Per-vertex sums of row i collect
// the prefix counts of row i with int32 atomics.  The host planner below
// builds the items for graphs held on the host; plan_gpu.hip builds the
// same plan on the GPU.
#include <omp.h>

#include <algorithm>
#include <atomic>
#include <cstring>
#include <mutex>
#include <numeric>

#include "device_util.h"
#include "internal.h"

namespace btcx {
namespace {

constexpr int kThreads = 256;
constexpr uint32_t kEmpty = 0xFFFFFFFFu;
// Plans with at least this many wedge checks (a wedge kernel of roughly
// 20 us or more) run the dense-tile kernel beside it on a second stream.
constexpr int64_t kOverlapChecks = 1 << 24;

__device__ __forceinline__ uint32_t hslot(uint32_t key, int shift) { return (key * 2654435761u) >> shift; }

// 16-byte load that does not keep the line in the caches (read-once data).
__device__ __forceinline__ uint4 load_nt(const uint4* p) {
    typedef unsigned int v4u __attribute__((ext_vector_type(4)));
    const v4u v = __builtin_nontemporal_load(reinterpret_cast<const v4u*>(p));
    return make_uint4(v.x, v.y, v.z, v.w);
}

// Pointers into the global address space.  Every pointer of a plan points
// to device memory, but the compiler cannot tell for pointers read from a
// plan descriptor and emits flat loads, which also count against the LDS
// wait counter, so each LDS wait would wait for them too.  Base pointers
// are cast once and indexed in the global space (a cast applied after the
// address arithmetic does not survive optimisation).
// (Builtin vector types: HIP's uint2 / uint4 classes cannot be copied out
// of an address-space qualified object.)
typedef unsigned int V2U __attribute__((ext_vector_type(2)));
typedef unsigned int V4U __attribute__((ext_vector_type(4)));
typedef const char __attribute__((address_space(1)))* GlobalChar;
typedef const int32_t __attribute__((address_space(1)))* GlobalI32;
typedef const V2U __attribute__((address_space(1)))* GlobalU2;
typedef const V4U __attribute__((address_space(1)))* GlobalU4;
__device__ __forceinline__ uint2 to_u2(V2U v) { return make_uint2(v.x, v.y); }
__device__ __forceinline__ uint4 to_u4(V4U v) { return make_uint4(v.x, v.y, v.z, v.w); }

// Membership tests for N_L(pivot).  Ids of a list's clipped prefix lie in
// [min N_L(j), max N_L(j)]; ids outside the prefix (partial 16-byte loads)
// are replaced by sentinel(), which every set reports absent.
struct BitmapSet {  // bit (k - base) of an LDS window; base is a multiple of 128
    static constexpr bool kBitmap = true;
    static constexpr bool kFull16 = false;
    const uint32_t* words;  // LDS window minus base / 32 words: words[k >> 5] is k's word
    uint32_t last;          // base + window bits - 1: never set
    __device__ __forceinline__ uint32_t sentinel() const { return last; }
    __device__ __forceinline__ uint32_t has(uint32_t k) const {
        // v_bfe_u32 reads the bit offset from k's low 5 bits
        return __builtin_amdgcn_ubfe(words[k >> 5], k, 1);
    }
};

// An LDS window that starts at id 0 and spans every 16-bit id, for pivots
// whose ids are all below kNarrowIds: any id read from the 16-bit copy of L
// indexes inside it, so ids outside a prefix need no clamping and are
// dropped afterwards by a mask.
struct FullWindow16 {
    static constexpr bool kBitmap = true;
    static constexpr bool kFull16 = true;
    const uint32_t* words;
    // bits of the four 16-bit ids packed in h (bit i = id i); v_bfe_u32
    // reads the bit offset from the low 5 bits of its operand, so the low
    // id of a dword needs no unpacking
    __device__ __forceinline__ uint32_t bits4(uint2 h) const {
        const uint32_t b0 = __builtin_amdgcn_ubfe(words[__builtin_amdgcn_ubfe(h.x, 5, 11)], h.x, 1);
        const uint32_t b1 = __builtin_amdgcn_ubfe(words[h.x >> 21], h.x >> 16, 1);
        const uint32_t b2 = __builtin_amdgcn_ubfe(words[__builtin_amdgcn_ubfe(h.y, 5, 11)], h.y, 1);
        const uint32_t b3 = __builtin_amdgcn_ubfe(words[h.y >> 21], h.y >> 16, 1);
        return b0 | (b1 << 1) | (b2 << 2) | (b3 << 3);
    }
};

struct HashSet {  // open addressing, linear probing, load <= 1/2
    static constexpr bool kBitmap = false;
    static constexpr bool kFull16 = false;
    const uint32_t* slots;
    int shift;
    uint32_t mask;
    __device__ __forceinline__ uint32_t sentinel() const { return kEmpty; }
    __device__ __forceinline__ uint32_t has(uint32_t k) const {
        uint32_t h = hslot(k, shift);
        for (;;) {
            const uint32_t v = slots[h];
            if (v == kEmpty) return 0u;  // before the match test: the sentinel is kEmpty
            if (v == k) return 1u;
            h = (h + 1u) & mask;
        }
    }
};

struct SearchSet {  // binary search of the sorted pivot row in HBM
    static constexpr bool kBitmap = false;
    static constexpr bool kFull16 = false;
    const int32_t* col;
    uint32_t t0, t1;
    __device__ __forceinline__ uint32_t sentinel() const { return kEmpty; }  // above every id
    __device__ __forceinline__ uint32_t has(uint32_t k) const {
        uint32_t lo = t0, hi = t1;
        while (lo < hi) {
            const uint32_t mid = (lo + hi) >> 1;
            if ((uint32_t)col[mid] < k) lo = mid + 1;
            else hi = mid;
        }
        return (lo < t1 && (uint32_t)col[lo] == k) ? 1u : 0u;
    }
};

// Device view of one plan's work lists.
struct PlanDev {
    const int32_t* col;
    const uint16_t* col16;     // col's low 16 bits, read for pivots whose ids are below kNarrowIds
    const uint4* lists;        // per list {s0, s_noncore, s1, row bitmap or kNoBitmap}
    const uint4* rowbits;      // core4 uint4 per row bitmap
    const int32_t* ur;
    const uint4* items;  // 2 x uint4 per item: {t0, t1, eb, ee}, {lo, hi, work, graph (batches) or 0}
    uint32_t n_block;    // items [0, n_block) are workgroup items, the rest wave items
    uint32_t n_items;
    unsigned long long* result;
    PvOut pv;
    uint32_t core4;  // hub-core ids / 128 (uint4 words of a row bitmap)
    // batches: per-graph sums that the GS kernels add each item's count to,
    // at the graph index its record holds (B.w)
    unsigned long long* gsum = nullptr;
};

// One launch counts one plan (nplans == 1: p0) or several (plans[], with map
// giving each global item its plan and local item: workgroup items of all
// plans first, then wave items).
struct WedgeArgs {
    PlanDev p0;
    const PlanDev* plans;
    const uint2* map;
    uint32_t nplans;
    uint32_t n_block;    // global workgroup items
    uint32_t n_items;    // global items
    uint32_t* counters;  // [0] workgroup-item cursor, [1] wave-item cursor (schedule 0)
    int lookup;          // 0 = auto, 1 = never bitmap, 2 = never hash
    int schedule;        // 0 = work stealing, 1 = static stride
    uint32_t share = 0, shares = 1;  // single plan: launch item t is plan item share + t * shares of its class
    unsigned long long* slots = nullptr;  // single plan: workgroup sums go to these kSlots slots, not p0.result
    const uint32_t* dcounts = nullptr;    // single plan, set: {n_block, n_items} read from device memory
};

// Barrier of the threads that share one LDS region: the workgroup (NT = 256)
// or one wave (NT = 64).
template <int NT>
__device__ __forceinline__ void team_sync() {
    if constexpr (NT == kThreads) {
        __syncthreads();
    } else {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
}

// Counts the lists e in [eb, ee) of one pivot against `set`.  A group of G
// lanes owns lists eb + g, eb + g + NT/G, ...  A list whose row has a core
// bitmap (and whose pivot is in an LDS bitmap) first ANDs the row's core
// bitmap with the pivot's window word by word (popcount) and then looks up
// only the prefix ids >= C; any other list looks up its whole clipped
// prefix.  Prefix ids are read with aligned loads, U in flight per lane:
// 16 bytes of 32-bit ids, or, when N16 (every id of the pivot below
// kNarrowIds, so every prefix id too), 8 bytes of the 16-bit copy, or, in a
// FullWindow16, 16 bytes (8 ids) of it; ids outside the prefix become the
// set's sentinel, or, in a FullWindow16, are masked out of the eight bits a
// lane reads.  The next
// list's record is fetched before the current list is scanned; records are
// read once per launch, so they are loaded non-temporally.  L is the
// group's first record, loaded by the caller.  For bitmap sets, wabs[w] is
// the window's word w of ids (absolute), and the core AND runs over the
// uint4 words [c4, we4).
// Byte offset of position q in an array of (1 << sh)-byte ids: 32-bit
// arithmetic (the loads take a uniform base plus a 32-bit offset) unless
// W64, which the plans of graphs with 2^30 or more ids in L take.
template <bool W64>
__device__ __forceinline__ auto byte_off(uint32_t q, int sh) {
    if constexpr (W64) return (uint64_t)q << sh;
    else return q << sh;
}

template <int NT, int G, int U, bool PV, bool W64, bool N16, class Set>
__device__ __forceinline__ uint64_t scan_lists(const Set& set, const PlanDev& a, uint32_t tid, uint32_t eb,
                                               uint32_t ee, uint4 L, const uint32_t* wabs, uint32_t c4,
                                               uint32_t we4) {
    static_assert(N16 || !Set::kFull16, "a full 16-bit window reads 16-bit ids");
    constexpr uint32_t NG = NT / G;
    constexpr uint32_t STEP = 4 * G * U;
    const uint32_t g = tid / G, gl = tid % G;
    // 32-bit byte offsets unless W64 (positions are u32)
    const GlobalChar colb = N16 ? (GlobalChar)a.col16 : (GlobalChar)a.col;
    uint64_t acc = 0;
    for (uint32_t e = eb + g; e < ee; e += NG) {
        const uint32_t en = e + NG;
        uint4 nL = make_uint4(0, 0, 0, 0);
        if (en < ee) nL = load_nt(a.lists + en);
        uint32_t hits = 0;
        uint32_t s0 = L.x;
        const uint32_t s1 = L.z;
        if (Set::kBitmap && L.w != kNoBitmap) {
            // core ids: row bitmap AND pivot window, 128 bits per lane step,
            // over the uint4 words [c4, we4): the pivot has no ids below
            // them or above them
            const GlobalU4 R = (GlobalU4)a.rowbits + (size_t)L.w * a.core4;
            for (uint32_t w4 = c4 + gl; w4 < we4; w4 += G) {
                const uint4 r = to_u4(R[w4]);
                const uint4 pw = *reinterpret_cast<const uint4*>(wabs + 4 * w4);
                hits += __popc(r.x & pw.x) + __popc(r.y & pw.y) + __popc(r.z & pw.z) + __popc(r.w & pw.w);
            }
            s0 = L.y;
        }
        if constexpr (Set::kFull16) {
            // 16 bytes (8 ids) per lane per load, U loads in flight
            for (uint32_t q = (s0 & ~7u) + 8 * gl; q < s1; q += 2 * STEP) {
                uint4 h[U];
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const uint32_t qq = q + 8 * G * u;
                    h[u] = make_uint4(0, 0, 0, 0);
                    if (qq < s1) h[u] = to_u4(*(GlobalU4)(colb + byte_off<W64>(qq, 1)));
                }
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const uint32_t qq = q + 8 * G * u;
                    if (qq < s1) {
                        const uint32_t b = set.bits4(make_uint2(h[u].x, h[u].y)) |
                                           (set.bits4(make_uint2(h[u].z, h[u].w)) << 4);
                        // the prefix starts d0 (0..7) ids into this load and
                        // ends d1 (1..8) ids into it (empty when d1 <= d0)
                        const uint32_t d0 = s0 > qq ? s0 - qq : 0u;
                        const uint32_t d1 = min(s1 - qq, 8u);
                        hits += __popc(b & (0xFFu << d0) & ((1u << d1) - 1u));
                    }
                }
            }
        } else {
            const uint32_t sent = set.sentinel();
            const uint32_t len = s1 > s0 ? s1 - s0 : 0u;
            for (uint32_t q = (s0 & ~3u) + 4 * gl; q < s1; q += STEP) {
                uint4 v[U];
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const uint32_t qq = q + 4 * G * u;
                    if constexpr (N16) {
                        uint2 h = make_uint2(0u, 0u);  // ids at or past s1 are masked below
                        if (qq < s1) h = to_u2(*(GlobalU2)(colb + byte_off<W64>(qq, 1)));
                        v[u] = make_uint4(h.x & 0xFFFFu, h.x >> 16, h.y & 0xFFFFu, h.y >> 16);
                    } else {
                        v[u] = make_uint4(sent, sent, sent, sent);
                        if (qq < s1) v[u] = to_u4(*(GlobalU4)(colb + byte_off<W64>(qq, 2)));
                    }
                }
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const uint32_t o = q + 4 * G * u - s0;  // prefix offset of v[u].x (wraps below s0)
                    hits += set.has(o + 0 < len ? v[u].x : sent);
                    hits += set.has(o + 1 < len ? v[u].y : sent);
                    hits += set.has(o + 2 < len ? v[u].z : sent);
                    hits += set.has(o + 3 < len ? v[u].w : sent);
                }
            }
        }
        acc += hits;
        if (PV) {
            uint32_t c = hits;
#pragma unroll
            for (int o = G / 2; o > 0; o >>= 1) c += __shfl_xor(c, o, G);
            if (gl == 0 && c) pv_add(a.pv, a.ur[e], c);
        }
        L = nL;
    }
    return acc;
}

// One work item (pivot row [t0, t1) of col, lists [eb, ee)) processed by NT
// threads that share W words of zeroed LDS.  N_L(j) goes into LDS as a
// bitmap window when max - min < 32 W (from id 0 for a workgroup whose
// pivot's ids are all below kNarrowIds), else into a hash table when
// d_L(j) <= hash_capacity(W), else stays in HBM for binary search.  The LDS
// words are zero again on return.
template <int NT, int W, int G, int U, bool PV, bool W64>
__device__ __forceinline__ uint64_t process_item(const PlanDev& a, int lookup, uint4 A, uint4 B, uint32_t* lds,
                                                 uint32_t tid) {
    const uint32_t t0 = A.x, t1 = A.y, eb = A.z, ee = A.w, lo = B.x, hi = B.y;
    const uint32_t d = t1 - t0;  // >= 1 for every listed pivot
    // this thread's first list record, fetched while the set is built
    uint4 L = make_uint4(0, 0, 0, 0);
    {
        const uint32_t e = eb + tid / G;
        if (e < ee) L = load_nt(a.lists + e);
    }
    uint64_t acc = 0;
    const bool narrow = hi < kNarrowIds;
    // a workgroup's window holds 32 W >= 2 kNarrowIds ids: from id 0 it
    // spans every 16-bit id
    static_assert(NT != kThreads || 32u * W >= 2u * kNarrowIds, "workgroup window below 2^17 ids");
    const bool full16 = NT == kThreads && narrow;
    const uint32_t base = full16 ? 0u : lo & ~127u;  // 128-aligned: window words line up with core bitmap uint4s
    if (hi - base < 32u * W - 1u && lookup != 1) {  // the window's last bit stays 0 (sentinel)
        const uint32_t bw = base >> 5;
        for (uint32_t p = t0 + tid; p < t1; p += NT) {
            const uint32_t k = (uint32_t)((GlobalI32)a.col)[p];
            atomicOr(&lds[(k >> 5) - bw], 1u << (k & 31u));
        }
        team_sync<NT>();
        // core AND over the uint4 words from the pivot's first 128-id
        // block up to the one holding its last id (or the core's end)
        const uint32_t c4 = lo >> 7, we4 = min(a.core4, (hi >> 7) + 1u);
        if constexpr (NT == kThreads) {
            if (full16) {
                acc = scan_lists<NT, G, U, PV, W64, true>(FullWindow16{lds}, a, tid, eb, ee, L, lds, c4, we4);
            } else {
                acc = scan_lists<NT, G, U, PV, W64, false>(BitmapSet{lds - bw, base + 32u * W - 1u}, a, tid, eb, ee, L,
                                                      lds - bw, c4, we4);
            }
        } else {
            const BitmapSet set{lds - bw, base + 32u * W - 1u};
            acc = narrow ? scan_lists<NT, G, U, PV, W64, true>(set, a, tid, eb, ee, L, lds - bw, c4, we4)
                         : scan_lists<NT, G, U, PV, W64, false>(set, a, tid, eb, ee, L, lds - bw, c4, we4);
        }
        team_sync<NT>();
        // zero the words that can hold bits, [0, (hi - base) / 32], with
        // 16-byte stores rather than reading the pivot row again (W is a
        // multiple of 4, so every store stays inside the window)
        static_assert(W % 4 == 0, "window of whole uint4 words");
        const uint32_t nw = ((hi - base) >> 5) + 1u;
        for (uint32_t w = 4 * tid; w < nw; w += 4 * NT) *reinterpret_cast<uint4*>(lds + w) = make_uint4(0, 0, 0, 0);
    } else if (d <= hash_capacity(W) && lookup != 2) {
        constexpr int kMaxLog = hash_log_max(W);
        const int logS = min(kMaxLog, max(6, 32 - __clz(4 * d - 1)));  // S >= 4d while it fits
        const int shift = 32 - logS;
        const uint32_t mask = (1u << logS) - 1u;
        for (uint32_t x = tid; x <= mask; x += NT) lds[x] = kEmpty;
        team_sync<NT>();
        for (uint32_t p = t0 + tid; p < t1; p += NT) {
            const uint32_t key = (uint32_t)((GlobalI32)a.col)[p];
            uint32_t h = hslot(key, shift);
            while (atomicCAS(&lds[h], kEmpty, key) != kEmpty) h = (h + 1u) & mask;
        }
        team_sync<NT>();
        const HashSet set{lds, shift, mask};
        acc = narrow ? scan_lists<NT, G, U, PV, W64, true>(set, a, tid, eb, ee, L, nullptr, 0u, 0u)
                     : scan_lists<NT, G, U, PV, W64, false>(set, a, tid, eb, ee, L, nullptr, 0u, 0u);
        team_sync<NT>();
        for (uint32_t x = tid; x <= mask; x += NT) lds[x] = 0;
    } else {
        const SearchSet set{a.col, t0, t1};
        acc = narrow ? scan_lists<NT, G, U, PV, W64, true>(set, a, tid, eb, ee, L, nullptr, 0u, 0u)
                     : scan_lists<NT, G, U, PV, W64, false>(set, a, tid, eb, ee, L, nullptr, 0u, 0u);
    }
    return acc;
}

__device__ __forceinline__ uint4 uniform4(uint4 v) {
    return make_uint4(uniform(v.x), uniform(v.y), uniform(v.z), uniform(v.w));
}

__device__ __forceinline__ PlanDev uniform_plan(const PlanDev* p) {
    PlanDev q = *p;
    q.n_block = uniform(q.n_block);
    q.n_items = uniform(q.n_items);
    return q;
}

// Item t of the launch: its plan's index and its two descriptor words
// (prefetched one item ahead; the plan descriptor itself is read when the
// item runs, which keeps the prefetch to a few registers).  Wave items are
// numbered after all workgroup items.
struct Item {
    uint32_t plan;
    uint4 A, B;
};

// nb = the launch's workgroup items, pnb = the plan's.
__device__ __forceinline__ Item fetch_item(const WedgeArgs& a, uint32_t t, uint32_t nb, uint32_t pnb) {
    Item r;
    const uint4* items = a.p0.items;
    uint32_t li = t;
    r.plan = 0;
    if (a.nplans == 1) {
        // single plan: every shares-th item of each class from share on
        // (the plan's own numbering when shares = 1)
        li = t < nb ? a.share + t * a.shares : pnb + a.share + (t - nb) * a.shares;
    } else {
        const uint2 m = a.map[t];
        r.plan = uniform(m.x);
        li = uniform(m.y);
        items = a.plans[r.plan].items;
    }
    r.A = uniform4(items[2 * li]);
    r.B = uniform4(items[2 * li + 1]);
    return r;
}

__device__ __forceinline__ PlanDev item_plan(const WedgeArgs& a, const Item& it) {
    return a.nplans == 1 ? a.p0 : uniform_plan(a.plans + it.plan);
}

// Adds a team's partial counts into a plan's total (multi-plan launches
// flush after every item; a single-plan launch once at the end).
__device__ __forceinline__ void flush(uint64_t& acc, unsigned long long* result) {
    const uint64_t s = wave_sum_u64(acc);
    if ((threadIdx.x & 63) == 0 && s) atomicAdd(result, (unsigned long long)s);
    acc = 0;
}

// The workgroup's partial counts (every thread calls it) added with one
// atomic into slot blockIdx % kSlots.
__device__ __forceinline__ void flush_slot(uint64_t acc, unsigned long long* slots) {
    __shared__ unsigned long long part[kThreads / 64];
    const uint64_t s = wave_sum_u64(acc);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long x = 0;
#pragma unroll
        for (int w = 0; w < kThreads / 64; ++w) x += part[w];
        if (x) atomicAdd(slots + kSlotStride * (blockIdx.x % kSlots), x);
    }
}

// Persistent grid.  Phase 1: workgroup items (heavy pivots, sorted by work,
// 16 KiB LDS each).  Phase 2: wave items (pivots with few lists), each wave
// with its own 4 KiB slice.  schedule 1 strides statically over the sorted
// items and prefetches the next item; schedule 0 takes items from an
// atomic cursor.
// Shapes with G <= 8, U = 1, totals only (the default is G = 4) are held
// to 64 VGPRs, which keeps 8 waves per SIMD resident at the cost of a few
// spilled registers; the others are left to the compiler.
template <int G, int U, bool PV>
constexpr int kMinWaves = (G <= 8 && U == 1 && !PV) ? 8 : 1;

// Adds one item's count to its graph's sum gsum[g] (batches): v is this
// thread's part of the item, shared by NT threads (a workgroup, a wave or a
// tiny-item group).  A workgroup reduces through part[4]; the caller's next
// barrier orders part's reuse.
template <int NT>
__device__ __forceinline__ void add_item_graph(uint64_t v, unsigned long long* gsum, uint32_t g,
                                               unsigned long long* part, uint32_t tid) {
    if constexpr (NT == kThreads) {
        v = wave_sum_u64(v);
        if ((tid & 63) == 0) part[tid >> 6] = v;
        __syncthreads();
        if (tid == 0) {
            const unsigned long long x = part[0] + part[1] + part[2] + part[3];
            if (x) atomicAdd(gsum + g, x);
        }
    } else if constexpr (NT == 64) {
        v = wave_sum_u64(v);
        if ((tid & 63) == 0 && v) atomicAdd(gsum + g, (unsigned long long)v);
    } else {
#pragma unroll
        for (int o = NT / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, NT);
        if (tid % NT == 0 && v) atomicAdd(gsum + g, (unsigned long long)v);
    }
}

// GS: every item's count also goes to its graph's sum (PV is then off).
template <int G, int U, bool PV, bool W64, bool GS>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(kMinWaves<G, U, PV>))) void wedge_kernel(
    WedgeArgs a) {
    static_assert(!(GS && PV), "per-graph item sums replace per-vertex sums");
    __shared__ __align__(16) uint32_t lds[kLdsWords];  // 16-byte reads of bitmap words
    __shared__ uint32_t s_item;
    __shared__ unsigned long long gs_part[kThreads / 64];
    const uint32_t tid = threadIdx.x;
    const int lane = tid & 63;
    const uint32_t wid = uniform(tid >> 6);
    const bool multi = a.nplans != 1;
    const uint32_t n_block = a.dcounts ? uniform(a.dcounts[0]) : a.n_block;
    const uint32_t n_items = a.dcounts ? uniform(a.dcounts[1]) : a.n_items;
    const uint32_t pnb = a.dcounts ? n_block : a.p0.n_block;
    for (int i = tid; i < kLdsWords; i += kThreads) lds[i] = 0;
    __syncthreads();
    uint64_t acc = 0;

    // phase 1: workgroup items
    if (a.schedule == 1) {
        uint32_t it = blockIdx.x;
        Item cur;
        if (it < n_block) cur = fetch_item(a, it, n_block, pnb);
        while (it < n_block) {
            const uint32_t nx = it + gridDim.x;
            Item nxt;
            if (nx < n_block) nxt = fetch_item(a, nx, n_block, pnb);
            const PlanDev P = item_plan(a, cur);
            const uint64_t v = process_item<kThreads, kLdsWords, G, U, PV, W64>(P, a.lookup, cur.A, cur.B, lds, tid);
            acc += v;
            if constexpr (GS) add_item_graph<kThreads>(v, P.gsum, cur.B.w, gs_part, tid);
            if (multi) flush(acc, P.result);
            __syncthreads();
            it = nx;
            cur = nxt;
        }
    } else {
        for (;;) {
            if (tid == 0) s_item = atomicAdd(a.counters, 1u);
            __syncthreads();
            const uint32_t it = uniform(s_item);
            if (it >= n_block) break;
            const Item cur = fetch_item(a, it, n_block, pnb);
            const PlanDev P = item_plan(a, cur);
            const uint64_t v = process_item<kThreads, kLdsWords, G, U, PV, W64>(P, a.lookup, cur.A, cur.B, lds, tid);
            acc += v;
            if constexpr (GS) add_item_graph<kThreads>(v, P.gsum, cur.B.w, gs_part, tid);
            if (multi) flush(acc, P.result);
            __syncthreads();
        }
    }

    // phase 2: wave items, LDS split into four zeroed slices
    uint32_t* slice = lds + wid * kWaveWords;
    const uint32_t nwave = n_items - n_block;
    if (a.schedule == 1) {
        const uint32_t stride = gridDim.x * (kThreads / 64);
        uint32_t it = blockIdx.x * (kThreads / 64) + wid;
        Item cur;
        if (it < nwave) cur = fetch_item(a, n_block + it, n_block, pnb);
        while (it < nwave) {
            const uint32_t nx = it + stride;
            Item nxt;
            if (nx < nwave) nxt = fetch_item(a, n_block + nx, n_block, pnb);
            const PlanDev P = item_plan(a, cur);
            const uint64_t v = process_item<64, kWaveWords, G, U, PV, W64>(P, a.lookup, cur.A, cur.B, slice,
                                                                           (uint32_t)lane);
            acc += v;
            if constexpr (GS) add_item_graph<64>(v, P.gsum, cur.B.w, nullptr, (uint32_t)lane);
            if (multi) flush(acc, P.result);
            team_sync<64>();
            it = nx;
            cur = nxt;
        }
    } else {
        for (;;) {
            uint32_t it = 0;
            if (lane == 0) it = atomicAdd(a.counters + 1, 1u);
            it = uniform(__shfl(it, 0));
            if (it >= nwave) break;
            const Item cur = fetch_item(a, n_block + it, n_block, pnb);
            const PlanDev P = item_plan(a, cur);
            const uint64_t v = process_item<64, kWaveWords, G, U, PV, W64>(P, a.lookup, cur.A, cur.B, slice,
                                                                           (uint32_t)lane);
            acc += v;
            if constexpr (GS) add_item_graph<64>(v, P.gsum, cur.B.w, nullptr, (uint32_t)lane);
            if (multi) flush(acc, P.result);
            team_sync<64>();
        }
    }
    if (!multi) {
        if (a.slots) flush_slot(acc, a.slots);
        else flush(acc, a.p0.result);
    }
}

// Tiny-pivot items [first, first + share + t * shares) of one plan (t < n):
// each group of kTinyLanes lanes takes one item with a kTinyWords-slot LDS
// hash table of its own (static stride over the work-sorted items), so a
// wave has 8 pivots' HBM round trips in flight where a wave item has one;
// within a group, kTinyListLanes lanes stream each prefix list, so a
// group's lists (a few short ones each) run 4 at a time.
// kTinyListLanes 2 against 8 (one list at a time): batched workload 6.22 ->
// 4.38 ms (1 lane: 4.39), R-MAT s22 count 2.751 -> 2.744 ms (1 lane: 2.759).
// Each of those lanes keeps kTinyLoads loads in flight (2 against 1: 4.41
// -> 4.31 ms on the batched workload).
constexpr int kTinyListLanes = 2;
constexpr int kTinyLoads = 2;
// The groups of a wave run different items; every step of process_item
// for a team of kTinyLanes lanes stays within the group.
template <bool PV, bool W64, bool GS>
__global__ __launch_bounds__(kThreads) void tiny_kernel(PlanDev P, int lookup, uint32_t first, uint32_t n,
                                                         uint32_t share, uint32_t shares,
                                                         unsigned long long* slots) {
    static_assert(!(GS && PV), "per-graph item sums replace per-vertex sums");
    constexpr uint32_t kGroups = kThreads / kTinyLanes;
    __shared__ __align__(16) uint32_t lds[kGroups * kTinyWords];
    const uint32_t tid = threadIdx.x, grp = tid / kTinyLanes, gl = tid % kTinyLanes;
    for (uint32_t i = tid; i < kGroups * kTinyWords; i += kThreads) lds[i] = 0;
    __syncthreads();
    uint32_t* table = lds + grp * kTinyWords;
    uint64_t acc = 0;
    for (uint32_t t = blockIdx.x * kGroups + grp; t < n; t += gridDim.x * kGroups) {
        const uint32_t li = first + share + t * shares;
        const uint4 A = P.items[2 * li], B = P.items[2 * li + 1];
        const uint64_t v =
            process_item<kTinyLanes, kTinyWords, kTinyListLanes, kTinyLoads, PV, W64>(P, lookup, A, B, table, gl);
        acc += v;
        if constexpr (GS) add_item_graph<kTinyLanes>(v, P.gsum, B.w, nullptr, tid);
    }
    if (slots) flush_slot(acc, slots);
    else flush(acc, P.result);
}

using WedgeFn = void (*)(WedgeArgs);

template <int G, int U>
WedgeFn pick_pv(bool pv, bool gs) {
    return pv ? wedge_kernel<G, U, true, false, false>
              : (gs ? wedge_kernel<G, U, false, false, true> : wedge_kernel<G, U, false, false, false>);
}

template <int G>
WedgeFn pick_u(int u, bool pv, bool gs) {
    switch (u) {
        case 1: return pick_pv<G, 1>(pv, gs);
        case 2: return pick_pv<G, 2>(pv, gs);
        default: return pick_pv<G, 4>(pv, gs);
    }
}

WedgeFn pick_wedge(int g, int u, bool pv, bool gs) {
    switch (g) {
        case 4: return pick_u<4>(u, pv, gs);
        case 16: return pick_u<16>(u, pv, gs);
        case 32: return pick_u<32>(u, pv, gs);
        default: return pick_u<8>(u, pv, gs);
    }
}

}  // namespace
}  // namespace btcx

namespace btcx {
namespace {

// col16[q] = low 16 bits of col[q], over L and its padding.
__global__ void narrow_col_kernel(const int32_t* __restrict__ col, Cnt n_, uint16_t* __restrict__ col16) {
    const uint32_t n = (uint32_t)n_.get();
    for (uint32_t q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x)
        col16[q] = (uint16_t)col[q];
}

void build_col16(btc_plan_s* p, hipStream_t s) {
    const uint32_t n = (uint32_t)p->g->m + (uint32_t)kColPad;
    p->col16.alloc(n);
    const unsigned grid = std::min<unsigned>((n + kThreads - 1) / kThreads, 1u << 16);
    hipLaunchKernelGGL(narrow_col_kernel, dim3(grid), dim3(kThreads), 0, s, p->g->d_col.p, n, p->col16.p);
    BTC_CHECK_LAUNCH("narrow_col_kernel");
}

// Marks the L entries that fall into dense tiles and uploads the dense tile
// list.  Leaves p->bsr empty when no tile is dense.
// Graphs up to this many edges are classified on the host (tile counts from
// the CSR), so a graph without dense tiles never builds a Bit-BSR.
constexpr int64_t kHostClassifyEdges = 8 << 20;

// Host classification: tiles of each row block enumerated in (I, J) order --
// the order of the GPU builder -- so tile t here is tile t of the Bit-BSR.
void classify_dense_host(btc_plan_s* p, hipStream_t s, std::vector<uint8_t>& dense) {
    btc_graph_s* g = p->g;
    const int K = p->K;
    const int32_t n_rb = (g->n + 15) / 16;
    std::vector<uint32_t> ntiles((size_t)n_rb + 1, 0);
    // pass 1: tiles per row block; pass 2: dense tiles and flags
    auto row_block = [&](int32_t I, bool mark, uint32_t t0, std::vector<uint32_t>* list, int64_t* dedges) {
        thread_local std::vector<uint32_t> cb;
        cb.clear();
        const int32_t r0 = I * 16, r1 = std::min(g->n, r0 + 16);
        for (int64_t q = g->rowptr[r0]; q < g->rowptr[r1]; ++q) cb.push_back((uint32_t)g->col[q] / K);
        std::sort(cb.begin(), cb.end());
        uint32_t tiles = 0;
        for (size_t a = 0; a < cb.size();) {
            size_t b = a;
            while (b < cb.size() && cb[b] == cb[a]) ++b;
            if (mark && (int)(b - a) > p->tau) {
                list->push_back(t0 + tiles);
                *dedges += (int64_t)(b - a);
                for (int32_t r = r0; r < r1; ++r)
                    for (int64_t q = g->rowptr[r]; q < g->rowptr[r + 1]; ++q)
                        if ((uint32_t)g->col[q] / K == cb[a]) dense[q] = 1;
            }
            ++tiles;
            a = b;
        }
        return tiles;
    };
#pragma omp parallel for schedule(dynamic, 256)
    for (int32_t I = 0; I < n_rb; ++I) ntiles[I + 1] = row_block(I, false, 0, nullptr, nullptr);
    for (int32_t I = 0; I < n_rb; ++I) ntiles[I + 1] += ntiles[I];
    std::vector<uint32_t> list;
    int64_t dense_edges = 0;
    for (int32_t I = 0; I < n_rb; ++I) row_block(I, true, ntiles[I], &list, &dense_edges);
    p->n_dense = (uint32_t)list.size();
    p->info.dense_tiles = p->n_dense;
    p->info.dense_edges = dense_edges;
    if (list.empty()) return;
    plan_bsr(p, s);
    BTC_REQUIRE(p->bsr->nblk == ntiles[n_rb], BTC_ERR_STATE, "tile count mismatch (%u vs %u)", p->bsr->nblk,
                ntiles[n_rb]);
    p->dense_list.alloc(list.size());
    BTC_HIP(hipMemcpyAsync(p->dense_list.p, list.data(), list.size() * 4, hipMemcpyHostToDevice, s));
    BTC_HIP(hipStreamSynchronize(s));
}

void classify_dense(btc_plan_s* p, hipStream_t s, std::vector<uint8_t>& dense) {
    btc_graph_s* g = p->g;
    dense.assign((size_t)g->m, 0);
    if (p->tau >= 16 * p->K || g->m == 0) return;
    if (g->m <= kHostClassifyEdges) {
        classify_dense_host(p, s, dense);
        return;
    }
    const btc_bitbsr_s* b = plan_bsr(p, s);
    DevBuf<uint32_t> d_pop;
    d_pop.alloc(b->nblk);
    launch_tile_popcount(b, d_pop.p, s);
    std::vector<uint32_t> pop(b->nblk), indptr((size_t)b->n_rb + 1), indices(b->nblk);
    BTC_HIP(hipMemcpyAsync(pop.data(), d_pop.p, (size_t)b->nblk * 4, hipMemcpyDeviceToHost, s));
    BTC_HIP(hipMemcpyAsync(indptr.data(), b->indptr.p, indptr.size() * 4, hipMemcpyDeviceToHost, s));
    BTC_HIP(hipMemcpyAsync(indices.data(), b->indices.p, (size_t)b->nblk * 4, hipMemcpyDeviceToHost, s));
    BTC_HIP(hipStreamSynchronize(s));
    std::vector<uint32_t> list;
    int64_t dense_edges = 0;
    for (uint32_t t = 0; t < b->nblk; ++t)
        if ((int)pop[t] > p->tau) {
            list.push_back(t);
            dense_edges += pop[t];
        }
    p->n_dense = (uint32_t)list.size();
    p->info.dense_tiles = p->n_dense;
    p->info.dense_edges = dense_edges;
    if (list.empty()) {
        drop_plan_bsr(p);
        return;
    }
    const int K = p->K;
#pragma omp parallel for schedule(dynamic, 1024)
    for (int32_t i = 0; i < g->n; ++i) {
        const uint32_t I = (uint32_t)i / 16;
        const uint32_t* b0 = indices.data() + indptr[I];
        const uint32_t* b1 = indices.data() + indptr[I + 1];
        for (int64_t q = g->rowptr[i]; q < g->rowptr[i + 1]; ++q) {
            const uint32_t jc = (uint32_t)g->col[q] / K;
            const uint32_t* f = std::lower_bound(b0, b1, jc);
            dense[q] = (int)pop[f - indices.data()] > p->tau;
        }
    }
    p->dense_list.alloc(list.size());
    BTC_HIP(hipMemcpyAsync(p->dense_list.p, list.data(), list.size() * 4, hipMemcpyHostToDevice, s));
    BTC_HIP(hipStreamSynchronize(s));
}

// Groups the sparse edges' clipped prefixes by pivot and cuts them into work
// items of about `chunk` wedge checks.
void build_wedge_lists(btc_plan_s* p, const std::vector<uint8_t>& dense, int64_t chunk, hipStream_t s) {
    btc_graph_s* g = p->g;
    const int32_t n = g->n;
    const int64_t* rp = g->rowptr.data();
    const int32_t* cl = g->col.data();
    std::vector<std::atomic<uint32_t>> cnt((size_t)n);
#pragma omp parallel for schedule(static)
    for (int32_t j = 0; j < n; ++j) cnt[j].store(0, std::memory_order_relaxed);

    // clipped prefix of entry q (row i, pivot j = col[q]); empty when nothing can match
    auto clip = [&](int32_t i, int64_t q, uint32_t& s0, uint32_t& s1) -> bool {
        const int32_t j = cl[q];
        if (rp[j + 1] == rp[j] || q == rp[i]) return false;
        const int32_t lo = cl[rp[j]], hi = cl[rp[j + 1] - 1];
        const int32_t* a = std::lower_bound(cl + rp[i], cl + q, lo);
        const int32_t* b = std::upper_bound(a, cl + q, hi);
        if (b <= a) return false;
        s0 = (uint32_t)(a - cl);
        s1 = (uint32_t)(b - cl);
        return true;
    };
    int64_t sparse_edges = 0, sparse_work = 0;
#pragma omp parallel for schedule(dynamic, 1024) reduction(+ : sparse_edges, sparse_work)
    for (int32_t i = 0; i < n; ++i)
        for (int64_t q = rp[i]; q < rp[i + 1]; ++q) {
            if (dense[q]) continue;
            ++sparse_edges;
            sparse_work += (rp[i + 1] - rp[i]) + (rp[cl[q] + 1] - rp[cl[q]]);
            uint32_t s0, s1;
            if (clip(i, q, s0, s1)) cnt[cl[q]].fetch_add(1, std::memory_order_relaxed);
        }
    std::vector<int64_t> off((size_t)n + 1, 0);
    for (int32_t j = 0; j < n; ++j) off[j + 1] = off[j] + cnt[j].load(std::memory_order_relaxed);
    const int64_t ne = off[n];
    BTC_REQUIRE(ne < (int64_t)0xFFFFFFFFLL, BTC_ERR_OVERFLOW, "too many wedge lists");
#pragma omp parallel for schedule(static)
    for (int32_t j = 0; j < n; ++j) cnt[j].store((uint32_t)off[j], std::memory_order_relaxed);
    // core bitmaps: rows with at least kCoreRowMin ids below C
    const uint32_t C = p->core_ids, CW = C / 32;
    std::vector<uint32_t> cend((size_t)n), rbi((size_t)n, kNoBitmap);
#pragma omp parallel for schedule(dynamic, 4096)
    for (int32_t i = 0; i < n; ++i)
        cend[i] = (uint32_t)(std::lower_bound(cl + rp[i], cl + rp[i + 1], (int32_t)C) - cl);
    uint32_t nbm = 0;
    for (int32_t i = 0; i < n; ++i)
        if (cend[i] - rp[i] >= kCoreRowMin) rbi[i] = nbm++;
    std::vector<uint32_t> rowbits((size_t)nbm * CW, 0u);
#pragma omp parallel for schedule(dynamic, 1024)
    for (int32_t i = 0; i < n; ++i) {
        if (rbi[i] == kNoBitmap) continue;
        uint32_t* b = rowbits.data() + (size_t)rbi[i] * CW;
        for (int64_t q = rp[i]; q < cend[i]; ++q) b[cl[q] >> 5] |= 1u << (cl[q] & 31);
    }
    std::vector<uint4> lst((size_t)ne);
    std::vector<int32_t> ur((size_t)ne);
    std::atomic<int64_t> core_lists{0};
#pragma omp parallel for schedule(dynamic, 1024)
    for (int32_t i = 0; i < n; ++i)
        for (int64_t q = rp[i]; q < rp[i + 1]; ++q) {
            if (dense[q]) continue;
            uint32_t s0, s1;
            if (!clip(i, q, s0, s1)) continue;
            // the core AND costs core_words(pivot) words against the
            // lookups of the core ids
            const int32_t j = cl[q];
            const uint32_t base = (uint32_t)cl[rp[j]] & ~127u;
            uint32_t sc = s0, rb = kNoBitmap;
            if (rbi[i] != kNoBitmap && base < C) {
                const uint32_t core_ids = std::min(s1, cend[i]) > s0 ? std::min(s1, cend[i]) - s0 : 0u;
                const uint32_t words = core_words((uint32_t)cl[rp[j]], (uint32_t)cl[rp[j + 1] - 1], C);
                if (take_core_bitmap(core_ids, words)) {
                    sc = std::max(s0, cend[i]);
                    rb = rbi[i];
                    core_lists.fetch_add(1, std::memory_order_relaxed);
                }
            }
            const uint32_t e = cnt[j].fetch_add(1, std::memory_order_relaxed);
            lst[e] = make_uint4(s0, sc, s1, rb);
            ur[e] = i;
        }
    // work items: consecutive lists of one pivot, about `chunk` checks each.
    // Pivots with few lists become wave items (tiny items when tiny_pivot);
    // the rest workgroup items.
    struct Item {
        uint4 a, b;
    };
    std::vector<Item> block_items, wave_items, tiny_items;
    int64_t checks = 0, pivot_words = 0, narrow_ids = 0;
    for (int32_t j = 0; j < n; ++j) {
        int64_t e = off[j], e1 = off[j + 1];
        if (e == e1) continue;
        const uint32_t t0 = (uint32_t)rp[j], t1 = (uint32_t)rp[j + 1];
        const uint32_t lo = (uint32_t)cl[rp[j]], hi = (uint32_t)cl[rp[j + 1] - 1];
        // cost of a list: ids looked up, plus core words ANDed
        auto cost = [&](int64_t f) -> int64_t {
            const uint4& r = lst[f];
            return (int64_t)(r.z > r.y ? r.z - r.y : 0) + (r.w != kNoBitmap ? (int64_t)core_words(lo, hi, C) : 0);
        };
        const uint32_t nl = (uint32_t)std::min<int64_t>(e1 - e, 0xFFFFFFFFLL);
        const bool small = wave_pivot(nl, lo, hi, t1 - t0);
        const bool tiny = p->use_tiny && small && tiny_pivot(nl, lo, hi, t1 - t0, p->tiny_all);
        while (e < e1) {
            int64_t w = 0, f = e;
            while (f < e1 && (small || w == 0 || w + cost(f) <= chunk)) {
                w += cost(f);
                ++f;
            }
            checks += w;
            pivot_words += t1 - t0;
            if (hi < kNarrowIds)
                for (int64_t x = e; x < f; ++x) narrow_ids += lst[x].z > lst[x].y ? lst[x].z - lst[x].y : 0;
            Item it{make_uint4(t0, t1, (uint32_t)e, (uint32_t)f),
                    make_uint4(lo, hi, (uint32_t)std::min<int64_t>(w, 0xFFFFFFFFLL), 0u)};
            (tiny ? tiny_items : small ? wave_items : block_items).push_back(it);
            e = f;
        }
    }
    auto by_work = [](const Item& x, const Item& y) { return x.b.z > y.b.z; };
    std::stable_sort(block_items.begin(), block_items.end(), by_work);
    std::stable_sort(wave_items.begin(), wave_items.end(), by_work);
    std::stable_sort(tiny_items.begin(), tiny_items.end(), by_work);
    std::vector<Item> items(block_items);
    items.insert(items.end(), wave_items.begin(), wave_items.end());
    items.insert(items.end(), tiny_items.begin(), tiny_items.end());
    p->nitems = (uint32_t)items.size();
    p->n_block_items = (uint32_t)block_items.size();
    p->n_tiny_items = (uint32_t)tiny_items.size();
    p->info.tiny_items = p->n_tiny_items;
    p->info.sparse_edges = sparse_edges;
    p->info.sparse_merge_work = sparse_work;
    p->info.wedge_checks = checks;
    p->info.core_lists = core_lists.load();
    p->info.core_rows = nbm;
    // looked-up ids (4 B, or 2 B below kNarrowIds), core words and pivot
    // rows (4 B), list records (16 B), items (32 B)
    p->info.stream_bytes = 4 * (checks + pivot_words) - 2 * narrow_ids + 16 * ne + 32 * (int64_t)items.size();
    p->info.work_items = (int64_t)items.size();
    // one device allocation: items, list records, row bitmaps (16-byte
    // aligned), then list rows
    const size_t ib = items.size() * sizeof(Item), lb = (size_t)ne * sizeof(uint4),
                 bb = rowbits.size() * 4, rb = (size_t)ne * 4;
    std::vector<unsigned char> host(ib + lb + bb + rb);
    std::memcpy(host.data(), items.data(), ib);
    std::memcpy(host.data() + ib, lst.data(), lb);
    std::memcpy(host.data() + ib + lb, rowbits.data(), bb);
    std::memcpy(host.data() + ib + lb + bb, ur.data(), rb);
    p->lists.alloc(host.size());
    if (!host.empty())
        BTC_HIP(hipMemcpyAsync(p->lists.p, host.data(), host.size(), hipMemcpyHostToDevice, s));
    p->items_p = reinterpret_cast<const uint4*>(p->lists.p);
    p->list_p = reinterpret_cast<const uint4*>(p->lists.p + ib);
    p->rowbits_p = reinterpret_cast<const uint4*>(p->lists.p + ib + lb);
    p->urow_p = reinterpret_cast<const int32_t*>(p->lists.p + ib + lb + bb);
    BTC_HIP(hipStreamSynchronize(s));
}

}  // namespace

// Dense-tile form of a count: dense_kernel 2 = i8 MFMA, else (0 = auto,
// 1) 64-bit AND + popcount, which the dense-unit kernel runs faster at every
// tile width (planted 2^17, rocprofv3: 23 / 32 / 40 us against 27 / 35 / 48 us
// with the MFMA form at K = 32 / 64 / 128).
inline bool dense_mfma(const btc_count_opts& o) { return o.dense_kernel == 2; }

void check_launch_opts(const btc_count_opts& o) {
    BTC_REQUIRE(o.lookup >= 0 && o.lookup <= 2, BTC_ERR_ARG, "lookup %d (expected 0, 1 or 2)", o.lookup);
    BTC_REQUIRE(o.group_lanes == 0 || o.group_lanes == 4 || o.group_lanes == 8 || o.group_lanes == 16 ||
                    o.group_lanes == 32,
                BTC_ERR_ARG, "group_lanes %d (expected 0, 4, 8, 16 or 32)", o.group_lanes);
    BTC_REQUIRE(o.unroll == 0 || o.unroll == 1 || o.unroll == 2 || o.unroll == 4, BTC_ERR_ARG,
                "unroll %d (expected 0, 1, 2 or 4)", o.unroll);
    BTC_REQUIRE(o.schedule == 0 || o.schedule == 1, BTC_ERR_ARG, "schedule %d (expected 0 or 1)", o.schedule);
    BTC_REQUIRE(o.waves_per_eu >= 0, BTC_ERR_ARG, "waves_per_eu %d < 0", o.waves_per_eu);
    BTC_REQUIRE(o.dense_kernel >= 0 && o.dense_kernel <= 2, BTC_ERR_ARG, "dense_kernel %d (expected 0, 1 or 2)",
                o.dense_kernel);
    BTC_REQUIRE(o.planner >= 0 && o.planner <= 2, BTC_ERR_ARG, "planner %d (expected 0, 1 or 2)", o.planner);
    BTC_REQUIRE(o.wide_offsets == 0 || o.wide_offsets == 1, BTC_ERR_ARG, "wide_offsets %d (expected 0 or 1)",
                o.wide_offsets);
}

namespace {

// W64: 64-bit byte offsets into L, in the default shape only (4 lanes per
// list, one load in flight).
WedgeFn wedge_fn(const btc_count_opts& o, bool pv, bool w64, unsigned* grid, bool gs = false) {
    const int G = o.group_lanes ? o.group_lanes : 4;
    const int U = o.unroll ? o.unroll : 1;
    WedgeFn fn = w64 ? (pv ? wedge_kernel<4, 1, true, true, false> : wedge_kernel<4, 1, false, true, false>)
                     : pick_wedge(G, U, pv, gs);
    int per_cu = 0;
    BTC_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(fn), kThreads, 0));
    per_cu = std::max(1, per_cu);
    if (o.waves_per_eu > 0) per_cu = std::min(per_cu, (int)o.waves_per_eu);
    *grid = (unsigned)(device_cu_count() * per_cu);
    return fn;
}

// Whether counts of p address L with 64-bit offsets: L holds 2^30 ids or
// more (32-bit byte offsets of 4-byte ids stop there), or the plan asks.
bool wide_offsets(const btc_plan_s* p) {
    return p->launch.wide_offsets != 0 || (uint64_t)p->g->m + kColPad >= (uint64_t(1) << 30);
}

PlanDev plan_dev(btc_plan_s* p, unsigned long long* d_total, PvOut d_pv) {
    return PlanDev{p->g->d_col.p, p->col16.p, p->list_p, p->rowbits_p, p->urow_p, p->items_p, p->n_block_items, p->nitems,
                   d_total, d_pv, p->core_ids / 128, p->item_graph_sums};
}

// Whether the wedge and tiny kernels add each item's count to its graph's
// sum (a batch's plan; not with 64-bit offsets, which keep per-vertex sums).
bool graph_item_sums(const btc_plan_s* p) { return p->item_graph_sums && !wide_offsets(p); }

// Items of one class (n of them) that fall to `share` of `shares`.
uint32_t share_count(uint32_t n, uint32_t share, uint32_t shares) {
    return n > share ? (n - share + shares - 1) / shares : 0u;
}

// Workgroup and wave items of a single plan (tiny items go to launch_tiny).
void launch_wedge(btc_plan_s* p, unsigned long long* d_total, uint32_t* d_counters, PvOut d_pv, hipStream_t s,
                  uint32_t share = 0, uint32_t shares = 1, unsigned long long* d_slots = nullptr) {
    const uint32_t nb = share_count(p->n_block_items, share, shares);
    const uint32_t nw = share_count(p->nitems - p->n_tiny_items - p->n_block_items, share, shares);
    if (nb + nw == 0) return;
    unsigned grid = 0;
    const bool gs = graph_item_sums(p);
    if (gs) d_pv = PvOut{};
    WedgeFn fn = wedge_fn(p->launch, d_pv.on(), wide_offsets(p), &grid, gs);
    WedgeArgs a{plan_dev(p, d_total, d_pv), nullptr, nullptr, 1u, nb, nb + nw, d_counters, (int)p->launch.lookup,
                (int)p->launch.schedule, share, shares, d_slots};
    hipLaunchKernelGGL(fn, dim3(grid), dim3(kThreads), 0, s, a);
    BTC_CHECK_LAUNCH("wedge_kernel");
}

void launch_tiny(btc_plan_s* p, unsigned long long* d_total, PvOut d_pv, hipStream_t s, uint32_t share = 0,
                 uint32_t shares = 1, unsigned long long* d_slots = nullptr) {
    const uint32_t n = share_count(p->n_tiny_items, share, shares);
    if (!n) return;
    const bool gs = graph_item_sums(p);
    if (gs) d_pv = PvOut{};
    auto fn = wide_offsets(p) ? (d_pv.on() ? tiny_kernel<true, true, false> : tiny_kernel<false, true, false>)
                              : (d_pv.on() ? tiny_kernel<true, false, false>
                                           : (gs ? tiny_kernel<false, false, true> : tiny_kernel<false, false, false>));
    int per_cu = 0;
    BTC_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(fn), kThreads, 0));
    constexpr uint32_t kGroups = kThreads / kTinyLanes;
    const unsigned grid =
        (unsigned)std::min<int64_t>((int64_t)device_cu_count() * std::max(1, per_cu), (n + kGroups - 1) / kGroups);
    hipLaunchKernelGGL(fn, dim3(grid), dim3(kThreads), 0, s, plan_dev(p, d_total, d_pv), (int)p->launch.lookup,
                       p->nitems - p->n_tiny_items, n, share, shares, d_slots);
    BTC_CHECK_LAUNCH("tiny_kernel");
}

// The plan's dense tiles (share `share` of a `shares`-way split: units
// share, share + shares, ...) on stream s.
// Their total lands in the kSlots slots at d_slots (launch_slot_sum folds them).
void launch_plan_dense(btc_plan_s* p, unsigned long long* d_slots, PvOut d_pv, hipStream_t s, uint32_t share = 0,
                       uint32_t shares = 1) {
    launch_dense_units(p->bsr, p->dense_units.p, share, share_count(p->n_units, share, shares), shares,
                       dense_mfma(p->launch), d_slots, d_pv, s);
}

}  // namespace

namespace {

// pv[0, ctl->nc) = 0
__global__ void zero_pv_kernel(unsigned long long* __restrict__ pv, const PipeCtl* __restrict__ c) {
    const uint32_t n = c->nc;
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x) pv[v] = 0;
}

// per_graph[g] = sum of pv over [csegs[g], csegs[g + 1]), g = blockIdx < ctl->graphs
__global__ __launch_bounds__(kThreads) void segment_sums_kernel(const unsigned long long* __restrict__ pv,
                                                                const uint32_t* __restrict__ csegs,
                                                                const PipeCtl* __restrict__ c,
                                                                unsigned long long* __restrict__ per_graph) {
    __shared__ uint64_t lds[kThreads / 64];
    const uint32_t g = blockIdx.x;
    if (g >= c->graphs) return;  // workgroup-uniform
    uint64_t x = 0;
    for (uint32_t v = csegs[g] + threadIdx.x; v < csegs[g + 1]; v += kThreads) x += pv[v];
    x = block_sum_u64<kThreads>(x, lds);
    if (threadIdx.x == 0) per_graph[g] = x;
}

// A several-graph run's per-graph sums, cleared before its count adds into
// them: per_graph[g] = 0 for g < max_graphs.
__global__ void clear_graph_sums_kernel(unsigned long long* __restrict__ per_graph, uint32_t max_graphs) {
    for (uint32_t g = blockIdx.x * blockDim.x + threadIdx.x; g < max_graphs; g += gridDim.x * blockDim.x)
        per_graph[g] = 0;
}

// Each of the run's items (ctl->nitems) gets its graph in B.w: the last g <
// ctl->graphs whose rows start at or before the item's pivot row in L.
__global__ void run_item_graphs_kernel(uint4* __restrict__ items, const uint32_t* __restrict__ rowptr,
                                       const uint32_t* __restrict__ csegs, const PipeCtl* __restrict__ c) {
    const uint32_t n = c->nitems, graphs = c->graphs;
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        const uint32_t t0 = items[2 * k].x;
        uint32_t lo = 0, hi = graphs;  // rowptr[csegs[lo]] <= t0 < rowptr[csegs[hi]]
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (rowptr[csegs[mid]] <= t0) lo = mid;
            else hi = mid;
        }
        items[2 * k + 1].w = lo;
    }
}

}  // namespace

void pipeline_count_enqueue(const PlanCap& c, const uint16_t* col16, const btc_count_opts& o, hipStream_t s,
                            unsigned long long* pv, const uint32_t* csegs, unsigned long long* per_graph,
                            int max_graphs) {
    const int64_t span = c.m_cap + (int64_t)kColPad;
    const bool w64 = o.wide_offsets != 0 || (uint64_t)span >= (uint64_t(1) << 30);
    // several graphs: each item's count goes to its graph (the item records
    // get their graph first), or, with 64-bit offsets, per vertex and then
    // summed over each graph's vertex range
    const bool gs = pv && !w64;
    PvOut pvo{};
    if (gs) {
        hipLaunchKernelGGL(clear_graph_sums_kernel, dim3(1), dim3(kThreads), 0, s, per_graph, (uint32_t)max_graphs);
        BTC_CHECK_LAUNCH("clear_graph_sums_kernel");
        const unsigned g = (unsigned)std::min<int64_t>(((int64_t)c.items_cap + kThreads - 1) / kThreads, 2048);
        hipLaunchKernelGGL(run_item_graphs_kernel, dim3(g), dim3(kThreads), 0, s, c.items, c.rowptr, csegs, c.ctl);
        BTC_CHECK_LAUNCH("run_item_graphs_kernel");
    } else if (pv) {
        pvo.p64 = pv;
        const unsigned g = (unsigned)std::min<int64_t>((c.n_cap + kThreads - 1) / kThreads, 2048);
        hipLaunchKernelGGL(zero_pv_kernel, dim3(g), dim3(kThreads), 0, s, pv, c.ctl);
        BTC_CHECK_LAUNCH("zero_pv_kernel");
    }
    unsigned grid = 0;
    WedgeFn fn = wedge_fn(o, pvo.on(), w64, &grid, gs);
    unsigned long long* work = c.work;
    const PlanDev pd{c.col, col16, c.lists, reinterpret_cast<const uint4*>(c.rowbits), c.ur, c.items, 0u, 0u,
                     work, pvo, c.core_ids / 128, gs ? per_graph : nullptr};
    WedgeArgs a{pd, nullptr, nullptr, 1u, 0u, 0u, reinterpret_cast<uint32_t*>(work + 1), (int)o.lookup,
                (int)o.schedule, 0u, 1u, work + kWorkSlots, c.ctl->item_counts};
    hipLaunchKernelGGL(fn, dim3(grid), dim3(kThreads), 0, s, a);
    BTC_CHECK_LAUNCH("wedge_kernel");
    launch_slot_sum(work + kWorkSlots, work, s);
    if (pv && !gs) {
        hipLaunchKernelGGL(segment_sums_kernel, dim3((unsigned)max_graphs), dim3(kThreads), 0, s, pv, csegs, c.ctl,
                           per_graph);
        BTC_CHECK_LAUNCH("segment_sums_kernel");
    }
}

void plans_count(btc_plan_s* const* P, int k, hipStream_t s, uint64_t* totals, MultiWorkspace& ws) {
    if (k <= 0) return;
    if (ws.totals.n < (size_t)k) ws.totals.alloc((size_t)k);
    if (ws.counters.n < 2) ws.counters.alloc(2);
    BTC_HIP(hipMemsetAsync(ws.totals.p, 0, (size_t)k * 8, s));
    BTC_HIP(hipMemsetAsync(ws.counters.p, 0, 8, s));
    std::vector<PlanDev> pd((size_t)k);
    uint32_t nb = 0, ni = 0;
    for (int j = 0; j < k; ++j) {
        pd[(size_t)j] = plan_dev(P[j], ws.totals.p + j, PvOut{});
        nb += P[j]->n_block_items;
        ni += P[j]->nitems;
    }
    for (int j = 0; j < k; ++j)
        if (P[j]->n_dense) {
            launch_zero_u32(reinterpret_cast<uint32_t*>(P[j]->scratch.p + kWorkSlots), 2 * kSlotWords, s);
            launch_plan_dense(P[j], P[j]->scratch.p + kWorkSlots, PvOut{}, s);
            launch_slot_sum(P[j]->scratch.p + kWorkSlots, ws.totals.p + j, s);
        }
    if (ni) {
        unsigned grid = 0;
        bool w64 = false;
        for (int j = 0; j < k; ++j) w64 = w64 || wide_offsets(P[j]);
        WedgeFn fn = wedge_fn(P[0]->launch, false, w64, &grid);
        WedgeArgs a{pd[0], nullptr, nullptr, 1u, nb, ni, ws.counters.p, (int)P[0]->launch.lookup,
                    (int)P[0]->launch.schedule};
        if (k > 1) {
            ws.h_map.clear();
            ws.h_map.reserve(ni);
            for (int j = 0; j < k; ++j)
                for (uint32_t li = 0; li < P[j]->n_block_items; ++li) ws.h_map.push_back(make_uint2((uint32_t)j, li));
            for (int j = 0; j < k; ++j)
                for (uint32_t li = P[j]->n_block_items; li < P[j]->nitems; ++li)
                    ws.h_map.push_back(make_uint2((uint32_t)j, li));
            const size_t pbytes = pd.size() * sizeof(PlanDev);
            if (ws.plans.n < pbytes) ws.plans.alloc(pbytes);
            if (ws.map.n < ni) ws.map.alloc(ni);
            BTC_HIP(hipMemcpyAsync(ws.plans.p, pd.data(), pbytes, hipMemcpyHostToDevice, s));
            BTC_HIP(hipMemcpyAsync(ws.map.p, ws.h_map.data(), (size_t)ni * sizeof(uint2), hipMemcpyHostToDevice, s));
            a.plans = reinterpret_cast<const PlanDev*>(ws.plans.p);
            a.map = ws.map.p;
            a.nplans = (uint32_t)k;
        }
        hipLaunchKernelGGL(fn, dim3(grid), dim3(kThreads), 0, s, a);
        BTC_CHECK_LAUNCH("wedge_kernel");
    }
    std::vector<unsigned long long> h((size_t)k);
    BTC_HIP(hipMemcpyAsync(h.data(), ws.totals.p, (size_t)k * 8, hipMemcpyDeviceToHost, s));
    BTC_HIP(hipStreamSynchronize(s));
    for (int j = 0; j < k; ++j) totals[j] = h[(size_t)j];
}

void ensure_side_stream(btc_plan_s* p) {
    if (p->side) return;
    BTC_HIP(hipStreamCreateWithFlags(&p->side, hipStreamNonBlocking));
    BTC_HIP(hipStreamCreateWithFlags(&p->side2, hipStreamNonBlocking));
    BTC_HIP(hipEventCreateWithFlags(&p->ev_fork, hipEventDisableTiming));
    BTC_HIP(hipEventCreateWithFlags(&p->ev_join, hipEventDisableTiming));
    BTC_HIP(hipEventCreateWithFlags(&p->ev_join2, hipEventDisableTiming));
}

void plan_enqueue(btc_plan_s* p, hipStream_t s, PvOut d_pv, EventTimer* tm, EventTimer* tw,
                  unsigned long long* d_work, bool make_streams) {
    BTC_REQUIRE(!d_pv.p32 || p->g->max_row <= kPv32MaxRow, BTC_ERR_STATE,
                "longest row %d: per-vertex counts need 64-bit slots", p->g->max_row);
    if (!d_work) d_work = p->scratch.p;
    launch_zero_u32(reinterpret_cast<uint32_t*>(d_work), 2 * kWorkWords, s);  // total, counters and slots
    unsigned long long* d_total = d_work;
    uint32_t* d_counters = reinterpret_cast<uint32_t*>(d_work + 1);
    unsigned long long* d_slots = d_work + kWorkSlots;
    if ((p->n_dense || p->n_tiny_items) && p->nitems && p->info.wedge_checks >= kOverlapChecks &&
        (make_streams || p->side)) {
        // the latency-bound tiny-pivot and dense-tile kernels run on two
        // more streams while the HBM-bound wedge kernel runs on s; all add
        // into d_total (for small plans the fork and join would cost more
        // than they hide).  Both are launched after the wedge kernel: they
        // then take the slots its persistent workgroups leave in its tail
        // (launched first, the tiny kernel slowed the wedge kernel by a
        // third at R-MAT s22).
        ensure_side_stream(p);
        if (tm) tm->start(s);
        BTC_HIP(hipEventRecord(p->ev_fork, s));
        if (tw) tw->start(s);
        launch_wedge(p, d_total, d_counters, d_pv, s, 0, 1, d_slots);
        if (tw) tw->stop(s);
        if (p->n_tiny_items) {
            BTC_HIP(hipStreamWaitEvent(p->side2, p->ev_fork, 0));
            launch_tiny(p, d_total, d_pv, p->side2, 0, 1, d_slots);
            BTC_HIP(hipEventRecord(p->ev_join2, p->side2));
        }
        if (p->n_dense) {
            BTC_HIP(hipStreamWaitEvent(p->side, p->ev_fork, 0));
            launch_plan_dense(p, d_slots, d_pv, p->side);
            BTC_HIP(hipEventRecord(p->ev_join, p->side));
            BTC_HIP(hipStreamWaitEvent(s, p->ev_join, 0));
        }
        if (p->n_tiny_items) BTC_HIP(hipStreamWaitEvent(s, p->ev_join2, 0));
        launch_slot_sum(d_slots, d_total, s);
        if (tm) tm->stop(s);
        return;
    }
    if (tm) tm->start(s);
    if (p->n_dense) launch_plan_dense(p, d_slots, d_pv, s);
    if (tw) tw->start(s);
    if (p->nitems) launch_wedge(p, d_total, d_counters, d_pv, s, 0, 1, d_slots);
    if (tw) tw->stop(s);
    if (p->n_tiny_items) launch_tiny(p, d_total, d_pv, s, 0, 1, d_slots);
    launch_slot_sum(d_slots, d_total, s);
    if (tm) tm->stop(s);
}

namespace {

// One count of p; per-vertex sums go to pv32 (int32) or pv64 (int64) when set.
void plan_count(btc_plan_s* p, hipStream_t s, uint64_t* total, int32_t* pv32, int64_t* pv64, float* kernel_ms) {
    btc_graph_s* g = p->g;
    const bool per_vertex = pv32 || pv64;
    if (!per_vertex && p->count_graph.exec) {
        // replay of the captured totals-only count
        EventTimer tm(kernel_ms != nullptr);
        tm.start(s);
        p->count_graph.launch(s);
        tm.stop(s);
        unsigned long long h = 0;
        BTC_HIP(hipMemcpyAsync(&h, p->scratch.p, 8, hipMemcpyDeviceToHost, s));
        BTC_HIP(hipStreamSynchronize(s));
        if (kernel_ms) {
            *kernel_ms = tm.ms();
            p->last_dense_ms = p->last_wedge_ms = -1.f;  // not split inside a replay
        }
        *total = h;
        return;
    }
    PvBuf pv;
    if (per_vertex) {
        pv.alloc((size_t)g->n, g->max_row);
        BTC_HIP(hipMemsetAsync(pv.data(), 0, pv.bytes(), s));
    }
    EventTimer tm(kernel_ms != nullptr), tw(kernel_ms != nullptr);
    plan_enqueue(p, s, pv.out(), &tm, &tw);
    unsigned long long h = 0;
    BTC_HIP(hipMemcpyAsync(&h, p->scratch.p, 8, hipMemcpyDeviceToHost, s));
    pv.copy_out(pv32, pv64, s);  // synchronises s
    if (kernel_ms) {
        *kernel_ms = tm.ms();
        p->last_wedge_ms = tw.ms();
        p->last_dense_ms = p->n_dense || p->n_tiny_items ? std::max(0.f, *kernel_ms - p->last_wedge_ms) : 0.f;
    }
    *total = h;
}

}  // namespace
}  // namespace btcx

extern "C" {

void btc_count_opts_default(btc_count_opts* o) {
    if (!o) return;
    std::memset(o, 0, sizeof(*o));
    o->tile_k = 0;
    o->tau = -1;
    o->schedule = 1;
}

}  // extern "C"

namespace btcx {
namespace {

// btc_plan_create; with ext (a Bit-BSR of g) the dense tiles are taken from
// it and the plan's tile width is ext's.
btc_plan_s* plan_create(btc_graph g, const btc_count_opts* opts, hipStream_t s, const btc_bitbsr_s* ext) {
    BTC_REQUIRE(g, BTC_ERR_ARG, "NULL argument");
    btc_count_opts o;
    btc_count_opts_default(&o);
    if (opts) o = *opts;
    btcx::tune_table_apply(g, &o);  // auto fields from the autotuner's table, when it has this graph's class
    if (ext) {
        BTC_REQUIRE(ext->n == g->n, BTC_ERR_ARG, "Bit-BSR of %d vertices for a graph of %d", ext->n, g->n);
        BTC_REQUIRE(!opts || opts->tile_k == 0 || opts->tile_k == ext->K, BTC_ERR_ARG,
                    "tile width K = %d differs from the Bit-BSR's %d", opts->tile_k, ext->K);
        o.tile_k = ext->K;
    }
    BTC_REQUIRE(o.tile_k == 0 || o.tile_k == 32 || o.tile_k == 64 || o.tile_k == 128, BTC_ERR_ARG,
                "tile width K = %d (expected 0, 32, 64 or 128)", o.tile_k);
    auto p = std::make_unique<btc_plan_s>();
    p->g = g;
    p->ext_bsr = ext;
    p->K = o.tile_k ? o.tile_k : 64;
    p->tau = o.tau >= 0 ? o.tau : btc_default_tau(p->K);
    BTC_REQUIRE(o.core_ids <= 0 || (o.core_ids % 128 == 0 && (uint32_t)o.core_ids <= btcx::kCoreIdsMax),
                BTC_ERR_ARG, "core_ids %d (expected a multiple of 128 up to %u, 0 = default, < 0 = none)",
                o.core_ids, btcx::kCoreIdsMax);
    p->core_ids = o.core_ids == 0 ? btcx::kCoreIdsDefault : o.core_ids < 0 ? 0u : (uint32_t)o.core_ids;
    BTC_REQUIRE(o.tiny_items >= -1 && o.tiny_items <= 2, BTC_ERR_ARG,
                "tiny_items %d (expected -1, 0, 1 or 2)", o.tiny_items);
    p->use_tiny = o.tiny_items > 0 || (o.tiny_items == 0 && g->wedges >= btcx::kTinyMinWedges);
    p->tiny_all = o.tiny_items == 2;
    btcx::check_launch_opts(o);
    p->launch = o;
    btcx::ensure_device(g, s);
    BTC_REQUIRE((uint64_t)g->m + btcx::kColPad < (1ull << 32), BTC_ERR_OVERFLOW,
                "nnz(L) = %lld: work lists hold u32 positions of L (limit 2^32 - %zu)", (long long)g->m,
                btcx::kColPad);
    btcx::build_col16(p.get(), s);  // ordered before the planners' closing synchronisation
    int64_t chunk = o.chunk_work > 0 ? o.chunk_work : 0;
    if (!chunk) {
        const int64_t cus = btcx::device_cu_count();
        // about 64 items per CU, from 8192 up to 2^20 checks each (R-MAT
        // s22: 2^20 gave 2.82 ms per count, 2^18 2.87 ms, 2^16 3.04 ms)
        chunk = std::max<int64_t>(8192, std::min<int64_t>(1 << 20, g->wedges / std::max<int64_t>(1, cus * 64)));
    }
    const bool on_device = o.planner == 2 || (o.planner == 0 && (!g->on_host || g->m > btcx::kHostPlanEdges));
    if (on_device) {
        btcx::plan_device(p.get(), chunk, s);
    } else {
        btcx::ensure_host(g);
        std::vector<uint8_t> dense;
        btcx::classify_dense(p.get(), s, dense);
        btcx::build_wedge_lists(p.get(), dense, chunk, s);
    }
    if (p->n_dense) p->n_units = btcx::build_dense_units(p->bsr, p->dense_list.p, p->n_dense, p->dense_units, s);
    p->scratch.alloc(btcx::kWorkWords);
    BTC_HIP(hipMemsetAsync(p->scratch.p, 0, p->scratch.bytes(), s));
    BTC_HIP(hipStreamSynchronize(s));  // counts may run on other streams (every count clears these words again)
    p->info.tile_k = p->K;
    p->info.tau = p->tau;
    p->info.core_ids = p->core_ids;
    int64_t bytes =
        (int64_t)(p->lists.bytes() + p->items_dev.bytes() + p->dense_list.bytes() + p->dense_units.bytes() +
                  p->scratch.bytes() +
                  p->col16.bytes());
    if (p->own_bsr)  // a caller's Bit-BSR is the caller's memory
        bytes += (int64_t)(p->own_bsr->indptr.bytes() + p->own_bsr->indices.bytes() +
                           p->own_bsr->row_indices.bytes() + p->own_bsr->tiles.bytes());
    p->info.device_bytes = bytes;
    return p.release();
}

}  // namespace
}  // namespace btcx

extern "C" {

int btc_plan_create(btc_graph g, const btc_count_opts* opts, void* stream, btc_plan* out) {
    return btcx::guard([&] {
        BTC_REQUIRE(out, BTC_ERR_ARG, "NULL argument");
        *out = nullptr;
        *out = btcx::plan_create(g, opts, btcx::as_stream(stream), nullptr);
    });
}

int btc_plan_create_on_bitbsr(btc_graph g, btc_bitbsr b, const btc_count_opts* opts, void* stream,
                              btc_plan* out) {
    return btcx::guard([&] {
        BTC_REQUIRE(out && b, BTC_ERR_ARG, "NULL argument");
        *out = nullptr;
        *out = btcx::plan_create(g, opts, btcx::as_stream(stream), b);
    });
}

int btc_plan_tune(btc_plan p, const btc_count_opts* opts) {
    return btcx::guard([&] {
        BTC_REQUIRE(p && opts, BTC_ERR_ARG, "NULL argument");
        btcx::check_launch_opts(*opts);
        p->launch = *opts;
        p->count_graph.reset();  // a captured count used the old launch shape
    });
}

int btc_plan_count_share(btc_plan p, int32_t shares, int32_t share, void* stream, uint64_t* total,
                         float* kernel_ms) {
    return btcx::guard([&] {
        BTC_REQUIRE(p && total, BTC_ERR_ARG, "NULL argument");
        BTC_REQUIRE(shares >= 1 && share >= 0 && share < shares, BTC_ERR_ARG, "share %d of %d", share, shares);
        hipStream_t s = btcx::as_stream(stream);
        const uint32_t k = (uint32_t)shares, j = (uint32_t)share;
        btcx::launch_zero_u32(reinterpret_cast<uint32_t*>(p->scratch.p), 2 * btcx::kWorkWords, s);
        btcx::EventTimer tm(kernel_ms != nullptr);
        tm.start(s);
        const uint32_t nd = btcx::share_count(p->n_units, j, k);
        const uint32_t nt = btcx::share_count(p->n_tiny_items, j, k);
        // the dense-tile and tiny-pivot parts beside the wedge part, as in
        // a full count
        hipStream_t ds = s, ts = s;
        if ((nd || nt) && p->nitems) {
            btcx::ensure_side_stream(p);
            BTC_HIP(hipEventRecord(p->ev_fork, s));
            if (nd) {
                BTC_HIP(hipStreamWaitEvent(p->side, p->ev_fork, 0));
                ds = p->side;
            }
            if (nt) {
                BTC_HIP(hipStreamWaitEvent(p->side2, p->ev_fork, 0));
                ts = p->side2;
            }
        }
        unsigned long long* slots = p->scratch.p + btcx::kWorkSlots;
        if (p->nitems)
            btcx::launch_wedge(p, p->scratch.p, reinterpret_cast<uint32_t*>(p->scratch.p + 1), btcx::PvOut{}, s, j,
                               k, slots);
        btcx::launch_tiny(p, p->scratch.p, btcx::PvOut{}, ts, j, k, slots);
        if (nd) btcx::launch_plan_dense(p, slots, btcx::PvOut{}, ds, j, k);
        if (ds != s) {
            BTC_HIP(hipEventRecord(p->ev_join, ds));
            BTC_HIP(hipStreamWaitEvent(s, p->ev_join, 0));
        }
        if (ts != s) {
            BTC_HIP(hipEventRecord(p->ev_join2, ts));
            BTC_HIP(hipStreamWaitEvent(s, p->ev_join2, 0));
        }
        btcx::launch_slot_sum(slots, p->scratch.p, s);
        tm.stop(s);
        unsigned long long h = 0;
        BTC_HIP(hipMemcpyAsync(&h, p->scratch.p, 8, hipMemcpyDeviceToHost, s));
        BTC_HIP(hipStreamSynchronize(s));
        if (kernel_ms) *kernel_ms = tm.ms();
        *total = h;
    });
}

int btc_plan_capture(btc_plan p, void* stream) {
    return btcx::guard([&] {
        BTC_REQUIRE(p, BTC_ERR_ARG, "NULL handle");
        (void)stream;  // captured on a private stream; replays run on the count's stream
        p->count_graph.capture([&](hipStream_t cs) { btcx::plan_enqueue(p, cs, btcx::PvOut{}, nullptr, nullptr); });
    });
}

int btc_plan_get_info(btc_plan p, btc_plan_info* info) {
    return btcx::guard([&] {
        BTC_REQUIRE(p && info, BTC_ERR_ARG, "NULL argument");
        *info = p->info;
    });
}

int btc_plan_count(btc_plan p, void* stream, uint64_t* total, int32_t* per_vertex, float* kernel_ms) {
    return btcx::guard([&] {
        BTC_REQUIRE(p && total, BTC_ERR_ARG, "NULL argument");
        btcx::plan_count(p, btcx::as_stream(stream), total, per_vertex, nullptr, kernel_ms);
    });
}

int btc_plan_count64(btc_plan p, void* stream, uint64_t* total, int64_t* per_vertex, float* kernel_ms) {
    return btcx::guard([&] {
        BTC_REQUIRE(p && total, BTC_ERR_ARG, "NULL argument");
        btcx::plan_count(p, btcx::as_stream(stream), total, nullptr, per_vertex, kernel_ms);
    });
}

int btc_plan_last_timing(btc_plan p, float* dense_ms, float* wedge_ms) {
    return btcx::guard([&] {
        BTC_REQUIRE(p, BTC_ERR_ARG, "NULL handle");
        if (dense_ms) *dense_ms = p->last_dense_ms;
        if (wedge_ms) *wedge_ms = p->last_wedge_ms;
    });
}

void btc_plan_free(btc_plan p) { delete p; }

int btc_count(btc_graph g, const btc_count_opts* opts, void* stream, uint64_t* total, int32_t* per_vertex,
              float* kernel_ms) {
    btc_plan p = nullptr;
    int rc = btc_plan_create(g, opts, stream, &p);
    if (rc != BTC_OK) return rc;
    rc = btc_plan_count(p, stream, total, per_vertex, kernel_ms);
    btc_plan_free(p);
    return rc;
}

}  // extern "C"

// ------------------------------------------------- count without allocation
// btc_plan_count_noalloc keeps the count's total, work-item counters and
// per-vertex sums in `scratch`, a static device array of the library (each
// device holds its own copy from module load until the process ends), so a
// count allocates nothing: for callers that cannot call hipMalloc (code
// under stream capture, tight serving loops).  Every such call on a device
// shares its copy; calls that overlap on one device race.

namespace btcx {
namespace {
// The count's words (kWorkWords: total, work-item counters, slots), then
// per-vertex int32 sums from byte kNoAllocPvOffset, at most kNoAllocPvBytes
// of them (2^22 vertices).
constexpr size_t kNoAllocPvOffset = (kWorkWords * 8 + 255) / 256 * 256;
constexpr size_t kNoAllocPvBytes = size_t(16) << 20;
constexpr int kNoAllocDevices = 64;
}  // namespace
}  // namespace btcx

static __device__ __align__(256) unsigned long long
    scratch[(btcx::kNoAllocPvOffset + btcx::kNoAllocPvBytes) / sizeof(unsigned long long)];

namespace btcx {
namespace {

// Device address of `scratch` per device (the array's name is a host-side
// stand-in, not a device address), looked up once per device.
struct NoAllocSlot {
    std::once_flag once;
    unsigned long long* ptr = nullptr;
    hipError_t err = hipSuccess;
};
NoAllocSlot g_noalloc[kNoAllocDevices];

}  // namespace

unsigned long long* noalloc_workspace() {
    int dev = 0;
    BTC_HIP(hipGetDevice(&dev));
    BTC_REQUIRE(dev >= 0 && dev < kNoAllocDevices, BTC_ERR_ARG, "device %d: the static workspace serves devices 0-%d",
                dev, kNoAllocDevices - 1);
    NoAllocSlot& slot = g_noalloc[dev];
    std::call_once(slot.once, [&slot] {
        void* p = nullptr;
        slot.err = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (slot.err == hipSuccess)
            slot.ptr = static_cast<unsigned long long*>(p);
        else
            (void)hipGetLastError();
    });
    BTC_REQUIRE(slot.ptr, BTC_ERR_HIP, "hipGetSymbolAddress of the static workspace on device %d: %s", dev,
                hipGetErrorString(slot.err));
    return slot.ptr;
}

}  // namespace btcx

extern "C" {

int btc_plan_count_noalloc(btc_plan p, void* stream, uint64_t* total, int32_t* per_vertex) {
    return btcx::guard([&] {
        BTC_REQUIRE(p && total, BTC_ERR_ARG, "NULL argument");
        hipStream_t s = btcx::as_stream(stream);
        unsigned long long* work = btcx::noalloc_workspace();
        const size_t n = (size_t)p->g->n;
        btcx::PvOut pv{};
        if (per_vertex) {
            BTC_REQUIRE(p->g->max_row <= btcx::kPv32MaxRow, BTC_ERR_OVERFLOW,
                        "longest row %d > %d: per-vertex sums need 64-bit slots, which this entry point does not "
                        "keep (btc_plan_count64 returns them)", p->g->max_row, btcx::kPv32MaxRow);
            BTC_REQUIRE(n * 4 <= btcx::kNoAllocPvBytes, BTC_ERR_ARG,
                        "per-vertex sums of %zu vertices exceed the static workspace (at most %zu vertices)", n,
                        btcx::kNoAllocPvBytes / 4);
            pv.p32 = reinterpret_cast<int32_t*>(reinterpret_cast<unsigned char*>(work) + btcx::kNoAllocPvOffset);
            btcx::launch_zero_u32(reinterpret_cast<uint32_t*>(pv.p32), n, s);
        }
        btcx::plan_enqueue(p, s, pv, nullptr, nullptr, work, false);
        unsigned long long h = 0;
        BTC_HIP(hipMemcpyAsync(&h, work, 8, hipMemcpyDeviceToHost, s));
        if (per_vertex && n) BTC_HIP(hipMemcpyAsync(per_vertex, pv.p32, n * 4, hipMemcpyDeviceToHost, s));
        BTC_HIP(hipStreamSynchronize(s));
        *total = h;
    });
}

}  // extern "C"
