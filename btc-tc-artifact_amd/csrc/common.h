// Internal definitions shared by the host (.cpp) and device (.hip) sources of
// libbtc_amd.so.  Errors inside the library are raised as btcx::Error and
// turned into status codes at the C ABI (btcx::guard); nothing exits.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <new>
#include <vector>

#include "btc_amd.h"

namespace btcx {

struct Error {
    int code;
};

void set_error(const char* fmt, ...) __attribute__((format(printf, 1, 2)));
void clear_error();
[[noreturn]] void fail(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));

#define BTC_HIP(call)                                                                        \
    do {                                                                                     \
        hipError_t e_ = (call);                                                              \
        if (e_ != hipSuccess)                                                                \
            ::btcx::fail(e_ == hipErrorOutOfMemory ? BTC_ERR_NOMEM : BTC_ERR_HIP,            \
                         "%s: %s (%s:%d)", #call, hipGetErrorString(e_), __FILE__, __LINE__); \
    } while (0)

#define BTC_CHECK_LAUNCH(name)                                                                \
    do {                                                                                      \
        hipError_t e_ = hipGetLastError();                                                    \
        if (e_ != hipSuccess)                                                                 \
            ::btcx::fail(BTC_ERR_HIP, "launch of %s: %s", name, hipGetErrorString(e_));       \
    } while (0)

#define BTC_REQUIRE(cond, code, ...)          \
    do {                                      \
        if (!(cond)) ::btcx::fail(code, __VA_ARGS__); \
    } while (0)

// Rethrows a failed status of a nested C-ABI call (its message is kept).
inline void check_status(int rc) {
    if (rc != BTC_OK) throw Error{rc};
}

// Runs f and maps library errors to status codes for the C ABI.
template <class F>
int guard(F&& f) {
    clear_error();
    try {
        f();
        return BTC_OK;
    } catch (const Error& e) {
        return e.code;
    } catch (const std::bad_alloc&) {
        set_error("host allocation failed");
        return BTC_ERR_NOMEM;
    } catch (...) {
        set_error("unexpected exception");
        return BTC_ERR_STATE;
    }
}

// Owning device buffer (hipMalloc / hipFree).
template <class T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    bool owned = true;  // false: a view into memory that another buffer owns (view())
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : p(o.p), n(o.n), owned(o.owned) { o.p = nullptr; o.n = 0; o.owned = true; }
    DevBuf& operator=(DevBuf&& o) noexcept {
        if (this != &o) {
            release();
            p = o.p;
            n = o.n;
            owned = o.owned;
            o.p = nullptr;
            o.n = 0;
            o.owned = true;
        }
        return *this;
    }
    // Points at count elements at q, which the buffer does not free.
    void view(T* q, size_t count) {
        release();
        p = q;
        n = count;
        owned = false;
    }
    ~DevBuf() { release(); }
    void alloc(size_t count) {
        release();
        if (count == 0) count = 1;  // keep a valid pointer for empty arrays
        hipError_t e = hipMalloc(reinterpret_cast<void**>(&p), count * sizeof(T));
        if (e != hipSuccess) {
            p = nullptr;
            (void)hipGetLastError();
            fail(BTC_ERR_NOMEM, "hipMalloc of %zu bytes failed: %s", count * sizeof(T),
                 hipGetErrorString(e));
        }
        n = count;
    }
    void release() {
        if (p && owned) (void)hipFree(p);
        p = nullptr;
        n = 0;
        owned = true;
    }
    size_t bytes() const { return n * sizeof(T); }
};

// Keeps up to 1 GiB of freed stream-ordered memory cached in the current
// device's default pool (once per device), so temporaries are reused.
void mem_pool_init();

// A device arena for the temporary buffers (TmpBuf) of one thread's work on
// one stream: buffers are carved from one device block by a bump pointer
// (released in LIFO order, or all at once when the scope ends) and reused
// in stream order, so a thread that prepares many graphs (a server's
// preparation thread) makes no allocation call per buffer.  ArenaScope
// installs it for the calling thread and stream; buffers that do not fit
// come from hipMallocAsync, and the block grows to the scope's demand (up
// to kArenaMax) when the scope ends.
struct Arena {
    char* base = nullptr;
    size_t cap = 0, top = 0, demand = 0;
    hipStream_t s = nullptr;
    Arena() = default;
    Arena(const Arena&) = delete;
    Arena& operator=(const Arena&) = delete;
    ~Arena();
};
constexpr size_t kArenaMax = size_t(1) << 30;

struct ArenaScope {
    Arena& a;
    Arena* prev;
    ArenaScope(Arena& arena, hipStream_t s);
    ~ArenaScope();
};

// The arena installed on this thread, or NULL.
Arena* current_arena();

// Stream-ordered temporary device buffer: from the thread's arena when one
// is installed for this stream, else hipMallocAsync / hipFreeAsync on the
// stream (freeing does not wait for the whole device, so threads that
// build plans side by side do not serialise on each other's kernels).
template <class T>
struct TmpBuf {
    T* p = nullptr;
    size_t n = 0;
    hipStream_t s = nullptr;
    size_t off = 0, bytes = 0;
    bool from_arena = false;
    TmpBuf() = default;
    TmpBuf(const TmpBuf&) = delete;
    TmpBuf& operator=(const TmpBuf&) = delete;
    ~TmpBuf() { release(); }
    void alloc(size_t count, hipStream_t stream) {
        release();
        if (count == 0) count = 1;
        s = stream;
        bytes = (count * sizeof(T) + 255) & ~size_t(255);
        Arena* a = current_arena();
        if (a && a->s == stream) {
            if (a->demand < a->top + bytes) a->demand = a->top + bytes;
            if (a->top + bytes <= a->cap) {
                off = a->top;
                p = reinterpret_cast<T*>(a->base + off);
                a->top += bytes;
                from_arena = true;
                n = count;
                return;
            }
        }
        mem_pool_init();
        hipError_t e = hipMallocAsync(reinterpret_cast<void**>(&p), count * sizeof(T), s);
        if (e != hipSuccess) {
            p = nullptr;
            (void)hipGetLastError();
            fail(BTC_ERR_NOMEM, "hipMallocAsync of %zu bytes failed: %s", count * sizeof(T), hipGetErrorString(e));
        }
        n = count;
    }
    void release() {
        if (p) {
            if (!from_arena) {
                (void)hipFreeAsync(p, s);
            } else {
                Arena* a = current_arena();
                if (a && off + bytes == a->top) a->top = off;  // the newest buffer: pop it
            }
        }
        p = nullptr;
        n = 0;
        from_arena = false;
    }
};

// Pinned host staging buffer, grown on demand (hipHostMalloc).
struct PinnedBuf {
    char* p = nullptr;
    size_t cap = 0;
    PinnedBuf() = default;
    PinnedBuf(const PinnedBuf&) = delete;
    PinnedBuf& operator=(const PinnedBuf&) = delete;
    ~PinnedBuf() {
        if (p) (void)hipHostFree(p);
    }
    char* reserve(size_t bytes);
};

// Copies count host arrays of `bytes` each into device memory (dev[k]) on
// stream s and waits: through a thread's cached pinned staging, two 16 MiB
// chunks at a time for large arrays (host threads fill one chunk while the
// DMA engine drains the other).
void upload_arrays(const void* const* host, void* const* dev, int count, size_t bytes, hipStream_t s);

// A HIP graph captured from a sequence of kernel launches, replayed with
// one hipGraphLaunch.  Captured work contains kernels only (buffers are
// cleared by fill kernels, not memset nodes).
struct CapturedGraph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    CapturedGraph() = default;
    CapturedGraph(const CapturedGraph&) = delete;
    CapturedGraph& operator=(const CapturedGraph&) = delete;
    ~CapturedGraph() { reset(); }
    void reset() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
    }
    // Captures what enqueue(stream) launches on a private stream.
    template <class F>
    void capture(F&& enqueue) {
        reset();
        hipStream_t cs = nullptr;
        BTC_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
        hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
        if (e == hipSuccess) {
            try {
                enqueue(cs);
            } catch (...) {
                hipGraph_t g = nullptr;
                (void)hipStreamEndCapture(cs, &g);
                if (g) (void)hipGraphDestroy(g);
                (void)hipStreamDestroy(cs);
                throw;
            }
            e = hipStreamEndCapture(cs, &graph);
        }
        (void)hipStreamDestroy(cs);
        if (e == hipSuccess) e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        if (e != hipSuccess) {
            reset();
            (void)hipGetLastError();
            fail(BTC_ERR_HIP, "graph capture failed: %s", hipGetErrorString(e));
        }
    }
    void launch(hipStream_t s) { BTC_HIP(hipGraphLaunch(exec, s)); }
};

// Kernel timing with HIP events recorded on the launch stream.
struct EventTimer {
    hipEvent_t a = nullptr, b = nullptr;
    bool on = false;
    explicit EventTimer(bool enable);
    ~EventTimer();
    void start(hipStream_t s);
    void stop(hipStream_t s);
    float ms();  // synchronises on the stop event
};

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

// Device address of the current device's static count workspace (looked up
// on the first call per device; btc_device_init makes that call).
unsigned long long* noalloc_workspace();

// Count totals of kernels with many workgroups: each workgroup adds its
// partial sum into one of kSlots u64 slots kSlotStride words (one 128-byte
// line) apart instead of into one address, where same-address atomics
// serialise in L2; launch_slot_sum then folds the slots into the total and
// clears them.  Every count clears its words (total, counters, slots) first.
constexpr int kSlots = 64;
constexpr int kSlotStride = 16;
constexpr size_t kSlotWords = (size_t)kSlots * kSlotStride;
// A count's device words: [0] total, [1] two u32 work-item counters, the
// slots from word kWorkSlots on.
constexpr size_t kWorkSlots = 16;
constexpr size_t kWorkWords = kWorkSlots + kSlotWords;

// Per-vertex sums on the device.  A vertex's count is at most C(d, 2) for
// its row length d, which fits int32 while d <= kPv32MaxRow; graphs with a
// longer row sum into 64-bit slots.  At most one pointer is set (none:
// totals only).  Every kernel adds a lane's or a group's partial sum with
// one atomic (pv_add).
constexpr int32_t kPv32MaxRow = 65536;

struct PvOut {
    int32_t* p32 = nullptr;
    unsigned long long* p64 = nullptr;
    __host__ __device__ bool on() const { return p32 != nullptr || p64 != nullptr; }
};

// A lane of the tile kernels sums its row's intersections with K / 4 tile
// columns in 32 bits; each is at most the longest row.
inline bool lane_sums_fit(int64_t max_row, int K) { return max_row * (K / 4) < (int64_t(1) << 32); }

// Device per-vertex sums for n vertices of a graph whose longest row is
// max_row (int32 slots when every count fits them, else 64-bit).
struct PvBuf {
    DevBuf<int32_t> b32;
    DevBuf<unsigned long long> b64;
    size_t n = 0;
    void alloc(size_t count, int32_t max_row) {
        n = count;
        if (max_row <= kPv32MaxRow) b32.alloc(count);
        else b64.alloc(count);
    }
    PvOut out() const { return PvOut{b32.p, b64.p}; }
    bool wide() const { return b64.p != nullptr; }
    void* data() const { return b64.p ? (void*)b64.p : (void*)b32.p; }
    size_t bytes() const { return n * (b64.p ? 8 : 4); }
    // Copies the sums to host int32 (h32; BTC_ERR_OVERFLOW when a count
    // does not fit) or int64 (h64) after the work on s; synchronises s.
    void copy_out(int32_t* h32, int64_t* h64, hipStream_t s) const;
};

}  // namespace btcx

// Canonical lower-triangular graph L (CSR, ascending columns) on host and HBM.
struct btc_graph_s {
    int32_t n = 0;
    int32_t n_input = 0;
    int64_t m = 0;
    int32_t max_row = 0;
    uint32_t flags = 0;
    int64_t wedges = 0;
    int64_t merge_work = 0;
    bool on_host = true;           // rowptr / col / labels below are valid
    std::vector<int64_t> rowptr;   // n + 1
    std::vector<int32_t> col;      // m
    std::vector<int32_t> labels;   // canonical -> input id
    btcx::DevBuf<uint32_t> d_rowptr;  // n + 1 (m < 2^32 is enforced)
    btcx::DevBuf<int32_t> d_col;      // m + padding
    btcx::DevBuf<int32_t> d_labels;   // n, for graphs canonicalised on the device
    // one allocation behind the device arrays above (views into it) and,
    // for an exported pipeline run, behind its plan's arrays too
    btcx::DevBuf<unsigned char> block;
};

namespace btcx {
// Fills the derived statistics (rows, wedges, merge work) of g.
void graph_finalize(btc_graph_s* g);
// Canonicalises device-resident COO pairs into g (CSR and labels in HBM,
// statistics filled, host copies left empty until ensure_host).  With
// segs (ascending vertex offsets of up to 64 independent graphs, segs[0] =
// 0, back = n), each graph stays one contiguous vertex range, relabelled by
// degree within itself; csegs receives the ranges after compaction.
void canonicalise_device(const int32_t* src, const int32_t* dst, int64_t nnz, int32_t n, uint32_t flags,
                         hipStream_t s, btc_graph_s* g, const std::vector<int64_t>* segs,
                         std::vector<int64_t>* csegs);
// Makes the host CSR and labels of g valid (copies them from HBM once).
void ensure_host(btc_graph_s* g);
// Fills the auto fields of o from the autotuner's table when an entry
// matches g's features; returns whether one did.
bool tune_table_apply(const btc_graph_s* g, btc_count_opts* o);
}  // namespace btcx
