// Error state, version, device selection, event timing and the scalar
// dispatch rules shared by every path.
#include <algorithm>
#include <cstdarg>
#include <cstring>
#include <mutex>

#include "common.h"

namespace btcx {

static thread_local char g_err[1024] = {0};

void set_error(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

void clear_error() { g_err[0] = 0; }

void fail(int code, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    throw Error{code};
}

EventTimer::EventTimer(bool enable) : on(enable) {
    if (!on) return;
    BTC_HIP(hipEventCreate(&a));
    BTC_HIP(hipEventCreate(&b));
}

EventTimer::~EventTimer() {
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
}

void EventTimer::start(hipStream_t s) {
    if (on) BTC_HIP(hipEventRecord(a, s));
}

void EventTimer::stop(hipStream_t s) {
    if (on) BTC_HIP(hipEventRecord(b, s));
}

float EventTimer::ms() {
    if (!on) return 0.f;
    BTC_HIP(hipEventSynchronize(b));
    float t = 0.f;
    BTC_HIP(hipEventElapsedTime(&t, a, b));
    return t;
}

}  // namespace btcx

namespace btcx {

void mem_pool_init() {
    static std::once_flag once[64];
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return;
    std::call_once(once[dev], [dev] {
        hipMemPool_t pool = nullptr;
        if (hipDeviceGetDefaultMemPool(&pool, dev) != hipSuccess) {
            (void)hipGetLastError();
            return;
        }
        uint64_t keep = 1ull << 30;
        (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
        (void)hipGetLastError();
    });
}

static thread_local Arena* tl_arena = nullptr;

Arena* current_arena() { return tl_arena; }

Arena::~Arena() {
    if (base) (void)hipFreeAsync(base, s);
}

ArenaScope::ArenaScope(Arena& arena, hipStream_t s) : a(arena), prev(tl_arena) {
    a.s = s;
    a.top = 0;
    a.demand = 0;
    tl_arena = &a;
}

ArenaScope::~ArenaScope() {
    tl_arena = prev;
    a.top = 0;
    if (a.demand > a.cap && a.cap < kArenaMax) {
        // grow to the scope's demand (a power of two, at least 64 MiB)
        size_t c = size_t(64) << 20;
        while (c < a.demand && c < kArenaMax) c <<= 1;
        if (a.base) (void)hipFreeAsync(a.base, a.s);
        a.base = nullptr;
        a.cap = 0;
        mem_pool_init();
        if (hipMallocAsync(reinterpret_cast<void**>(&a.base), c, a.s) == hipSuccess) a.cap = c;
        else (void)hipGetLastError();  // the arena stays empty: buffers come from the pool
    }
}

char* PinnedBuf::reserve(size_t bytes) {
    if (bytes <= cap) return p;
    size_t c = size_t(1) << 20;
    while (c < bytes) c <<= 1;
    if (p) (void)hipHostFree(p);
    p = nullptr;
    cap = 0;
    BTC_HIP(hipHostMalloc(reinterpret_cast<void**>(&p), c, hipHostMallocDefault));
    cap = c;
    return p;
}

void upload_arrays(const void* const* host, void* const* dev, int count, size_t bytes, hipStream_t s) {
    constexpr size_t kChunk = size_t(16) << 20;
    if (bytes == 0) return;
    if (bytes <= kChunk / 4) {  // small: one staging copy each
        static thread_local PinnedBuf small;
        char* b = small.reserve(bytes * (size_t)count);
        for (int k = 0; k < count; ++k) {
            std::memcpy(b + (size_t)k * bytes, host[k], bytes);
            BTC_HIP(hipMemcpyAsync(dev[k], b + (size_t)k * bytes, bytes, hipMemcpyHostToDevice, s));
        }
        BTC_HIP(hipStreamSynchronize(s));
        return;
    }
    // two pinned chunks: host threads fill one while the DMA engine drains
    // the other (each chunk's event says when it may be refilled)
    static thread_local PinnedBuf stage;
    char* buf = stage.reserve(2 * kChunk);
    hipEvent_t done[2] = {nullptr, nullptr};
    BTC_HIP(hipEventCreateWithFlags(&done[0], hipEventDisableTiming));
    hipError_t e = hipEventCreateWithFlags(&done[1], hipEventDisableTiming);
    if (e != hipSuccess) {
        (void)hipEventDestroy(done[0]);
        BTC_HIP(e);
    }
    bool used[2] = {false, false};
    int slot = 0;
    e = hipSuccess;
    for (int k = 0; k < count && e == hipSuccess; ++k) {
        const char* src = static_cast<const char*>(host[k]);
        char* dst = static_cast<char*>(dev[k]);
        for (size_t off = 0; off < bytes && e == hipSuccess; off += kChunk) {
            const size_t len = std::min(kChunk, bytes - off);
            char* c = buf + (size_t)slot * kChunk;
            if (used[slot]) e = hipEventSynchronize(done[slot]);
            if (e != hipSuccess) break;
            const int parts = (int)std::max<size_t>(1, std::min<size_t>(16, len >> 20));
#pragma omp parallel for num_threads(parts) schedule(static)
            for (int t = 0; t < parts; ++t) {
                const size_t a = len * (size_t)t / (size_t)parts, b = len * (size_t)(t + 1) / (size_t)parts;
                std::memcpy(c + a, src + off + a, b - a);
            }
            e = hipMemcpyAsync(dst + off, c, len, hipMemcpyHostToDevice, s);
            if (e == hipSuccess) e = hipEventRecord(done[slot], s);
            used[slot] = true;
            slot ^= 1;
        }
    }
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    (void)hipEventDestroy(done[0]);
    (void)hipEventDestroy(done[1]);
    BTC_HIP(e);
}

void PvBuf::copy_out(int32_t* h32, int64_t* h64, hipStream_t s) const {
    if (n == 0 || (!h32 && !h64)) {
        BTC_HIP(hipStreamSynchronize(s));
        return;
    }
    if (!wide() && h32) {
        BTC_HIP(hipMemcpyAsync(h32, b32.p, n * 4, hipMemcpyDeviceToHost, s));
        BTC_HIP(hipStreamSynchronize(s));
        return;
    }
    if (wide() && h64) {
        BTC_HIP(hipMemcpyAsync(h64, b64.p, n * 8, hipMemcpyDeviceToHost, s));
        BTC_HIP(hipStreamSynchronize(s));
        return;
    }
    if (h64) {  // int32 slots widened
        std::vector<int32_t> t(n);
        BTC_HIP(hipMemcpyAsync(t.data(), b32.p, n * 4, hipMemcpyDeviceToHost, s));
        BTC_HIP(hipStreamSynchronize(s));
        for (size_t v = 0; v < n; ++v) h64[v] = t[v];
        return;
    }
    std::vector<unsigned long long> t(n);  // 64-bit slots narrowed, each checked
    BTC_HIP(hipMemcpyAsync(t.data(), b64.p, n * 8, hipMemcpyDeviceToHost, s));
    BTC_HIP(hipStreamSynchronize(s));
    for (size_t v = 0; v < n; ++v) {
        BTC_REQUIRE(t[v] <= (unsigned long long)INT32_MAX, BTC_ERR_OVERFLOW,
                    "vertex %zu closes %llu triangles, beyond int32 per-vertex output (the int64 entry points "
                    "return it)", v, t[v]);
        h32[v] = (int32_t)t[v];
    }
}

}  // namespace btcx

extern "C" {

const char* btc_last_error(void) { return btcx::g_err; }

const char* btc_version(void) { return "0.1.0-gfx950"; }

int btc_device_init(int32_t device) {
    return btcx::guard([&] {
        int count = 0;
        hipError_t e = hipGetDeviceCount(&count);
        if (e != hipSuccess || count == 0) {
            (void)hipGetLastError();
            btcx::fail(BTC_ERR_HIP, "no HIP device available");
        }
        BTC_REQUIRE(device >= 0 && device < count, BTC_ERR_ARG, "device %d out of range [0,%d)",
                    device, count);
        BTC_HIP(hipSetDevice(device));
        btcx::noalloc_workspace();  // once per device, with the device current
    });
}

// apps/btc_tc_lite.cu:45-65: 16x32 if the average L-row is below 15 or the
// graph has fewer than 50,000 rows, else 16x128; BTC_MODE 1/2 force 32/128.
int32_t btc_lite_tile_k(int32_t n, int64_t m, int32_t mode) {
    if (mode == 1) return 32;
    if (mode == 2) return 128;
    if (n <= 0) return 32;
    float avg = (float)m / (float)n;
    return (avg < 15.0f || n < 50000) ? 32 : 128;
}

// tau literals: tc_16x128_mma_v5.h:51 (512 of 2048 bits), tc_16x32_mma_v6.h:52
// (64 of 512 bits); 16x64 is not in the reference and takes 12.5 % (128 of 1024).
int32_t btc_default_tau(int32_t tile_k) {
    switch (tile_k) {
        case 128: return 512;
        case 64: return 128;
        case 32: return 64;
        default: return -1;
    }
}

}  // extern "C"
