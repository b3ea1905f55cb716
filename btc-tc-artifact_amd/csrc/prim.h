// Device-wide sort and scan used by the device canonicaliser and planner
// (rocPRIM underneath; kept in one translation unit for compile time).
// Every function runs on `s` and returns once the work is enqueued; its
// temporary storage is allocated and freed in stream order on `s`.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace btcx {

// Sorts keys[0, n) ascending on bits [0, end_bit); tmp holds n keys.
// Returns the buffer that holds the result (keys or tmp).
uint64_t* sort_u64(uint64_t* keys, uint64_t* tmp, size_t n, int end_bit, hipStream_t s);

// Stable sort of (key, value) pairs on key bits [0, end_bit); kt / vt hold n
// entries each.  Returns whether the result is in (kt, vt) rather than
// (keys, vals).
bool sort_pairs_u32(uint32_t* keys, uint32_t* vals, uint32_t* kt, uint32_t* vt, size_t n, int end_bit,
                    hipStream_t s);

// out[i] = sum of in[0, i) for i <= n (out has n + 1 entries; out[n] = total).
void exclusive_scan_u32(const uint32_t* in, uint32_t* out, size_t n, hipStream_t s);
void exclusive_scan_u64(const uint64_t* in, uint64_t* out, size_t n, hipStream_t s);

// Capturable forms (devprim.hip).  The element count n is read from device
// memory (*d_n) when the kernels run, buffers hold `cap` elements (n <= cap),
// and the caller's workspace `ws` (dprim_ws_bytes(cap) bytes, 256-byte
// aligned, zeroed once before its first use) holds every temporary: the
// kernels are the only work enqueued, so one stream capture serves every n
// up to cap.  Each call leaves the workspace's tickets, totals and statuses
// zeroed again for the next call on the same stream; calls sharing a
// workspace must not overlap.

size_t dprim_ws_bytes(size_t cap);

// The workspace: its device address and the capacity it was sized for
// (its layout follows from that, whatever the capacity of a call).
struct DevWs {
    void* p = nullptr;
    size_t cap = 0;
};

// out[i] = sum of in[0, i) for i <= n; out[n] and *d_total (when set) get
// the total.
void dscan_u32(const uint32_t* in, uint32_t* out, size_t cap, const uint32_t* d_n, uint32_t* d_total,
               const DevWs& ws, hipStream_t s);
void dscan_u64(const uint64_t* in, uint64_t* out, size_t cap, const uint32_t* d_n, uint64_t* d_total,
               const DevWs& ws, hipStream_t s);

// Word of ws that a sort or scan sets (and nothing clears) if a look-back
// ever ran past its spin limit -- a result that cannot be trusted.
const uint32_t* dsort_error_word(const DevWs& ws);

// Stable radix sorts on bits [0, end_bit), as sort_u64 / sort_pairs_u32.
uint64_t* dsort_u64(uint64_t* keys, uint64_t* tmp, size_t cap, const uint32_t* d_n, int end_bit, const DevWs& ws,
                    hipStream_t s);
bool dsort_pairs_u32(uint32_t* keys, uint32_t* vals, uint32_t* kt, uint32_t* vt, size_t cap, const uint32_t* d_n,
                     int end_bit, const DevWs& ws, hipStream_t s);

}  // namespace btcx
