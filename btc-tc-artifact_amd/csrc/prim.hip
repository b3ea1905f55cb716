// rocPRIM-backed device sort and scan (see prim.h).
#include <cstring>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include "common.h"
#include "prim.h"

namespace btcx {

namespace {

// out[n] = out[n - 1] + in[n - 1] completes an exclusive scan into a total.
template <class T>
__global__ void scan_total(const T* in, T* out, size_t n) {
    if (threadIdx.x == 0 && blockIdx.x == 0) out[n] = n ? out[n - 1] + in[n - 1] : T(0);
}

template <class T>
void exclusive_scan_t(const T* in, T* out, size_t n, hipStream_t s) {
    if (n) {
        size_t bytes = 0;
        BTC_HIP(rocprim::exclusive_scan(nullptr, bytes, in, out, T(0), n, rocprim::plus<T>(), s));
        TmpBuf<unsigned char> tmp;
        tmp.alloc(bytes, s);
        BTC_HIP(rocprim::exclusive_scan(tmp.p, bytes, in, out, T(0), n, rocprim::plus<T>(), s));
        hipLaunchKernelGGL(scan_total<T>, dim3(1), dim3(64), 0, s, in, out, n);
        BTC_CHECK_LAUNCH("scan_total");
    } else {
        BTC_HIP(hipMemsetAsync(out, 0, sizeof(T), s));
    }
}

}  // namespace

uint64_t* sort_u64(uint64_t* keys, uint64_t* tmp, size_t n, int end_bit, hipStream_t s) {
    if (n < 2 || end_bit <= 0) return keys;
    rocprim::double_buffer<uint64_t> db(keys, tmp);
    size_t bytes = 0;
    BTC_HIP(rocprim::radix_sort_keys(nullptr, bytes, db, n, 0u, (unsigned)end_bit, s));
    TmpBuf<unsigned char> t;
    t.alloc(bytes, s);
    BTC_HIP(rocprim::radix_sort_keys(t.p, bytes, db, n, 0u, (unsigned)end_bit, s));
    return db.current();
}

bool sort_pairs_u32(uint32_t* keys, uint32_t* vals, uint32_t* kt, uint32_t* vt, size_t n, int end_bit,
                    hipStream_t s) {
    if (n < 2 || end_bit <= 0) return false;
    rocprim::double_buffer<uint32_t> dk(keys, kt), dv(vals, vt);
    size_t bytes = 0;
    BTC_HIP(rocprim::radix_sort_pairs(nullptr, bytes, dk, dv, n, 0u, (unsigned)end_bit, s));
    TmpBuf<unsigned char> t;
    t.alloc(bytes, s);
    BTC_HIP(rocprim::radix_sort_pairs(t.p, bytes, dk, dv, n, 0u, (unsigned)end_bit, s));
    return dk.current() == kt;
}

void exclusive_scan_u32(const uint32_t* in, uint32_t* out, size_t n, hipStream_t s) {
    exclusive_scan_t<uint32_t>(in, out, n, s);
}

void exclusive_scan_u64(const uint64_t* in, uint64_t* out, size_t n, hipStream_t s) {
    exclusive_scan_t<uint64_t>(in, out, n, s);
}

}  // namespace btcx
