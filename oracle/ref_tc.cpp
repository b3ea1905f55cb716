// Test infrastructure only: exposes the reference's own CPU verifier,
// btc::cpu_tc_intersection_lower_triangular (btc/operations/triangle_count_cpu.h:60-101),
// compiled from its header where it lies in the reference tree, so the
// oracle and the CPU baseline can be checked against the reference's code.
// The template reads mat.num_rows, mat.row_pointers[i], mat.column_indices[i];
// CsrView supplies exactly those over caller arrays.
#include <btc/operations/triangle_count_cpu.h>

#include <cstdint>

namespace {
struct IntArray {
    const int* p;
    int operator[](size_t i) const { return p[i]; }
};
struct CsrView {
    int num_rows;
    IntArray row_pointers;
    IntArray column_indices;
};
}  // namespace

extern "C" uint64_t ref_cpu_tc_lower(const int32_t* rowptr, const int32_t* col, int32_t n, int32_t threads) {
    if (threads > 0) omp_set_num_threads(threads);
    CsrView m{n, {rowptr}, {col}};
    return (uint64_t)btc::cpu_tc_intersection_lower_triangular(m);
}
