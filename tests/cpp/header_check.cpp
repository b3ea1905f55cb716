// Host-only check of the C++ drop-in surface (btc/btc.h): read a Matrix
// Market file, preprocess, and count on the CPU verify path.  Prints
// "triangles N edges M rows R".
#include <btc/btc.h>

int main(int argc, char** argv) {
    if (argc < 2) return 64;
    btc::CsrMatrix<int, float, btc::device_memory> A;
    btc::CooMatrix<int, float, btc::device_memory> C;
    btc::read_matrix_file(A, argv[1]);
    btc::convert_csr_to_coo(C, A);
    btc::preprocess_for_triangle_counting(C);
    btc::CsrMatrix<int, float, btc::host_memory> L;
    btc::convert_coo_to_csr(L, C);
    std::printf("triangles %zu edges %d rows %d\n", btc::cpu_tc_intersection_lower_triangular(L), C.num_entries,
                C.num_rows);
    return 0;
}
