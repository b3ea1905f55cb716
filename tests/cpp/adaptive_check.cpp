// GPU check of the reference's count entry points through btc/btc.h: read a
// graph file, preprocess it (BTC_REORDER=0 keeps the input's id order),
// build the Bit-BSR at both reference tile widths and count with the
// reference-semantics kernels (v4 / v5 at 16x128, v5 / v6 at 16x32) and with
// the adaptive calls, which run the engine on that Bit-BSR (twice: the
// second call reuses the plan kept in the handle).  Prints "<call> <count>".
#include <btc/btc.h>

int main(int argc, char** argv) try {
    if (argc < 2) return 64;
    btc::check(btc_device_init(0));
    btc::CsrMatrix<int, float, btc::device_memory> A;
    btc::CooMatrix<int, float, btc::device_memory> C;
    btc::read_matrix_file(A, argv[1]);
    btc::convert_csr_to_coo(C, A);
    btc::preprocess_for_triangle_counting(C);
    {
        btc::BCSR_16x128_Device b;
        btc::convert_coo_to_bcsr_16x128_gpu(b, C);
        std::printf("v4_16x128 %llu\n", btc::count_triangles_16x128_v4(b));
        std::printf("v5_16x128 %llu\n", btc::count_triangles_16x128_v5(b));
        float ms = -1.f;
        std::printf("adaptive_16x128 %llu\n", btc::count_triangles_16x128_adaptive(b, 2048, &ms));
        std::printf("adaptive_16x128_again %llu\n", btc::count_triangles_16x128_adaptive(b));
        std::printf("adaptive_16x128_ms_set %d\n", ms >= 0.f ? 1 : 0);
        b.free();
    }
    {
        btc::BCSR_16x32_Device b;
        btc::convert_coo_to_bcsr_16x32_gpu(b, C);
        std::printf("v5_16x32 %llu\n", btc::count_triangles_16x32_v5(b));
        std::printf("v6_16x32 %llu\n", btc::count_triangles_16x32_v6(b));
        std::printf("adaptive_16x32 %llu\n", btc::count_triangles_16x32_adaptive(b));
        std::printf("adaptive_16x32_again %llu\n", btc::count_triangles_16x32_adaptive(b, 40000, nullptr));
        b.free();
    }
    return 0;
} catch (const btc::Error& e) {
    std::fprintf(stderr, "adaptive_check: %s (status %d)\n", e.what(), e.code);
    return 1;
}
