// Host-only build of the library's CPU code (generators, canonicalisation,
// CPU count, graph handles, Lite policy), the readers of btc/btc.h and the
// oracle, linked without any device code so AddressSanitizer and
// UndefinedBehaviorSanitizer watch every host path.  Each case is checked
// against the oracle; bad inputs must come back as status codes or
// btc::Error.  Prints "sanitize ok" on success.
//
// Usage: sanitize_host <scratch directory for reader files>
#include <btc/btc.h>

#include <cstdio>
#include <fstream>
#include <random>
#include <string>
#include <vector>

#include "common.h"

extern "C" {
#include "oracle.h"
}

// Device-side entry points of the full library, left out of this build: a
// graph canonicalised on the host never needs them.
namespace btcx {
void ensure_host(btc_graph_s* g) {
    if (!g->on_host) fail(BTC_ERR_STATE, "no device path in the host-only build");
}
void canonicalise_device(const int32_t*, const int32_t*, int64_t, int32_t, uint32_t, hipStream_t, btc_graph_s*,
                         const std::vector<int64_t>*, std::vector<int64_t>*) {
    fail(BTC_ERR_STATE, "no device path in the host-only build");
}
unsigned long long* noalloc_workspace() { fail(BTC_ERR_STATE, "no device path in the host-only build"); }
}  // namespace btcx

namespace {

int failures = 0;

#define EXPECT(cond)                                                              \
    do {                                                                          \
        if (!(cond)) {                                                            \
            std::fprintf(stderr, "%s:%d: check failed: %s\n", __FILE__, __LINE__, #cond); \
            ++failures;                                                           \
        }                                                                         \
    } while (0)

struct Pairs {
    std::vector<int32_t> s, d;
    int32_t n;
};


Pairs make(const char* kind, int32_t a, int32_t b = 0) {
    Pairs p;
    int64_t m = 0;
    if (std::string(kind) == "rmat") {
        m = btc_gen_rmat(a, b, 0.57, 0.19, 0.19, 7, nullptr, nullptr);
        p.s.resize((size_t)m);
        p.d.resize((size_t)m);
        btc_gen_rmat(a, b, 0.57, 0.19, 0.19, 7, p.s.data(), p.d.data());
        p.n = 1 << a;
    } else if (std::string(kind) == "er") {
        m = btc_gen_er(a, (double)b, 3, nullptr, nullptr);
        p.s.resize((size_t)m);
        p.d.resize((size_t)m);
        btc_gen_er(a, (double)b, 3, p.s.data(), p.d.data());
        p.n = a;
    } else if (std::string(kind) == "grid2") {
        m = btc_gen_stencil2d(a, b, nullptr, nullptr);
        p.s.resize((size_t)m);
        p.d.resize((size_t)m);
        btc_gen_stencil2d(a, b, p.s.data(), p.d.data());
        p.n = a * b;
    } else if (std::string(kind) == "grid3") {
        m = btc_gen_stencil3d(a, a, a, nullptr, nullptr);
        p.s.resize((size_t)m);
        p.d.resize((size_t)m);
        btc_gen_stencil3d(a, a, a, p.s.data(), p.d.data());
        p.n = a * a * a;
    } else {
        m = btc_gen_planted(a, b, 0.5, 4.0, 5, nullptr, nullptr);
        p.s.resize((size_t)m);
        p.d.resize((size_t)m);
        btc_gen_planted(a, b, 0.5, 4.0, 5, p.s.data(), p.d.data());
        p.n = a;
    }
    EXPECT(m >= 0);
    return p;
}

// Library canonicalisation and CPU count against the oracle, every flag set.
void check_graph(const Pairs& p) {
    for (uint32_t flags = 0; flags <= BTC_CANON_DEFAULT; ++flags) {
        btc_graph g = nullptr;
        EXPECT(btc_graph_from_coo(p.s.data(), p.d.data(), (int64_t)p.s.size(), p.n, flags, &g) == BTC_OK);
        if (!g) continue;
        btc_graph_info info;
        EXPECT(btc_graph_get_info(g, &info) == BTC_OK);
        std::vector<int64_t> rp((size_t)info.n + 1);
        std::vector<int32_t> col((size_t)info.m), lab((size_t)info.n);
        EXPECT(btc_graph_copy_csr(g, rp.data(), col.data()) == BTC_OK);
        EXPECT(btc_graph_copy_labels(g, lab.data()) == BTC_OK);
        int64_t *orp = nullptr;
        int32_t *ocol = nullptr, *olab = nullptr, on = 0;
        const int64_t om = orc_canon(p.s.data(), p.d.data(), (int64_t)p.s.size(), p.n, flags, &orp, &ocol, &olab, &on);
        EXPECT(om == info.m && on == info.n);
        if (om == info.m && on == info.n) {
            EXPECT(std::equal(rp.begin(), rp.end(), orp));
            EXPECT(std::equal(col.begin(), col.end(), ocol));
            EXPECT(std::equal(lab.begin(), lab.end(), olab));
        }
        std::vector<int64_t> rows((size_t)info.n);
        const uint64_t want = orc_tc_merge(rp.data(), col.data(), info.n, 2, rows.data());
        for (int32_t threads : {1, 3}) {
            uint64_t t = 0;
            std::vector<int32_t> pv((size_t)info.n);
            EXPECT(btc_count_cpu(g, threads, &t, pv.data()) == BTC_OK);
            EXPECT(t == want);
            for (int32_t v = 0; v < info.n; ++v) EXPECT(pv[(size_t)v] == rows[(size_t)v]);
        }
        // the oracle's tile count at every width, tau and row lookup
        if (flags == BTC_CANON_DEFAULT) {
            for (int K : {32, 64, 128}) {
                const int64_t nb = orc_bitbsr_build(rp.data(), col.data(), info.n, K, nullptr, nullptr, nullptr, nullptr);
                EXPECT(nb >= 0);
                const int32_t n_rb = (info.n + 15) / 16;
                std::vector<uint32_t> indptr((size_t)n_rb + 1), idx((size_t)std::max<int64_t>(nb, 1)),
                    rws((size_t)std::max<int64_t>(nb, 1));
                std::vector<uint64_t> tiles((size_t)std::max<int64_t>(nb * (K / 4), 1));
                orc_bitbsr_build(rp.data(), col.data(), info.n, K, indptr.data(), idx.data(), rws.data(), tiles.data());
                for (int32_t tau : {btc_default_tau(K), 0, 16 * K})
                    for (int32_t variant : {1, 2}) {
                        std::vector<int64_t> pr((size_t)n_rb * 16);
                        int64_t nd = 0;
                        const uint64_t t = orc_bitbsr_count(K, info.n, indptr.data(), idx.data(), rws.data(),
                                                            tiles.data(), (uint32_t)nb, tau, variant, pr.data(), &nd);
                        EXPECT(t == want);
                    }
            }
            if (info.n <= 600) EXPECT(orc_tc_brute(rp.data(), col.data(), info.n, nullptr) == want);
        }
        orc_free(orp);
        orc_free(ocol);
        orc_free(olab);
        btc_graph_free(g);
    }
}

void check_bad_inputs() {
    btc_graph g = nullptr;
    const int32_t s[3] = {0, 1, 5}, d[3] = {1, 2, 0};
    EXPECT(btc_graph_from_coo(s, d, 3, 4, BTC_CANON_DEFAULT, &g) == BTC_ERR_ARG);  // id 5 >= n
    const int32_t sn[2] = {0, -1}, dn[2] = {1, 2};
    EXPECT(btc_graph_from_coo(sn, dn, 2, 4, BTC_CANON_DEFAULT, &g) == BTC_ERR_ARG);  // negative id
    EXPECT(btc_graph_from_coo(s, d, -1, 4, BTC_CANON_DEFAULT, &g) != BTC_OK);
    EXPECT(btc_graph_from_coo(nullptr, d, 3, 8, BTC_CANON_DEFAULT, &g) != BTC_OK);
    EXPECT(btc_graph_from_coo(s, d, 3, 8, BTC_CANON_DEFAULT, nullptr) != BTC_OK);
    EXPECT(btc_graph_from_coo(s, d, 0, 8, BTC_CANON_DEFAULT, &g) == BTC_OK);  // empty graph
    if (g) {
        uint64_t t = 1;
        EXPECT(btc_count_cpu(g, 2, &t, nullptr) == BTC_OK && t == 0);
        btc_graph_free(g);
    }
    // CSR input: unsorted row, upper-triangle entry, bad row pointers
    const int64_t rp[4] = {0, 0, 1, 3};
    const int32_t unsorted[3] = {0, 1, 0}, upper[3] = {0, 2, 1}, ok[3] = {0, 0, 1};
    EXPECT(btc_graph_from_csr(rp, unsorted, 3, &g) != BTC_OK);
    EXPECT(btc_graph_from_csr(rp, upper, 3, &g) != BTC_OK);
    const int64_t rp_bad[4] = {0, 2, 1, 3};
    EXPECT(btc_graph_from_csr(rp_bad, ok, 3, &g) != BTC_OK);
    EXPECT(btc_graph_from_csr(rp, ok, 3, &g) == BTC_OK);
    if (g) {
        uint64_t t = 0;
        EXPECT(btc_count_cpu(g, 1, &t, nullptr) == BTC_OK && t == 1);
        btc_graph_free(g);
    }
    EXPECT(btc_lite_tile_k(16384, 130000, 0) == 32 && btc_lite_tile_k(2400000, 64000000, 0) == 128);
    EXPECT(btc_lite_tile_k(10, 1000, 2) == 128 && btc_lite_tile_k(100000, 3000000, 1) == 32);
    EXPECT(btc_default_tau(128) == 512 && btc_default_tau(64) == 128 && btc_default_tau(32) == 64);
    EXPECT(btc_last_error() != nullptr);
}

template <class F>
bool throws(F&& f) {
    try {
        f();
    } catch (const btc::Error&) {
        return true;
    }
    return false;
}

void check_readers(const std::string& dir) {
    using Csr = btc::CsrMatrix<int, float, btc::device_memory>;
    using Coo = btc::CooMatrix<int, float, btc::device_memory>;
    // the 4-clique 0-1-2-3 plus a pendant edge, written in every format
    {
        std::ofstream f(dir + "/k4.mtx");
        f << "%%MatrixMarket matrix coordinate pattern symmetric\n% comment\n5 5 7\n2 1\n3 1\n4 1\n3 2\n4 2\n4 3\n5 4\n";
    }
    {
        std::ofstream f(dir + "/k4.el");
        f << "# Nodes: 5 Edges: 7\n0 1\n0 2\n0 3\n1 2\n1 3\n2 3\n3 4\n";
    }
    {
        // .csr: nrow, nnz, rowptr, col (lower triangle of the same graph)
        const int32_t hdr[2] = {5, 7}, rp[6] = {0, 0, 1, 3, 6, 7}, col[7] = {0, 0, 1, 0, 1, 2, 3};
        std::ofstream f(dir + "/k4.csr", std::ios::binary);
        f.write((const char*)hdr, 8);
        f.write((const char*)rp, 24);
        f.write((const char*)col, 28);
        // .bin: nrow, ncol, nnz, rowptr, col, values
        const int32_t hb[3] = {5, 5, 7};
        const float val[7] = {1, 1, 1, 1, 1, 1, 1};
        std::ofstream b(dir + "/k4.bin", std::ios::binary);
        b.write((const char*)hb, 12);
        b.write((const char*)rp, 24);
        b.write((const char*)col, 28);
        b.write((const char*)val, 28);
    }
    for (const char* name : {"k4.mtx", "k4.el", "k4.csr", "k4.bin"}) {
        Csr A;
        Coo C;
        btc::read_matrix_file(A, dir + "/" + name);
        btc::convert_csr_to_coo(C, A);
        btc::preprocess_for_triangle_counting(C);
        Csr L;
        btc::convert_coo_to_csr(L, C);
        EXPECT(btc::cpu_tc_intersection_lower_triangular(L) == 4);
    }
    // malformed files are reported, never read past
    {
        const int32_t hdr[2] = {3, 2}, rp[4] = {0, 2, 1, 2}, col[2] = {0, 1};
        std::ofstream f(dir + "/bad_rp.csr", std::ios::binary);
        f.write((const char*)hdr, 8);
        f.write((const char*)rp, 16);
        f.write((const char*)col, 8);
        const int32_t rp2[4] = {0, 1, 1, 2}, col2[2] = {0, 9};
        std::ofstream g(dir + "/bad_col.csr", std::ios::binary);
        g.write((const char*)hdr, 8);
        g.write((const char*)rp2, 16);
        g.write((const char*)col2, 8);
        std::ofstream h(dir + "/short.csr", std::ios::binary);
        h.write((const char*)hdr, 8);
        std::ofstream m(dir + "/neg.mtx");
        m << "%%MatrixMarket matrix coordinate pattern general\n-3 4 1\n1 1\n";
        std::ofstream o(dir + "/out.mtx");
        o << "%%MatrixMarket matrix coordinate pattern general\n3 3 1\n7 1\n";
        std::ofstream a(dir + "/array.mtx");
        a << "%%MatrixMarket matrix array real general\n2 2\n1\n2\n3\n4\n";
    }
    for (const char* name : {"bad_rp.csr", "bad_col.csr", "short.csr", "neg.mtx", "out.mtx", "array.mtx", "none.el",
                             "k4.txt"}) {
        Csr A;
        EXPECT(throws([&] { btc::read_matrix_file(A, dir + "/" + name); }));
    }
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s <scratch dir>\n", argv[0]);
        return 2;
    }
    check_graph(make("rmat", 10, 8));
    check_graph(make("er", 700, 12));
    check_graph(make("grid2", 23, 17));
    check_graph(make("grid3", 9));
    check_graph(make("planted", 1024, 32));
    check_bad_inputs();
    check_readers(argv[1]);
    if (failures) {
        std::fprintf(stderr, "%d checks failed\n", failures);
        return 1;
    }
    std::printf("sanitize ok\n");
    return 0;
}
