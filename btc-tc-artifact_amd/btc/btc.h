// btc/btc.h -- drop-in C++ surface of the reference's `btc` namespace
// (btc/btc.h of fanna1234/btc-tc-artifact) over the C ABI of libbtc_amd.so.
//
// Same names, argument meaning and results as the reference for the
// triangle-counting path; containers hold host vectors (the C ABI takes
// host buffers and keeps device copies behind handles).  Errors are
// reported as btc::Error exceptions carrying the library status code
// instead of exit()/abort().
//
//   reference                                   here
//   CsrMatrix / CooMatrix (formats/csr.h, coo.h)  same fields, std::vector storage
//   program_options (utils/option.h:25-66)        same flags -i -v -w -r
//   read_matrix_file (utils/io/read.h:395-416)    .mtx .csr .el .bin
//   convert_csr_to_coo / convert_coo_to_csr       same (convert.h:27-45)
//   preprocess_for_triangle_counting (:601-692)   btc_graph_from_coo
//   BCSR_16x128_Device / BCSR_16x32_Device        same public fields
//   convert_coo_to_bcsr_16x{128,32}_gpu           btc_bitbsr_build
//   count_triangles_16x128_{v4,v5,adaptive}       btc_bitbsr_count
//   count_triangles_16x32_{v5,v6,adaptive}        btc_bitbsr_count
//   cpu_tc_intersection_lower_triangular          btc_count_cpu
//   (new) count_triangles_hybrid                  btc_count (the MI355X engine)
#pragma once

#include <btc_amd.h>
#include <getopt.h>

#include <algorithm>
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <numeric>
#include <sstream>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace btc {

constexpr int VERSION_MAJOR = 0;
constexpr int VERSION_MINOR = 1;
constexpr int VERSION_PATCH = 0;

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

inline void check(int rc) {
    if (rc != BTC_OK) throw Error(rc, btc_last_error());
}

struct host_memory {};
struct device_memory {};  // accepted for source compatibility; storage is host

template <typename IndexType, typename ValueType, typename MemorySpace>
class CsrMatrix {
public:
    using index_type = IndexType;
    using value_type = ValueType;
    using memory_space = MemorySpace;
    IndexType num_rows = 0, num_cols = 0, num_entries = 0;
    std::vector<IndexType> row_pointers;
    std::vector<IndexType> column_indices;
    std::vector<ValueType> values;

    void resize(IndexType nrow, IndexType ncol, IndexType nnz) {
        num_rows = nrow;
        num_cols = ncol;
        num_entries = nnz;
        row_pointers.resize((size_t)nrow + 1);
        column_indices.resize((size_t)nnz);
        values.resize((size_t)nnz);
    }
    template <typename Other>
    CsrMatrix& operator=(const CsrMatrix<IndexType, ValueType, Other>& o) {
        num_rows = o.num_rows;
        num_cols = o.num_cols;
        num_entries = o.num_entries;
        row_pointers = o.row_pointers;
        column_indices = o.column_indices;
        values = o.values;
        return *this;
    }
    CsrMatrix() = default;
    CsrMatrix(const CsrMatrix&) = default;
    CsrMatrix& operator=(const CsrMatrix&) = default;
    void free() { *this = CsrMatrix(); }
};

template <typename IndexType, typename ValueType, typename MemorySpace>
class CooMatrix {
public:
    using index_type = IndexType;
    using value_type = ValueType;
    using memory_space = MemorySpace;
    IndexType num_rows = 0, num_cols = 0, num_entries = 0;
    std::vector<IndexType> row_indices;
    std::vector<IndexType> column_indices;
    std::vector<ValueType> values;

    void resize(IndexType nrow, IndexType ncol, IndexType nnz) {
        num_rows = nrow;
        num_cols = ncol;
        num_entries = nnz;
        row_indices.resize((size_t)nnz);
        column_indices.resize((size_t)nnz);
        values.resize((size_t)nnz);
    }
    // (row, col) lexicographic order, values carried along (coo.h:43-51)
    void sort_columns_per_row() {
        std::vector<size_t> o((size_t)num_entries);
        std::iota(o.begin(), o.end(), size_t(0));
        std::stable_sort(o.begin(), o.end(), [&](size_t a, size_t b) {
            return row_indices[a] != row_indices[b] ? row_indices[a] < row_indices[b]
                                                    : column_indices[a] < column_indices[b];
        });
        auto perm = [&](auto& v) {
            auto t = v;
            for (size_t k = 0; k < o.size(); ++k) v[k] = t[o[k]];
        };
        perm(row_indices);
        perm(column_indices);
        perm(values);
    }
    void free() { *this = CooMatrix(); }
};

// ---------------------------------------------------------------- options
struct Config {
    std::string input_file;
    bool verify = false;
    int warmup = 1;
    int repeat = 5;
};

inline const char* option_hints =
    "              [-i input_file]\n"
    "              [-v verify (1 or 0)]\n"
    "              [-w warmup (default 1)]\n"
    "              [-r repeat (default 5)]\n";

// utils/option.h:25-66: same flags, defaults and banner.
inline Config program_options(int argc, char* argv[]) {
    Config config;
    int opt;
    if (argc == 1) {
        printf("Usage: %s ... \n%s", argv[0], option_hints);
        std::exit(EXIT_FAILURE);
    }
    optind = 1;
    while ((opt = getopt(argc, argv, "v:i:w:r:")) != -1) {
        switch (opt) {
            case 'i': config.input_file = optarg; break;
            case 'v': config.verify = std::stoi(optarg); break;
            case 'w': config.warmup = std::stoi(optarg); break;
            case 'r': config.repeat = std::stoi(optarg); break;
            default:
                printf("Usage: %s ... \n%s", argv[0], option_hints);
                std::exit(EXIT_FAILURE);
        }
    }
    printf("\n-----------------config-----------------\n");
    if (config.input_file.empty()) {
        printf("input file is not specified\n");
        std::exit(EXIT_FAILURE);
    }
    printf("input path: %s\n", config.input_file.c_str());
    if (config.verify) printf("verify with CPU result\n");
    printf("warmup: %d, repeat: %d\n", config.warmup, config.repeat);
    printf("----------------------------------------\n\n");
    return config;
}

// --------------------------------------------------------------------- IO
namespace detail {

inline bool ends_with(const std::string& s, const std::string& t) {
    return s.size() >= t.size() && s.compare(s.size() - t.size(), std::string::npos, t) == 0;
}

template <class Csr>
void csr_from_pairs(Csr& m, int32_t n, std::vector<std::pair<int32_t, int32_t>>& e, bool sort_rows) {
    m.resize(n, n, (typename Csr::index_type)e.size());
    std::fill(m.row_pointers.begin(), m.row_pointers.end(), 0);
    for (auto& p : e) {
        if (p.first < 0 || p.first >= n || p.second < 0 || p.second >= n)
            throw Error(BTC_ERR_ARG, "edge (" + std::to_string(p.first) + ", " + std::to_string(p.second) +
                                         ") outside [0, " + std::to_string(n) + ")");
        m.row_pointers[(size_t)p.first + 1]++;
    }
    for (int32_t i = 0; i < n; ++i) m.row_pointers[(size_t)i + 1] += m.row_pointers[(size_t)i];
    std::vector<typename Csr::index_type> pos(m.row_pointers.begin(), m.row_pointers.end() - 1);
    for (auto& p : e) m.column_indices[(size_t)pos[(size_t)p.first]++] = p.second;
    if (sort_rows)
        for (int32_t i = 0; i < n; ++i)
            std::sort(m.column_indices.begin() + m.row_pointers[(size_t)i],
                      m.column_indices.begin() + m.row_pointers[(size_t)i + 1]);
    std::fill(m.values.begin(), m.values.end(), 1);
}

// Matrix Market coordinate file (1-based), any field; symmetric, skew and
// hermitian inputs are expanded to both triangles with the diagonal once
// (utils/io/read.h:34-189).  Values are not kept: counting ignores them.
template <class Csr>
void read_mtx(Csr& m, const std::string& path) {
    std::ifstream f(path);
    if (!f) throw Error(BTC_ERR_ARG, "cannot open " + path);
    std::string line;
    if (!std::getline(f, line) || line.rfind("%%MatrixMarket", 0) != 0)
        throw Error(BTC_ERR_ARG, path + ": missing %%MatrixMarket banner");
    std::string lower = line;
    std::transform(lower.begin(), lower.end(), lower.begin(), ::tolower);
    if (lower.find("coordinate") == std::string::npos)
        throw Error(BTC_ERR_ARG, path + ": only coordinate Matrix Market files are supported");
    const bool sym = lower.find("symmetric") != std::string::npos || lower.find("hermitian") != std::string::npos ||
                     lower.find("skew") != std::string::npos;
    while (std::getline(f, line))
        if (!line.empty() && line[0] != '%') break;
    long long nr = 0, nc = 0, nz = 0;
    std::istringstream hdr(line);
    if (!(hdr >> nr >> nc >> nz)) throw Error(BTC_ERR_ARG, path + ": bad size line");
    if (nr < 0 || nc < 0 || nz < 0 || nr > INT32_MAX || nc > INT32_MAX || nz > INT32_MAX)
        throw Error(BTC_ERR_ARG, path + ": sizes outside [0, 2^31)");
    std::vector<std::pair<int32_t, int32_t>> e;
    e.reserve((size_t)(sym ? 2 * nz : nz));
    for (long long k = 0; k < nz && std::getline(f, line); ++k) {
        if (line.empty() || line[0] == '%') {
            --k;
            continue;
        }
        long long i = 0, j = 0;
        std::istringstream is(line);
        if (!(is >> i >> j)) throw Error(BTC_ERR_ARG, path + ": bad entry line");
        e.emplace_back((int32_t)(i - 1), (int32_t)(j - 1));
        if (sym && i != j) e.emplace_back((int32_t)(j - 1), (int32_t)(i - 1));
    }
    csr_from_pairs(m, (int32_t)std::max(nr, nc), e, true);
}

// SNAP edge list: "# Nodes: N Edges: M" header, "u v" lines, 0-based
// (utils/io/read.h:302-392).
template <class Csr>
void read_el(Csr& m, const std::string& path) {
    std::ifstream f(path);
    if (!f) throw Error(BTC_ERR_ARG, "cannot open " + path);
    std::string line;
    int32_t n = -1, mx = -1;
    std::vector<std::pair<int32_t, int32_t>> e;
    while (std::getline(f, line)) {
        if (line.empty()) continue;
        if (line[0] == '#') {
            auto a = line.find("Nodes:");
            if (a != std::string::npos) n = (int32_t)std::stol(line.substr(a + 6));
            continue;
        }
        std::istringstream is(line);
        long long u, v;
        if (!(is >> u >> v)) continue;
        e.emplace_back((int32_t)u, (int32_t)v);
        mx = std::max<int32_t>(mx, (int32_t)std::max(u, v));
    }
    csr_from_pairs(m, n >= 0 ? n : mx + 1, e, true);
}

// .csr: int32 nrow, nnz, rowptr[nrow+1], col[nnz]; .bin: int32 nrow, ncol,
// nnz, rowptr, col, float values (utils/io/read.h:225-299).
template <class Csr>
void read_binary(Csr& m, const std::string& path, bool with_ncol_values) {
    std::ifstream f(path, std::ios::binary);
    if (!f) throw Error(BTC_ERR_ARG, "cannot open " + path);
    int32_t nrow = 0, ncol = 0, nnz = 0;
    f.read(reinterpret_cast<char*>(&nrow), 4);
    if (with_ncol_values) f.read(reinterpret_cast<char*>(&ncol), 4);
    f.read(reinterpret_cast<char*>(&nnz), 4);
    if (!f || nrow < 0 || nnz < 0) throw Error(BTC_ERR_ARG, path + ": bad header");
    m.resize(nrow, with_ncol_values ? ncol : nrow, nnz);
    std::vector<int32_t> rp((size_t)nrow + 1), cl((size_t)nnz);
    f.read(reinterpret_cast<char*>(rp.data()), (std::streamsize)rp.size() * 4);
    f.read(reinterpret_cast<char*>(cl.data()), (std::streamsize)cl.size() * 4);
    if (!f || rp[0] != 0 || rp[(size_t)nrow] != nnz) throw Error(BTC_ERR_ARG, path + ": truncated or rowptr[n] != nnz");
    for (int32_t i = 0; i < nrow; ++i)
        if (rp[(size_t)i + 1] < rp[(size_t)i]) throw Error(BTC_ERR_ARG, path + ": rowptr decreases");
    const int32_t nc = with_ncol_values ? ncol : nrow;
    for (int32_t c : cl)
        if (c < 0 || c >= nc) throw Error(BTC_ERR_ARG, path + ": column id outside [0, ncol)");
    std::copy(rp.begin(), rp.end(), m.row_pointers.begin());
    std::copy(cl.begin(), cl.end(), m.column_indices.begin());
    std::fill(m.values.begin(), m.values.end(), 1);
}

}  // namespace detail

// utils/io/read.h:395-416
template <class CsrMatrixT>
void read_matrix_file(CsrMatrixT& csr, std::string input) {
    if (input.empty()) return;
    if (detail::ends_with(input, ".mtx")) detail::read_mtx(csr, input);
    else if (detail::ends_with(input, ".csr")) detail::read_binary(csr, input, false);
    else if (detail::ends_with(input, ".el")) detail::read_el(csr, input);
    else if (detail::ends_with(input, ".bin")) detail::read_binary(csr, input, true);
    else throw Error(BTC_ERR_ARG, "input file is NOT supported: " + input);
}

// --------------------------------------------------------------- convert
// convert.h:37-45
template <typename CSR, typename COO>
void convert_csr_to_coo(COO& coo, const CSR& csr) {
    coo.resize(csr.num_rows, csr.num_cols, csr.num_entries);
    for (typename CSR::index_type i = 0; i < csr.num_rows; ++i)
        for (auto p = csr.row_pointers[(size_t)i]; p < csr.row_pointers[(size_t)i + 1]; ++p)
            coo.row_indices[(size_t)p] = i;
    std::copy(csr.column_indices.begin(), csr.column_indices.end(), coo.column_indices.begin());
    std::copy(csr.values.begin(), csr.values.end(), coo.values.begin());
    coo.sort_columns_per_row();
}

// convert.h:27-34 (rows of coo must be sorted, csr_helpers.h:24)
template <typename CSR, typename COO>
void convert_coo_to_csr(CSR& csr, const COO& coo) {
    csr.resize(coo.num_rows, coo.num_cols, coo.num_entries);
    std::fill(csr.row_pointers.begin(), csr.row_pointers.end(), 0);
    for (size_t k = 0; k < coo.row_indices.size(); ++k) {
        if (k && coo.row_indices[k] < coo.row_indices[k - 1])
            throw Error(BTC_ERR_ARG, "convert_coo_to_csr: rows are not sorted");
        csr.row_pointers[(size_t)coo.row_indices[k] + 1]++;
    }
    for (size_t i = 0; i + 1 < csr.row_pointers.size(); ++i) csr.row_pointers[i + 1] += csr.row_pointers[i];
    csr.column_indices = coo.column_indices;
    csr.values = coo.values;
}

// Owning wrapper of a canonical graph handle.
struct Graph {
    btc_graph h = nullptr;
    Graph() = default;
    explicit Graph(btc_graph g) : h(g) {}
    Graph(const Graph&) = delete;
    Graph& operator=(const Graph&) = delete;
    Graph(Graph&& o) noexcept : h(o.h) { o.h = nullptr; }
    Graph& operator=(Graph&& o) noexcept {
        std::swap(h, o.h);
        return *this;
    }
    ~Graph() {
        if (h) btc_graph_free(h);
    }
    btc_graph_info info() const {
        btc_graph_info i;
        check(btc_graph_get_info(h, &i));
        return i;
    }
};

template <typename COO>
Graph graph_from_lower_coo(const COO& coo) {
    // coo holds a canonical L sorted by (row, col)
    const int32_t n = coo.num_rows;
    std::vector<int64_t> rp((size_t)n + 1, 0);
    for (auto r : coo.row_indices) rp[(size_t)r + 1]++;
    for (int32_t i = 0; i < n; ++i) rp[(size_t)i + 1] += rp[(size_t)i];
    std::vector<int32_t> cl(coo.column_indices.begin(), coo.column_indices.end());
    btc_graph g = nullptr;
    check(btc_graph_from_csr(rp.data(), cl.data(), n, &g));
    return Graph(g);
}

// convert.h:601-692.  Symmetrise to the strictly lower triangle, drop
// self-loops, deduplicate and compact ids in order; then, unless
// BTC_REORDER=0, relabel by descending degree (this engine's ordering; the
// reference's HashOrder modes are not provided).  force_cpu is accepted for
// source compatibility (canonicalisation runs on the host); lower_triangular
// must be true.
template <typename COO>
void preprocess_for_triangle_counting(COO& coo, bool force_cpu = false, bool lower_triangular = true,
                                      const std::string& input_file = "") {
    (void)force_cpu;
    (void)input_file;
    if (!lower_triangular) throw Error(BTC_ERR_ARG, "only the lower-triangular orientation is provided");
    printf("[Preprocess] Starting symmetrization for triangle counting...\n");
    printf("[Preprocess] Input: %d rows, %zu entries\n", (int)coo.num_rows, (size_t)coo.num_entries);
    const char* env = std::getenv("BTC_REORDER");
    const uint32_t flags = (env && std::atoi(env) == 0) ? BTC_CANON_COMPACT : BTC_CANON_DEFAULT;
    std::vector<int32_t> s(coo.row_indices.begin(), coo.row_indices.end());
    std::vector<int32_t> d(coo.column_indices.begin(), coo.column_indices.end());
    btc_graph g = nullptr;
    check(btc_graph_from_coo(s.data(), d.data(), (int64_t)s.size(), coo.num_rows, flags, &g));
    Graph G(g);
    const auto info = G.info();
    std::vector<int64_t> rp((size_t)info.n + 1);
    std::vector<int32_t> cl((size_t)info.m);
    check(btc_graph_copy_csr(g, rp.data(), cl.data()));
    coo.resize(info.n, info.n, (typename COO::index_type)info.m);
    for (int32_t i = 0; i < info.n; ++i)
        for (int64_t p = rp[(size_t)i]; p < rp[(size_t)i + 1]; ++p) {
            coo.row_indices[(size_t)p] = i;
            coo.column_indices[(size_t)p] = cl[(size_t)p];
        }
    std::fill(coo.values.begin(), coo.values.end(), 1);
    printf("[Symmetrize] Original edges: %zu, After symmetrize & lower-tri: %zu\n", s.size(), (size_t)info.m);
    printf("[Preprocess] Complete.\n");
}

// ------------------------------------------------------------ Bit-BSR
// bcsr_16x128.h:24-51 / bcsr_16x32.h:109-138: the public fields point at
// device memory owned by the library (blocks = the 16 x K tiles viewed as
// 32-bit words, the reference's layout); result points at nothing: counts
// are returned by value.
template <int K>
struct BCSR_Device {
    uint32_t* indptr = nullptr;
    uint32_t* indices = nullptr;
    uint32_t* blocks = nullptr;
    uint32_t* row_indices = nullptr;
    unsigned long long* result = nullptr;
    int n = 0;
    int n_row_blocks = 0;
    uint32_t num_blocks = 0;
    btc_bitbsr handle = nullptr;
    Graph graph;
    btc_plan plan = nullptr;  // the engine's plan on this Bit-BSR, made by the first adaptive count
    unsigned long long last = 0;

    void free() {
        if (plan) btc_plan_free(plan);  // before the Bit-BSR it reads
        plan = nullptr;
        if (handle) btc_bitbsr_free(handle);
        handle = nullptr;
        graph = Graph();
        indptr = indices = blocks = row_indices = nullptr;
    }
    void reset_result() { last = 0; }
    unsigned long long get_result() const { return last; }
};
using BCSR_16x128_Device = BCSR_Device<128>;
using BCSR_16x32_Device = BCSR_Device<32>;

template <int K, typename COO>
void convert_coo_to_bcsr_gpu(BCSR_Device<K>& b, const COO& coo) {
    b.free();
    b.graph = graph_from_lower_coo(coo);
    check(btc_bitbsr_build(b.graph.h, K, nullptr, &b.handle));
    btc_bitbsr_view v;
    check(btc_bitbsr_get_view(b.handle, &v));
    b.indptr = const_cast<uint32_t*>(v.indptr);
    b.indices = const_cast<uint32_t*>(v.indices);
    b.row_indices = const_cast<uint32_t*>(v.row_indices);
    b.blocks = reinterpret_cast<uint32_t*>(const_cast<uint64_t*>(v.tiles));
    b.n = v.n;
    b.n_row_blocks = v.n_row_blocks;
    b.num_blocks = v.num_blocks;
    printf("[BCSR 16x%d] num_blocks=%u, blocks=%.1f MB\n", K, v.num_blocks,
           (double)v.num_blocks * v.words_per_tile * 8 / (1024.0 * 1024.0));
}

template <typename COO>
void convert_coo_to_bcsr_16x128_gpu(BCSR_16x128_Device& b, const COO& coo) {
    convert_coo_to_bcsr_gpu<128>(b, coo);
}
template <typename COO>
void convert_coo_to_bcsr_16x32_gpu(BCSR_16x32_Device& b, const COO& coo) {
    convert_coo_to_bcsr_gpu<32>(b, coo);
}

template <int K>
unsigned long long count_bitbsr(BCSR_Device<K>& b, int variant, int threshold, float* kernel_ms) {
    uint64_t t = 0;
    check(btc_bitbsr_count(b.handle, variant, threshold, -1, nullptr, &t, nullptr, kernel_ms));
    b.last = t;
    return t;
}

// The MI355X engine on the Bit-BSR's graph: dense tiles of width K from the
// Bit-BSR itself, every other edge by LDS wedge checks; the plan is made on
// the first call and kept in the handle, so later calls only count.
template <int K>
void prepare_engine(BCSR_Device<K>& b) {
    if (!b.plan) check(btc_plan_create_on_bitbsr(b.graph.h, b.handle, nullptr, nullptr, &b.plan));
}
template <int K>
unsigned long long count_engine(BCSR_Device<K>& b, float* kernel_ms) {
    prepare_engine(b);
    uint64_t t = 0;
    check(btc_plan_count(b.plan, nullptr, &t, nullptr, kernel_ms));
    b.last = t;
    return t;
}

// tc_16x128_mma_v4.h:222-260, tc_16x128_mma_v5.h:213-252: the
// reference-semantics kernels (one wave per mask tile, tau routing).
inline unsigned long long count_triangles_16x128_v4(BCSR_16x128_Device& b, float* kernel_ms = nullptr) {
    return count_bitbsr(b, BTC_VARIANT_BINSEARCH, 0, kernel_ms);
}
inline unsigned long long count_triangles_16x128_v5(BCSR_16x128_Device& b, float* kernel_ms = nullptr) {
    return count_bitbsr(b, BTC_VARIANT_DIRECT, 0, kernel_ms);
}
// tc_16x128_adaptive.h:12-29.  The reference picks v4 or v5 by the number of
// tiles (threshold_blocks); on gfx950 the call runs the engine instead, which
// gives the same count.  threshold_blocks applies to count_bitbsr only.
inline unsigned long long count_triangles_16x128_adaptive(BCSR_16x128_Device& b, int threshold_blocks,
                                                          float* kernel_ms) {
    (void)threshold_blocks;
    return count_engine(b, kernel_ms);
}
inline unsigned long long count_triangles_16x128_adaptive(BCSR_16x128_Device& b, int threshold_blocks = 2048) {
    return count_triangles_16x128_adaptive(b, threshold_blocks, nullptr);
}
// tc_16x32_mma_v5.h:246-284, tc_16x32_mma_v6.h:240-279
inline unsigned long long count_triangles_16x32_v5(BCSR_16x32_Device& b, float* kernel_ms = nullptr) {
    return count_bitbsr(b, BTC_VARIANT_BINSEARCH, 0, kernel_ms);
}
inline unsigned long long count_triangles_16x32_v6(BCSR_16x32_Device& b, float* kernel_ms = nullptr) {
    return count_bitbsr(b, BTC_VARIANT_DIRECT, 0, kernel_ms);
}
// tc_16x32_adaptive.h:14-31 (v5 below 40000 tiles, else v6), on the engine
inline unsigned long long count_triangles_16x32_adaptive(BCSR_16x32_Device& b, int threshold_blocks,
                                                         float* kernel_ms) {
    (void)threshold_blocks;
    return count_engine(b, kernel_ms);
}
inline unsigned long long count_triangles_16x32_adaptive(BCSR_16x32_Device& b, int threshold_blocks = 40000) {
    return count_triangles_16x32_adaptive(b, threshold_blocks, nullptr);
}

// The MI355X engine on a preprocessed COO (dense tiles + LDS wedge checks).
template <typename COO>
unsigned long long count_triangles_hybrid(const COO& coo, float* kernel_ms = nullptr,
                                          const btc_count_opts* opts = nullptr) {
    Graph g = graph_from_lower_coo(coo);
    uint64_t t = 0;
    check(btc_count(g.h, opts, nullptr, &t, nullptr, kernel_ms));
    return t;
}

// triangle_count_cpu.h:60-101 on a CSR of L (ascending rows).
template <typename CSR>
size_t cpu_tc_intersection_lower_triangular(const CSR& csr) {
    const int32_t n = csr.num_rows;
    std::vector<int64_t> rp(csr.row_pointers.begin(), csr.row_pointers.end());
    std::vector<int32_t> cl(csr.column_indices.begin(), csr.column_indices.end());
    btc_graph g = nullptr;
    check(btc_graph_from_csr(rp.data(), cl.data(), n, &g));
    Graph G(g);
    uint64_t t = 0;
    check(btc_count_cpu(g, 0, &t, nullptr));
    return (size_t)t;
}

// Host wall-clock timer (utils/timer.h:44-65).
struct CPUTimer {
    std::chrono::high_resolution_clock::time_point a, b;
    void start() { a = std::chrono::high_resolution_clock::now(); }
    void stop() { b = std::chrono::high_resolution_clock::now(); }
    float elapsed() const { return std::chrono::duration<float, std::milli>(b - a).count(); }
};

}  // namespace btc
