// Seeded synthetic graph generators (host, OpenMP).
//
// Every random number is a counter-based SplitMix64 draw (Steele, Lea and
// Flood, "Fast splittable pseudorandom number generators", OOPSLA 2014):
//     draw(seed, ctr) = mix64(seed + (ctr + 1) * 0x9E3779B97F4A7C15)
// so each output is a pure function of (seed, counter) and the result does
// not depend on the thread count.  The counters used by each generator are
// fixed below; changing them changes every seeded workload.
#include <omp.h>

#include <cmath>
#include <cstring>

#include "common.h"

namespace {

constexpr uint64_t kGamma = 0x9E3779B97F4A7C15ULL;

inline uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

inline uint64_t draw(uint64_t seed, uint64_t ctr) { return mix64(seed + (ctr + 1) * kGamma); }

// Probability -> 32-bit threshold: P(r < t) = t / 2^32 for uniform 32-bit r.
inline uint64_t thresh32(double p) {
    if (p <= 0) return 0;
    if (p >= 1) return 1ULL << 32;
    return (uint64_t)(p * 4294967296.0);
}

// Uniform double in (0, 1] from the top 53 bits of a draw.
inline double unit_open0(uint64_t w) { return (double)((w >> 11) + 1) * (1.0 / 9007199254740992.0); }

// Exclusive prefix sum of per-part counts; returns the total.
int64_t exclusive_scan(std::vector<int64_t>& v) {
    int64_t s = 0;
    for (auto& x : v) {
        int64_t t = x;
        x = s;
        s += t;
    }
    return s;
}

// G(n, p) pairs (i, j), j < i, by geometric skipping along each row i.
// Row i uses counters (i << 32) + k, k = 0, 1, ... (one per skip).
template <class Emit>
void er_row(uint64_t seed, int32_t i, double p, double inv_log_q, Emit&& emit) {
    if (p <= 0.0) return;
    if (p >= 1.0) {
        for (int32_t j = 0; j < i; ++j) emit(i, j);
        return;
    }
    int64_t j = -1;
    uint64_t k = 0;
    for (;;) {
        double u = unit_open0(draw(seed, ((uint64_t)i << 32) + k++));
        double skip = std::floor(std::log(u) * inv_log_q);
        if (skip >= (double)i) break;
        j += 1 + (int64_t)skip;
        if (j >= i) break;
        emit(i, (int32_t)j);
    }
}

int64_t gen_er_impl(int32_t n, double p, uint64_t seed, int32_t* src, int32_t* dst, int64_t base) {
    const double inv_log_q = (p > 0 && p < 1) ? 1.0 / std::log1p(-p) : 0.0;
    std::vector<int64_t> cnt((size_t)n, 0);
#pragma omp parallel for schedule(dynamic, 256)
    for (int32_t i = 1; i < n; ++i) {
        int64_t c = 0;
        er_row(seed, i, p, inv_log_q, [&](int32_t, int32_t) { ++c; });
        cnt[i] = c;
    }
    int64_t total = exclusive_scan(cnt);
    if (src && dst) {
#pragma omp parallel for schedule(dynamic, 256)
        for (int32_t i = 1; i < n; ++i) {
            int64_t o = base + cnt[i];
            er_row(seed, i, p, inv_log_q, [&](int32_t a, int32_t b) {
                src[o] = a;
                dst[o] = b;
                ++o;
            });
        }
    }
    return total;
}

}  // namespace

extern "C" {

// R-MAT (Chakrabarti, Zhan, Faloutsos 2004).  Edge e (of edge_factor << scale)
// descends `scale` levels; level l uses the 32-bit half (l even: high, odd:
// low) of draw(seed, e * ceil(scale/2) + l/2), picks quadrant (0,0) below
// a, (0,1) below a+b, (1,0) below a+b+c, else (1,1), and sets bit l of the
// source / destination id.
int64_t btc_gen_rmat(int32_t scale, int32_t edge_factor, double a, double b, double c,
                     uint64_t seed, int32_t* src, int32_t* dst) {
    if (scale < 1 || scale > 30 || edge_factor < 1 || a < 0 || b < 0 || c < 0 ||
        a + b + c > 1.0 + 1e-12) {
        btcx::set_error("btc_gen_rmat: bad arguments (scale %d, edge_factor %d)", scale, edge_factor);
        return -1;
    }
    const int64_t m = (int64_t)edge_factor << scale;
    if (!src || !dst) return m;
    const uint64_t ta = thresh32(a), tab = thresh32(a + b), tabc = thresh32(a + b + c);
    const uint64_t per_edge = (uint64_t)(scale + 1) / 2;
#pragma omp parallel for schedule(static, 65536)
    for (int64_t e = 0; e < m; ++e) {
        uint32_t u = 0, v = 0;
        uint64_t w = 0;
        for (int l = 0; l < scale; ++l) {
            if ((l & 1) == 0) w = draw(seed, (uint64_t)e * per_edge + (uint64_t)(l >> 1));
            uint64_t r = (l & 1) ? (w & 0xFFFFFFFFULL) : (w >> 32);
            uint32_t bu, bv;
            if (r < ta) { bu = 0; bv = 0; }
            else if (r < tab) { bu = 0; bv = 1; }
            else if (r < tabc) { bu = 1; bv = 0; }
            else { bu = 1; bv = 1; }
            u |= bu << l;
            v |= bv << l;
        }
        src[e] = (int32_t)u;
        dst[e] = (int32_t)v;
    }
    return m;
}

// Erdos-Renyi G(n, p) with p = avg_degree / (n - 1): every pair j < i is an
// edge independently with probability p (geometric skipping per row).
int64_t btc_gen_er(int32_t n, double avg_degree, uint64_t seed, int32_t* src, int32_t* dst) {
    if (n < 1 || avg_degree < 0) {
        btcx::set_error("btc_gen_er: bad arguments (n %d)", n);
        return -1;
    }
    double p = n > 1 ? avg_degree / (double)(n - 1) : 0.0;
    return gen_er_impl(n, p, seed, src, dst, 0);
}

// 9-point stencil on an nx x ny grid, vertex (x, y) = y * nx + x; each of the
// 4 forward offsets (1,0), (-1,1), (0,1), (1,1) emits one pair.
int64_t btc_gen_stencil2d(int32_t nx, int32_t ny, int32_t* src, int32_t* dst) {
    if (nx < 1 || ny < 1 || (int64_t)nx * ny > INT32_MAX) {
        btcx::set_error("btc_gen_stencil2d: bad grid %d x %d", nx, ny);
        return -1;
    }
    static const int off[4][2] = {{1, 0}, {-1, 1}, {0, 1}, {1, 1}};
    int64_t total = 0;
    for (int32_t y = 0; y < ny; ++y)
        for (int32_t x = 0; x < nx; ++x)
            for (auto& o : off) {
                int32_t x2 = x + o[0], y2 = y + o[1];
                if (x2 < 0 || x2 >= nx || y2 >= ny) continue;
                if (src && dst) {
                    src[total] = y * nx + x;
                    dst[total] = y2 * nx + x2;
                }
                ++total;
            }
    return total;
}

// 27-point stencil on nx x ny x nz, vertex (x, y, z) = (z * ny + y) * nx + x;
// the 13 lexicographically positive offsets in {-1,0,1}^3 emit one pair each.
int64_t btc_gen_stencil3d(int32_t nx, int32_t ny, int32_t nz, int32_t* src, int32_t* dst) {
    if (nx < 1 || ny < 1 || nz < 1 || (int64_t)nx * ny * nz > INT32_MAX) {
        btcx::set_error("btc_gen_stencil3d: bad grid %d x %d x %d", nx, ny, nz);
        return -1;
    }
    int off[13][3];
    int k = 0;
    for (int dz = -1; dz <= 1; ++dz)
        for (int dy = -1; dy <= 1; ++dy)
            for (int dx = -1; dx <= 1; ++dx)
                if (dz > 0 || (dz == 0 && dy > 0) || (dz == 0 && dy == 0 && dx > 0)) {
                    off[k][0] = dx; off[k][1] = dy; off[k][2] = dz; ++k;
                }
    // Per-plane counts make the fill parallel and deterministic.
    std::vector<int64_t> cnt((size_t)nz, 0);
    auto plane = [&](int32_t z, bool write, int64_t o) {
        int64_t c = 0;
        for (int32_t y = 0; y < ny; ++y)
            for (int32_t x = 0; x < nx; ++x)
                for (auto& d : off) {
                    int32_t x2 = x + d[0], y2 = y + d[1], z2 = z + d[2];
                    if (x2 < 0 || x2 >= nx || y2 < 0 || y2 >= ny || z2 >= nz) continue;
                    if (write) {
                        src[o + c] = (z * ny + y) * nx + x;
                        dst[o + c] = (z2 * ny + y2) * nx + x2;
                    }
                    ++c;
                }
        return c;
    };
#pragma omp parallel for schedule(dynamic, 1)
    for (int32_t z = 0; z < nz; ++z) cnt[z] = plane(z, false, 0);
    int64_t total = exclusive_scan(cnt);
    if (src && dst) {
#pragma omp parallel for schedule(dynamic, 1)
        for (int32_t z = 0; z < nz; ++z) plane(z, true, cnt[z]);
    }
    return total;
}

// Planted dense communities: vertices [c*C, min((c+1)*C, n)) form community
// c; each pair inside it is an edge when the high 32 bits of
// draw(seed, pair_index) fall below p_in * 2^32 (pair_index enumerates the
// pairs of all communities in order).  G(n, noise/(n-1)) noise is added with
// seed mix64(seed ^ 0x6E6F697365).
int64_t btc_gen_planted(int32_t n, int32_t community, double p_in, double noise_degree,
                        uint64_t seed, int32_t* src, int32_t* dst) {
    if (n < 1 || community < 2 || p_in < 0 || p_in > 1 || noise_degree < 0) {
        btcx::set_error("btc_gen_planted: bad arguments");
        return -1;
    }
    const int32_t ncomm = (n + community - 1) / community;
    const uint64_t t = thresh32(p_in);
    const uint64_t full_pairs = (uint64_t)community * (community - 1) / 2;
    std::vector<int64_t> cnt((size_t)ncomm, 0);
    auto comm = [&](int32_t c, bool write, int64_t o) {
        int32_t lo = c * community, hi = std::min(n, lo + community);
        uint64_t pidx = (uint64_t)c * full_pairs;
        int64_t k = 0;
        for (int32_t i = lo + 1; i < hi; ++i)
            for (int32_t j = lo; j < i; ++j, ++pidx) {
                if ((draw(seed, pidx) >> 32) < t) {
                    if (write) {
                        src[o + k] = i;
                        dst[o + k] = j;
                    }
                    ++k;
                }
            }
        return k;
    };
#pragma omp parallel for schedule(dynamic, 4)
    for (int32_t c = 0; c < ncomm; ++c) cnt[c] = comm(c, false, 0);
    int64_t intra = exclusive_scan(cnt);
    if (src && dst) {
#pragma omp parallel for schedule(dynamic, 4)
        for (int32_t c = 0; c < ncomm; ++c) comm(c, true, cnt[c]);
    }
    double p = n > 1 ? noise_degree / (double)(n - 1) : 0.0;
    int64_t noise = gen_er_impl(n, p, mix64(seed ^ 0x6E6F697365ULL), src, dst, intra);
    return intra + noise;
}

}  // extern "C"
