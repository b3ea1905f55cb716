// Autotuner: picks the dense tile width K, the density threshold tau, the
// dense-tile kernel form (AND + popcount or i8 MFMA), the hub-core size and
// the wedge kernel's launch shape (lanes per list, loads in flight, resident
// workgroups per CU, static stride or work stealing from one counter) by
// timing candidates on the device, instead
// of the reference's fixed rules (apps/btc_tc_lite.cu:45-65 for K,
// tc_16x128_mma_v5.h:51 / tc_16x32_mma_v6.h:52 for tau).
//
// Results are kept in a table keyed by coarse graph features (log2 of n,
// m, wedge count and longest row), which btc_plan_create consults when K,
// tau or the launch shape are left on auto.  The table is saved and loaded
// as text, one entry per line:
//   btc-tune-3 <n_bits> <m_bits> <w_bits> <r_bits> <K> <tau> <G> <U> <wpe> <dense> <core> <ms>
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <tuple>

#include "common.h"

namespace btcx {
namespace {

using Key = std::tuple<int, int, int, int>;

struct Tuned {
    btc_count_opts opts;
    float ms;
};

std::mutex g_mu;
std::map<Key, Tuned> g_table;

int bits(int64_t x) { return x <= 1 ? 0 : 64 - __builtin_clzll((uint64_t)x - 1); }

Key features(const btc_graph_s* g) { return Key{bits(g->n), bits(g->m), bits(g->wedges), bits(g->max_row)}; }

}  // namespace

// Fills the auto fields of o from the table; returns whether an entry matched.
bool tune_table_apply(const btc_graph_s* g, btc_count_opts* o) {
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_table.find(features(g));
    if (it == g_table.end()) return false;
    const btc_count_opts& t = it->second.opts;
    if (o->tile_k == 0) o->tile_k = t.tile_k;
    if (o->tau < 0) o->tau = t.tau;
    if (o->group_lanes == 0) o->group_lanes = t.group_lanes;
    if (o->unroll == 0) o->unroll = t.unroll;
    if (o->waves_per_eu == 0) o->waves_per_eu = t.waves_per_eu;
    if (o->dense_kernel == 0) o->dense_kernel = t.dense_kernel;
    if (o->core_ids == 0) o->core_ids = t.core_ids;
    return true;
}

}  // namespace btcx

extern "C" {

void btc_tune_opts_default(btc_tune_opts* t) {
    if (!t) return;
    std::memset(t, 0, sizeof(*t));
    t->budget_ms = 4000;
    t->reps = 3;
    t->tile_ks_mask = 7;
}

int btc_autotune(btc_graph g, const btc_tune_opts* topts, void* stream, btc_count_opts* best, float* best_ms) {
    return btcx::guard([&] {
        BTC_REQUIRE(g && best, BTC_ERR_ARG, "NULL argument");
        btc_tune_opts t;
        btc_tune_opts_default(&t);
        if (topts) t = *topts;
        BTC_REQUIRE(t.reps > 0 && t.budget_ms > 0, BTC_ERR_ARG, "reps and budget_ms must be positive");
        const auto t0 = std::chrono::steady_clock::now();
        auto elapsed_ms = [&] {
            return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        };
        float best_t = 1e30f;
        btc_count_opts best_o;
        btc_count_opts_default(&best_o);
        bool any = false;
        // plan-level candidates: K x tau (default tau, twice it, no dense tiles)
        const int ks[3] = {64, 128, 32};
        const int kbit[3] = {2, 4, 1};
        for (int ki = 0; ki < 3; ++ki) {
            if (!(t.tile_ks_mask & kbit[ki])) continue;
            const int K = ks[ki];
            const int taus[3] = {btc_default_tau(K), 2 * btc_default_tau(K), 16 * K};
            for (int tau : taus) {
                if (any && elapsed_ms() > t.budget_ms) break;
                btc_count_opts o;
                btc_count_opts_default(&o);
                o.tile_k = K;
                o.tau = tau;
                btc_plan p = nullptr;
                btcx::check_status(btc_plan_create(g, &o, stream, &p));
                // launch-level candidates on the same plan
                auto time_opts = [&](const btc_count_opts& lo) {
                    int rc = btc_plan_tune(p, &lo);
                    uint64_t total = 0;
                    float ms = 0, mn = 1e30f;
                    if (rc == BTC_OK) rc = btc_plan_count(p, stream, &total, nullptr, &ms);  // warm-up
                    for (int r = 0; r < t.reps && rc == BTC_OK; ++r) {
                        rc = btc_plan_count(p, stream, &total, nullptr, &ms);
                        mn = std::min(mn, ms);
                    }
                    if (rc != BTC_OK) {
                        btc_plan_free(p);
                        btcx::check_status(rc);
                    }
                    if (mn < best_t) {
                        best_t = mn;
                        best_o = lo;
                        any = true;
                    }
                    return mn;
                };
                // {lanes per list, loads in flight, workgroups per CU, schedule}
                const int shapes[6][4] = {{4, 1, 0, 1}, {8, 1, 0, 1}, {4, 2, 0, 1},
                                          {16, 1, 0, 1}, {4, 1, 6, 1}, {4, 1, 0, 0}};
                btc_count_opts plan_best = o;
                float plan_ms = 1e30f;
                for (const auto& sh : shapes) {
                    btc_count_opts lo = o;
                    lo.group_lanes = sh[0];
                    lo.unroll = sh[1];
                    lo.waves_per_eu = sh[2];
                    lo.schedule = sh[3];
                    lo.dense_kernel = 1;
                    const float ms = time_opts(lo);
                    if (ms < plan_ms) {
                        plan_ms = ms;
                        plan_best = lo;
                    }
                }
                // the dense-tile kernel form, on the plan's best shape
                btc_plan_info pi;
                btcx::check_status(btc_plan_get_info(p, &pi));
                if (pi.dense_tiles > 0) {
                    plan_best.dense_kernel = 2;
                    time_opts(plan_best);
                }
                btc_plan_free(p);
            }
        }
        BTC_REQUIRE(any, BTC_ERR_ARG, "no tile width selected by tile_ks_mask");
        // hub-core size on the best plan's tile width, tau and launch shape
        for (int core : {1024, 4096, -1}) {
            if (elapsed_ms() > t.budget_ms) break;
            btc_count_opts o = best_o;
            o.core_ids = core;
            btc_plan p = nullptr;
            btcx::check_status(btc_plan_create(g, &o, stream, &p));
            uint64_t total = 0;
            float ms = 0, mn = 1e30f;
            int rc = btc_plan_count(p, stream, &total, nullptr, &ms);  // warm-up
            for (int r = 0; r < t.reps && rc == BTC_OK; ++r) {
                rc = btc_plan_count(p, stream, &total, nullptr, &ms);
                mn = std::min(mn, ms);
            }
            btc_plan_free(p);
            btcx::check_status(rc);
            if (mn < best_t) {
                best_t = mn;
                best_o = o;
            }
        }
        *best = best_o;
        if (best_ms) *best_ms = best_t;
        std::lock_guard<std::mutex> lk(btcx::g_mu);
        btcx::g_table[btcx::features(g)] = btcx::Tuned{best_o, best_t};
    });
}

int btc_tune_lookup(btc_graph g, btc_count_opts* opts) {
    return btcx::guard([&] {
        BTC_REQUIRE(g && opts, BTC_ERR_ARG, "NULL argument");
        btc_count_opts_default(opts);
        BTC_REQUIRE(btcx::tune_table_apply(g, opts), BTC_ERR_STATE, "no tuned entry for this graph's features");
    });
}

int btc_tune_table_save(const char* path) {
    return btcx::guard([&] {
        BTC_REQUIRE(path, BTC_ERR_ARG, "NULL path");
        FILE* f = std::fopen(path, "w");
        BTC_REQUIRE(f, BTC_ERR_ARG, "cannot write %s", path);
        std::lock_guard<std::mutex> lk(btcx::g_mu);
        for (const auto& kv : btcx::g_table) {
            const auto& o = kv.second.opts;
            std::fprintf(f, "btc-tune-3 %d %d %d %d %d %d %d %d %d %d %d %.4f\n", std::get<0>(kv.first),
                         std::get<1>(kv.first), std::get<2>(kv.first), std::get<3>(kv.first), o.tile_k, o.tau,
                         o.group_lanes, o.unroll, o.waves_per_eu, o.dense_kernel, o.core_ids, kv.second.ms);
        }
        std::fclose(f);
    });
}

int btc_tune_table_load(const char* path) {
    return btcx::guard([&] {
        BTC_REQUIRE(path, BTC_ERR_ARG, "NULL path");
        FILE* f = std::fopen(path, "r");
        BTC_REQUIRE(f, BTC_ERR_ARG, "cannot read %s", path);
        char tag[32];
        int a, b, c, d, K, tau, G, U, w, D, core;
        float ms;
        int line = 0;
        std::map<btcx::Key, btcx::Tuned> tab;
        for (;;) {
            const int r = std::fscanf(f, "%31s %d %d %d %d %d %d %d %d %d %d %d %f", tag, &a, &b, &c, &d, &K, &tau,
                                      &G, &U, &w, &D, &core, &ms);
            if (r == EOF) break;
            ++line;
            if (r != 13 || std::strcmp(tag, "btc-tune-3") != 0 || !(K == 32 || K == 64 || K == 128) ||
                !(G == 4 || G == 8 || G == 16 || G == 32) || !(U == 1 || U == 2 || U == 4) || w < 0 || D < 0 || D > 2 ||
                !(core <= 0 || (core % 128 == 0 && core <= 32768))) {
                std::fclose(f);
                btcx::fail(BTC_ERR_ARG, "%s: bad entry on line %d", path, line);
            }
            btc_count_opts o;
            btc_count_opts_default(&o);
            o.tile_k = K;
            o.tau = tau;
            o.group_lanes = G;
            o.unroll = U;
            o.waves_per_eu = w;
            o.dense_kernel = D;
            o.core_ids = core;
            tab[btcx::Key{a, b, c, d}] = btcx::Tuned{o, ms};
        }
        std::fclose(f);
        std::lock_guard<std::mutex> lk(btcx::g_mu);
        for (auto& kv : tab) btcx::g_table[kv.first] = kv.second;
    });
}

void btc_tune_table_clear(void) {
    std::lock_guard<std::mutex> lk(btcx::g_mu);
    btcx::g_table.clear();
}

}  // extern "C"
