// C++ host-layer test: the reference-shaped entry points of
// include/katric/cetric.hpp against known answers (examples/examples.yaml of
// the reference) and closed forms.  Needs a GPU; prints one line per check.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "../../include/katric/cetric.hpp"
#include "../../include/katric_graphio.h"

static int failures = 0;
#define CHECK(cond)                                                      \
    do {                                                                 \
        if (!(cond)) {                                                   \
            std::printf("FAIL %s:%d %s\n", __FILE__, __LINE__, #cond);   \
            ++failures;                                                  \
        }                                                                \
    } while (0)

static katric::DistributedGraph load(const std::string& path) {
    kt_host_graph h{};
    if (kt_read_metis(path.c_str(), &h) != 0) {
        std::printf("cannot read %s\n", path.c_str());
        std::exit(2);
    }
    std::vector<uint64_t> off(h.offsets, h.offsets + h.n + 1);
    std::vector<uint32_t> hd(h.heads, h.heads + h.nnz);
    kt_host_graph_free(&h);
    return katric::DistributedGraph(std::move(off), std::move(hd));
}

static katric::DistributedGraph complete(uint32_t n) {
    std::vector<uint64_t> off(1, 0);
    std::vector<uint32_t> hd;
    for (uint32_t v = 0; v < n; ++v) {
        for (uint32_t u = 0; u < n; ++u)
            if (u != v) hd.push_back(u);
        off.push_back(hd.size());
    }
    return katric::DistributedGraph(std::move(off), std::move(hd));
}

int main(int argc, char** argv) {
    const std::string dir = argc > 1 ? argv[1] : "tests/golden";
    katric::Config conf;
    katric::Context ctx(0);
    struct {
        const char* file;
        size_t triangles;
    } known[] = {{"triangle.metis", 1}, {"metis-sample.metis", 4}, {"hiv.metis", 45}};
    std::vector<katric::DistributedGraph> graphs;
    for (auto& k : known) {
        auto G = load(dir + "/" + k.file);
        katric::profiling::Statistics stats;
        size_t t = katric::run_shmem(ctx, G, stats, conf);
        std::printf("run_shmem %s = %zu\n", k.file, t);
        CHECK(t == k.triangles);
        CHECK(stats.counted_triangles == k.triangles);
        CHECK(stats.local.intersection_size_local >= stats.local.wedge_checks);
        graphs.push_back(G);
    }
    {
        auto K = complete(40);
        std::vector<uint64_t> pv;
        uint64_t t = katric::count_triangles(ctx, K, &pv);
        CHECK(t == 40ull * 39 * 38 / 6);
        for (uint64_t x : pv) CHECK(x == 39ull * 38 / 2);
        graphs.push_back(K);
        std::printf("K_40 = %llu\n", (unsigned long long)t);
    }
    {
        auto totals = katric::count_batch(ctx, graphs);
        CHECK(totals.size() == 4 && totals[0] == 1 && totals[1] == 4 && totals[2] == 45 && totals[3] == 9880);
        std::printf("batch = %llu %llu %llu %llu\n", (unsigned long long)totals[0], (unsigned long long)totals[1],
                    (unsigned long long)totals[2], (unsigned long long)totals[3]);
    }
    {
        // hiv split over 3 ranks, both distributed algorithms: ranks sum to 45
        const auto& G = graphs[2];
        const uint32_t n = (uint32_t)G.local_node_count();
        std::vector<uint64_t> bounds = {0, n / 3, 2 * n / 3, n};
        for (uint32_t algo : {KT_SHARD_CETRIC, KT_SHARD_DITRIC}) {
            uint64_t sum = 0;
            for (uint32_t r = 0; r < 3; ++r) {
                kt_shard_host h{};
                CHECK(kt_shard_make(n, G.first_out().data(), G.head().data(), 3, bounds.data(), r, algo, &h) == 0);
                kt_shard_in in{algo, r, 3, 0, bounds.data(), h.offsets, h.heads, h.n_ghost, h.ghost_ids,
                               h.ghost_degree, h.n_msg, h.msg_src, h.msg_offsets, h.msg_heads};
                auto rc = katric::count_rank(ctx, in, true);
                sum += rc.local_triangles + rc.type3_triangles;
                kt_shard_host_free(&h);
            }
            std::printf("shard algo %u over 3 ranks = %llu\n", algo, (unsigned long long)sum);
            CHECK(sum == 45);
        }
    }
    {
        // the same ranks with their inputs built on the device from the graph
        const auto& G = graphs[2];
        const uint32_t n = (uint32_t)G.local_node_count();
        std::vector<uint64_t> bounds = {0, n / 3, 2 * n / 3, n};
        for (uint32_t algo : {KT_SHARD_CETRIC, KT_SHARD_DITRIC}) {
            uint64_t sum = 0, pv_sum = 0;
            for (uint32_t r = 0; r < 3; ++r) {
                auto rc = katric::count_rank(ctx, G, 3, r, algo, bounds, true);
                sum += rc.local_triangles + rc.type3_triangles;
                for (uint64_t x : rc.per_vertex) pv_sum += x;
            }
            std::printf("device-produced shards algo %u over 3 ranks = %llu\n", algo, (unsigned long long)sum);
            CHECK(sum == 45);
            CHECK(pv_sum == 3 * 45);
        }
    }
    {
        // run_local-shaped listing: every triangle once, tail first in the
        // (degree, id) order, all three edges present
        const auto& G = graphs[2];
        std::vector<katric::Triangle> seen;
        katric::profiling::Statistics stats;
        katric::run_local(ctx, G, [&](const katric::Triangle& t) { seen.push_back(t); }, stats);
        CHECK(seen.size() == 45 && stats.counted_triangles == 45);
        auto has_edge = [&](uint32_t a, uint32_t b) {
            for (uint64_t e = G.first_out()[a]; e < G.first_out()[a + 1]; ++e)
                if (G.head()[e] == b) return true;
            return false;
        };
        auto before = [&](uint32_t a, uint32_t b) {
            return G.degree(a) < G.degree(b) || (G.degree(a) == G.degree(b) && a < b);
        };
        std::vector<uint64_t> keys;
        for (const auto& t : seen) {
            CHECK(has_edge(t.x, t.y) && has_edge(t.x, t.z) && has_edge(t.y, t.z));
            CHECK(before(t.x, t.y) && before(t.y, t.z));
            uint32_t a[3] = {t.x, t.y, t.z};
            std::sort(a, a + 3);
            keys.push_back(((uint64_t)a[0] << 42) | ((uint64_t)a[1] << 21) | a[2]);
        }
        std::sort(keys.begin(), keys.end());
        CHECK(std::adjacent_find(keys.begin(), keys.end()) == keys.end());
        auto K = complete(30);
        CHECK(katric::list_triangles(ctx, K).size() == 30ull * 29 * 28 / 6);
        std::printf("run_local hiv = %zu triangles listed\n", seen.size());
    }
    {
        bool threw = false;
        try {
            katric::DistributedGraph bad({0, 1, 1}, {1});  // asymmetric
            katric::profiling::Statistics stats;
            katric::run_shmem(ctx, bad, stats, conf);
        } catch (const katric::Error& e) {
            threw = e.status() == KT_EINVAL;
        }
        CHECK(threw);
    }
    {
        // asymmetric although half of the slots are kept: 0 -> 1 and 1 -> 2
        bool threw = false;
        try {
            katric::DistributedGraph bad({0, 1, 2, 2, 2}, {1, 2});
            katric::profiling::Statistics stats;
            katric::run_shmem(ctx, bad, stats, conf);
        } catch (const katric::Error& e) {
            threw = e.status() == KT_EINVAL;
        }
        CHECK(threw);
        // a batch graph whose head lies outside it
        threw = false;
        try {
            std::vector<katric::DistributedGraph> bg = {complete(3), katric::DistributedGraph({0, 1, 2}, {1, 2})};
            katric::count_batch(ctx, bg);
        } catch (const katric::Error& e) {
            threw = e.status() == KT_EINVAL;
        }
        CHECK(threw);
    }
    {
        // the reference's own arrays: heads as RankEncodedNodeId values with
        // the rank bits set (0xffff before find_ghost_ranks, or a rank)
        const auto& G = graphs[2];
        std::vector<uint64_t> enc(G.head().size());
        for (size_t i = 0; i < enc.size(); ++i)
            enc[i] = (uint64_t)G.head()[i] | ((uint64_t)(i % 3 ? 0xffffu : 7u) << katric::graph::NON_RANK_BITS);
        auto H = katric::DistributedGraph::from_rank_encoded(G.first_out(), enc);
        katric::profiling::Statistics stats;
        CHECK(katric::run_shmem(ctx, H, stats, conf) == 45);
        bool threw = false;
        try {
            enc[0] = (uint64_t)G.local_node_count() | (3ull << katric::graph::NON_RANK_BITS);
            katric::DistributedGraph::from_rank_encoded(G.first_out(), enc);
        } catch (const katric::Error& e) {
            threw = e.status() == KT_EINVAL;
        }
        CHECK(threw);
        std::printf("rank-encoded hiv = 45\n");
    }
    {
        // Config knobs: every accepted intersection method gives the same
        // count; pseudo2core skips the tails with fewer than two out-neighbours
        const auto& G = graphs[2];
        for (auto m : {katric::IntersectionMethod::merge, katric::IntersectionMethod::binary_search,
                       katric::IntersectionMethod::hybrid}) {
            katric::Config c2;
            c2.intersection_method = m;
            katric::profiling::Statistics stats;
            CHECK(katric::run_shmem(ctx, const_cast<katric::DistributedGraph&>(G), stats, c2) == 45);
            CHECK(stats.local.skipped_nodes == 0);
        }
        katric::Config c3;
        c3.pseudo2core = true;
        katric::profiling::Statistics s3;
        CHECK(katric::run_shmem(ctx, const_cast<katric::DistributedGraph&>(G), s3, c3) == 45);
        // out-degrees under the (degree, id) order, counted here
        uint64_t skipped = 0;
        for (uint64_t v = 0; v < G.local_node_count(); ++v) {
            uint64_t dplus = 0;
            for (uint64_t e = G.first_out()[v]; e < G.first_out()[v + 1]; ++e) {
                const uint32_t u = G.head()[e];
                dplus += (G.degree(v) < G.degree(u) || (G.degree(v) == G.degree(u) && v < u)) ? 1 : 0;
            }
            skipped += dplus < 2 ? 1 : 0;
        }
        CHECK(s3.local.skipped_nodes == skipped && skipped > 0);
        std::printf("pseudo2core hiv: skipped %llu\n", (unsigned long long)s3.local.skipped_nodes);
        bool threw = false;
        try {
            katric::Config c4;
            c4.intersection_method = katric::IntersectionMethod::binary;
            katric::profiling::Statistics s4;
            katric::run_shmem(ctx, const_cast<katric::DistributedGraph&>(G), s4, c4);
        } catch (const katric::Error& e) {
            threw = e.status() == KT_EINVAL;
        }
        CHECK(threw);
        threw = false;
        try {
            katric::Config c5;
            c5.num_threads = 0;
            katric::profiling::Statistics s5;
            katric::run_shmem(ctx, const_cast<katric::DistributedGraph&>(G), s5, c5);
        } catch (const katric::Error& e) {
            threw = e.status() == KT_EINVAL;
        }
        CHECK(threw);
        // the context's own knobs are back
        kt_params p{};
        kt_get_params(ctx.get(), &p);
        CHECK(p.pseudo2core == 0 && p.bsearch_min_deg == 8192);
    }
    {
        // cetric::intersection's hand vectors (tests/intersection_test.cpp:41-52),
        // both argument orders
        struct Case {
            std::vector<uint32_t> a, b, want;
        } cases[] = {{{0, 1, 3, 6, 9, 478}, {1, 6, 7, 8, 478}, {1, 6, 478}},
                     {{}, {1, 6, 7, 8, 478}, {}},
                     {{}, {}, {}},
                     {{1}, {}, {}},
                     {{1}, {1}, {1}},
                     {{1}, {42}, {}},
                     {{13, 46, 54, 58, 64}, {6, 21, 43, 47, 75}, {}},
                     {{1, 2, 5, 9, 10}, {1, 2, 4, 5, 6, 7, 8}, {1, 2, 5}},
                     {{2, 3, 4, 7, 10}, {3, 5, 6}, {3}},
                     {{1, 2, 3, 4, 5}, {6, 7, 8, 9, 10, 11, 12}, {}}};
        for (const auto& cs : cases) {
            for (int swap = 0; swap < 2; ++swap) {
                std::vector<uint32_t> got;
                katric::intersection(ctx, swap ? cs.b : cs.a, swap ? cs.a : cs.b,
                                     [&](uint32_t w) { got.push_back(w); });
                CHECK(got == cs.want);
            }
        }
        bool threw = false;
        try {
            katric::intersection(ctx, std::vector<uint32_t>{3, 2}, std::vector<uint32_t>{2}, [](uint32_t) {});
        } catch (const katric::Error& e) {
            threw = e.status() == KT_EINVAL;
        }
        CHECK(threw);
        std::printf("intersection hand vectors: %zu cases\n", sizeof(cases) / sizeof(cases[0]));
    }
    std::printf(failures ? "FAILED %d\n" : "ALL OK\n", failures);
    return failures ? 1 : 0;
}
