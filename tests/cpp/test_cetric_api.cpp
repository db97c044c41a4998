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
    std::printf(failures ? "FAILED %d\n" : "ALL OK\n", failures);
    return failures ? 1 : 0;
}
