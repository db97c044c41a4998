// katric_amd C++ host layer: the reference's entry points for the local
// counting path, over the C ABI of katric.h / katric_shard.h.
//
// Mirrors (same names and argument meaning; paths relative to the reference
// root):
//   cetric::run_shmem(DistributedGraph<>&, profiling::Statistics&, const Config&)
//       include/cetric/counters/shared_memory_cetric.h:105-170
//   cetric::profiling::Statistics counters used on this path
//       include/cetric/statistics.h:193-203
//   cetric::Config knobs of this path (intersection method, threads)
//       include/cetric/config.h:140-197
//   DistributedGraph construction from a (first_out, head) CSR, rank 0 of 1
//       include/cetric/datastructures/distributed/distributed_graph.h:797-819
// and adds the rank-shard and batch entry points.
//
//   CetricEdgeIterator::run_local(emit, stats, ...) with its Triangle{v, u, w}
//       callback: counters/cetric_edge_iterator.h:173-298 (emit at :289),
//       datastructures/graph_definitions.h:187-190 (Triangle)
//
// Differences, by design: run_local lists the triangles on the device and
// then calls emit on the host, one triangle after another (the reference
// calls it concurrently from its worker threads, in no particular order).
// Config's knobs map to device knobs (kt_params) as device_params() states:
// merge and hybrid -> the degree-binned LDS hash probe, with binary search
// for sets beyond the LDS (itself a hybrid rule); binary_search -> binary
// search for every tail; the recursive binary intersection has no device
// form and is rejected; pseudo2core -> the same skip on the device;
// num_threads (host worker threads) must be at least 1 and does not change
// the device run.  Counts are identical for every choice.  Errors raise
// katric::Error (the reference aborts through KASSERT in debug builds and has
// no status).
#pragma once

#include <algorithm>
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../katric.h"
#include "../katric_shard.h"

namespace katric {

class Error : public std::runtime_error {
public:
    Error(int status, const std::string& what) : std::runtime_error(what), status_(status) {}
    int status() const { return status_; }

private:
    int status_;
};

enum class IntersectionMethod { merge, binary_search, binary, hybrid };  // config.h:129

struct Config {  // the fields of cetric::Config read on this path (config.h:140-197)
    IntersectionMethod intersection_method = IntersectionMethod::merge;
    size_t binary_intersection_cutoff = 1000;
    double hybrid_cutoff_scale = 1.0;
    bool pseudo2core = false;
    size_t num_threads = 1;
    int device = 0;
    uint64_t hbm_budget_bytes = 0;
};

namespace profiling {
struct LocalStatistics {
    uint64_t skipped_nodes = 0;
    uint64_t wedge_checks = 0;
    uint64_t intersection_size_local = 0;
    uint64_t intersection_size_global = 0;
    uint64_t local_triangles = 0;
    uint64_t type3_triangles = 0;
    double preprocessing_time = 0;  // seconds, device
    double local_phase_time = 0;    // seconds, device
};
struct Statistics {
    uint64_t counted_triangles = 0;
    LocalStatistics local;
};
}  // namespace profiling

// RAII owner of a kt_ctx.
class Context {
public:
    explicit Context(int device = 0, uint64_t hbm_budget_bytes = 0) {
        kt_ctx* c = nullptr;
        int rc = kt_create(&c, device, hbm_budget_bytes);
        if (rc != KT_OK) throw Error(rc, std::string("kt_create: ") + kt_strerror(rc));
        ctx_.reset(c);
    }
    kt_ctx* get() const { return ctx_.get(); }
    void check(int rc, const char* what) const {
        if (rc != KT_OK) throw Error(rc, std::string(what) + ": " + kt_last_error(ctx_.get()));
    }

private:
    struct Del {
        void operator()(kt_ctx* c) const { kt_destroy(c); }
    };
    std::unique_ptr<kt_ctx, Del> ctx_;
};

namespace graph {
// Bits of a RankEncodedNodeId (datastructures/graph_definitions.h:48-90): the
// id in bits 0-47, the owner rank in bits 48-63 (0xffff: not yet set).
constexpr unsigned RANK_BITS = 16;
constexpr unsigned NON_RANK_BITS = 64 - RANK_BITS;
constexpr uint64_t RANK_MASK = ((1ull << RANK_BITS) - 1) << NON_RANK_BITS;

// A symmetric CSR held on the host; one rank (rank 0 of 1), as constructed by
// DistributedGraph(graphio::Graph&&) (distributed_graph.h:797-819).
class DistributedGraph {
public:
    DistributedGraph() : first_out_(1, 0) {}
    DistributedGraph(std::vector<uint64_t> first_out, std::vector<uint32_t> head)
        : first_out_(std::move(first_out)), head_(std::move(head)) {
        if (first_out_.empty() || first_out_.front() != 0 || first_out_.back() != head_.size())
            throw Error(KT_EINVAL, "DistributedGraph: offsets do not describe the head array");
    }
    // From the reference's own arrays: first_out_ (n + 1 offsets) and head_
    // as RankEncodedNodeId values (distributed_graph.h:1251-1254).  The rank
    // bits are stripped (this graph is one rank's whole view, ids global);
    // an id that does not fit 32 bits or is not below n is rejected.
    static DistributedGraph from_rank_encoded(std::vector<uint64_t> first_out,
                                              const std::vector<uint64_t>& rank_encoded_head) {
        if (first_out.empty()) throw Error(KT_EINVAL, "DistributedGraph: empty offsets");
        const uint64_t n = first_out.size() - 1;
        if (n > 0xffffffffull) throw Error(KT_EINVAL, "DistributedGraph: more than 2^32 vertices");
        std::vector<uint32_t> head(rank_encoded_head.size());
        for (size_t i = 0; i < head.size(); ++i) {
            const uint64_t id = rank_encoded_head[i] & ~RANK_MASK;
            if (id >= n) throw Error(KT_EINVAL, "DistributedGraph: head id " + std::to_string(id) + " >= n");
            head[i] = (uint32_t)id;
        }
        return DistributedGraph(std::move(first_out), std::move(head));
    }
    uint64_t local_node_count() const { return first_out_.size() - 1; }
    uint64_t local_edge_count() const { return head_.size(); }
    uint64_t degree(uint64_t v) const { return first_out_[v + 1] - first_out_[v]; }
    const std::vector<uint64_t>& first_out() const { return first_out_; }
    const std::vector<uint32_t>& head() const { return head_; }

private:
    std::vector<uint64_t> first_out_;
    std::vector<uint32_t> head_;
};
}  // namespace graph

using graph::DistributedGraph;

namespace detail {
inline Context& default_context(const Config& conf) {
    static Context ctx(conf.device, conf.hbm_budget_bytes);
    return ctx;
}
inline void fill_stats(const Context& ctx, profiling::Statistics& stats, uint64_t counted) {
    kt_stats s{};
    kt_timing t{};
    kt_last_stats(ctx.get(), &s);
    kt_last_timing(ctx.get(), &t);
    stats.counted_triangles = counted;
    stats.local.skipped_nodes = s.skipped_nodes;
    stats.local.wedge_checks = s.wedge_checks;
    stats.local.intersection_size_local = s.intersection_size_local;
    stats.local.intersection_size_global = s.intersection_size_global;
    stats.local.local_triangles = s.local_triangles;
    stats.local.type3_triangles = s.type3_triangles;
    stats.local.preprocessing_time = t.prep_ms * 1e-3;
    stats.local.local_phase_time = t.count_ms * 1e-3;
}

// The device knobs for conf, starting from the context's current ones (the
// autotuned walk and bins stay): see the mapping at the top of this file.
inline kt_params device_params(const Context& ctx, const Config& conf) {
    if (conf.num_threads < 1) throw Error(KT_EINVAL, "Config::num_threads must be at least 1");
    kt_params p{};
    ctx.check(kt_get_params(ctx.get(), &p), "kt_get_params");
    switch (conf.intersection_method) {
        case IntersectionMethod::merge:
        case IntersectionMethod::hybrid: p.bsearch_min_deg = 8192; break;
        case IntersectionMethod::binary_search: p.bsearch_min_deg = 1; break;
        case IntersectionMethod::binary:
            throw Error(KT_EINVAL, "Config::intersection_method = binary has no device form; counts do not "
                                   "depend on the method (use merge)");
    }
    p.pseudo2core = conf.pseudo2core ? 1u : 0u;
    return p;
}

// Runs f with the knobs of conf installed, then restores the context's own.
template <class F>
auto with_config(Context& ctx, const Config& conf, F&& f) {
    kt_params saved{};
    ctx.check(kt_get_params(ctx.get(), &saved), "kt_get_params");
    const kt_params p = device_params(ctx, conf);
    ctx.check(kt_set_params(ctx.get(), &p), "kt_set_params");
    struct Restore {
        Context& c;
        kt_params s;
        ~Restore() { kt_set_params(c.get(), &s); }
    } restore{ctx, saved};
    return f();
}
}  // namespace detail

// cetric::run_shmem: number of triangles of G (preprocessing + local phase on
// the device).  G is not modified (the reference orients it in place; here the
// oriented CSR lives in device memory only).
inline size_t run_shmem(Context& ctx, DistributedGraph& G, profiling::Statistics& stats, const Config& conf) {
    return detail::with_config(ctx, conf, [&] {
        uint64_t total = 0;
        ctx.check(kt_count(ctx.get(), (uint32_t)G.local_node_count(), G.first_out().data(), G.head().data(), 0,
                           &total, nullptr),
                  "kt_count");
        detail::fill_stats(ctx, stats, total);
        return (size_t)total;
    });
}

inline size_t run_shmem(DistributedGraph& G, profiling::Statistics& stats, const Config& conf) {
    return run_shmem(detail::default_context(conf), G, stats, conf);
}

// Total plus per-vertex counts (t[x] = triangles containing x).
inline uint64_t count_triangles(Context& ctx, const DistributedGraph& G, std::vector<uint64_t>* per_vertex) {
    uint64_t total = 0;
    if (per_vertex) per_vertex->assign(G.local_node_count(), 0);
    ctx.check(kt_count(ctx.get(), (uint32_t)G.local_node_count(), G.first_out().data(), G.head().data(), 0, &total,
                       per_vertex ? per_vertex->data() : nullptr),
              "kt_count");
    return total;
}

// A triangle as the reference's emit callback receives it: x = v (the tail,
// lowest in the (degree, id) order), y = u in N+(v), z = w in N+(v) and N+(u).
struct Triangle {
    uint32_t x, y, z;
};

// Every triangle of G once (kt_list_triangles): one counting pass sizes the
// buffer, the listing pass fills it.
inline std::vector<Triangle> list_triangles(Context& ctx, const DistributedGraph& G) {
    const uint32_t n = (uint32_t)G.local_node_count();
    uint64_t total = 0;
    int rc = kt_list_triangles(ctx.get(), n, G.first_out().data(), G.head().data(), 0, nullptr, 0, &total);
    if (rc != KT_OK && rc != KT_ERANGE) ctx.check(rc, "kt_list_triangles");
    std::vector<Triangle> out(total);
    if (total)
        ctx.check(kt_list_triangles(ctx.get(), n, G.first_out().data(), G.head().data(), 0,
                                    reinterpret_cast<uint32_t*>(out.data()), total, &total),
                  "kt_list_triangles");
    return out;
}

// CetricEdgeIterator::run_local-shaped entry: emit(Triangle) for every
// triangle of G, then the statistics of the device run.
template <class Emit>
void run_local(Context& ctx, const DistributedGraph& G, Emit&& emit, profiling::Statistics& stats) {
    const std::vector<Triangle> tri = list_triangles(ctx, G);
    for (const Triangle& t : tri) emit(t);
    detail::fill_stats(ctx, stats, tri.size());
}

// One rank's share of CeTriC ("cetric-x") or DiTriC ("patric"): the emits of
// its local phase and of its contracted / distributed phase (run_cetric_new,
// counters/cetric.h:665-1002; run_patric, :160-316).
struct RankCounts {
    uint64_t local_triangles = 0;
    uint64_t type3_triangles = 0;
    std::vector<uint64_t> per_vertex;  // owned vertices, then ghosts
};

inline RankCounts count_rank(Context& ctx, const kt_shard_in& in, bool per_vertex) {
    RankCounts r;
    const uint64_t nl = in.range_bounds[in.rank + 1] - in.range_bounds[in.rank];
    if (per_vertex) r.per_vertex.assign(nl + in.n_ghost, 0);
    kt_shard_out out{0, 0, per_vertex ? r.per_vertex.data() : nullptr};
    ctx.check(kt_count_shard(ctx.get(), &in, &out), "kt_count_shard");
    r.local_triangles = out.local_triangles;
    r.type3_triangles = out.type3_triangles;
    return r;
}

// Rank `rank` of `num_ranks` contiguous ranges of G, its inputs built on the
// device from G (kt_count_shard_graph); per_vertex (nullable) receives n
// counts indexed by global id.
inline RankCounts count_rank(Context& ctx, const DistributedGraph& G, uint32_t num_ranks, uint32_t rank,
                             uint32_t algorithm, const std::vector<uint64_t>& range_bounds, bool per_vertex) {
    RankCounts r;
    if (per_vertex) r.per_vertex.assign(G.local_node_count(), 0);
    kt_shard_out out{0, 0, per_vertex ? r.per_vertex.data() : nullptr};
    ctx.check(kt_count_shard_graph(ctx.get(), (uint32_t)G.local_node_count(), G.first_out().data(), G.head().data(),
                                   0, algorithm, num_ranks, range_bounds.data(), rank, &out),
              "kt_count_shard_graph");
    r.local_triangles = out.local_triangles;
    r.type3_triangles = out.type3_triangles;
    return r;
}

// Many independent graphs in one device pipeline; per-graph totals.
inline std::vector<uint64_t> count_batch(Context& ctx, const std::vector<DistributedGraph>& graphs) {
    std::vector<uint64_t> vbase(graphs.size() + 1, 0), off(1, 0);
    std::vector<uint32_t> heads;
    for (size_t g = 0; g < graphs.size(); ++g) {
        const auto& G = graphs[g];
        const uint64_t ng = G.local_node_count();
        vbase[g + 1] = vbase[g] + ng;
        if (vbase[g + 1] >= 0xffffffffull) throw Error(KT_EINVAL, "count_batch: more than 2^32 - 1 vertices");
        const uint64_t eb = heads.size();
        for (uint64_t v = 0; v < ng; ++v) off.push_back(eb + G.first_out()[v + 1]);
        for (uint32_t h : G.head()) {
            // a head outside its own graph must not become a vertex of the next one
            if (h >= ng) throw Error(KT_EINVAL, "count_batch: graph " + std::to_string(g) + " has a head id >= n");
            heads.push_back((uint32_t)(h + vbase[g]));
        }
    }
    std::vector<uint64_t> totals(graphs.size(), 0);
    ctx.check(kt_batch_count(ctx.get(), (uint32_t)graphs.size(), vbase.data(), off.data(), heads.data(), 0,
                             totals.data(), nullptr),
              "kt_batch_count");
    return totals;
}

// cetric::intersection for lists of the caller's (intersection.h:40-79):
// on_intersection(w) for every common element of the strictly increasing
// ranges a and b, ascending.  Many pairs are better served by one
// kt_intersect call; this form intersects one pair on the device.
template <class OnIntersection>
void intersection(Context& ctx, const std::vector<uint32_t>& a, const std::vector<uint32_t>& b,
                  OnIntersection&& on_intersection) {
    const uint64_t aoff[2] = {0, a.size()}, boff[2] = {0, b.size()};
    uint64_t count = 0, total = 0;
    std::vector<uint32_t> out(std::max<size_t>(std::min(a.size(), b.size()), 1));
    ctx.check(kt_intersect(ctx.get(), 1, aoff, a.data(), boff, b.data(), 0, &count, out.data(), out.size(), &total),
              "kt_intersect");
    for (uint64_t i = 0; i < total; ++i) on_intersection(out[i]);
}

}  // namespace katric
