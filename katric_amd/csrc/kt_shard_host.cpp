// Host-side producer of one rank's shard inputs from a full symmetric CSR.
//
// In the reference these inputs reach a rank through MPI: its local CSR by
// the 1D partitioned read, the ghost degrees by DegreeCommunicator
// (datastructures/distributed/graph_communicator.h:48-266), and the
// contracted-phase messages [a, N+cut(a)] by the message queue
// (counters/cetric_edge_iterator.h:1328-1379, 1632-1661).  Here they are
// computed directly from the whole graph so that one GPU can count one
// rank's share; the counting itself never uses the full graph.
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <vector>

#include "../../include/katric_shard.h"

namespace {

inline bool deg_less(const uint64_t* off, uint32_t a, uint32_t b) {
    uint64_t da = off[a + 1] - off[a], db = off[b + 1] - off[b];
    return da < db || (da == db && a < b);
}

inline int owner_of(uint64_t x, const uint64_t* bounds, int P) {
    // upper_bound over the range ends (helpers.h:83-101)
    int lo = 0, hi = P;
    while (lo < hi) {
        int mid = (lo + hi) / 2;
        if (x < bounds[mid + 1])
            hi = mid;
        else
            lo = mid + 1;
    }
    return lo;
}

int nthreads() {
    const char* e = getenv("KATRIC_THREADS");
    if (!e) e = getenv("OMP_NUM_THREADS");  // the host's CPU share, where a scheduler sets it
    int t = e ? atoi(e) : (int)std::thread::hardware_concurrency();
    return std::max(1, std::min(t, 64));
}

template <class F>
void pfor(uint64_t n, F&& f) {
    int nt = nthreads();
    std::vector<std::thread> th;
    for (int t = 0; t < nt; ++t)
        th.emplace_back([&, t] {
            for (uint64_t i = n * t / nt; i < n * (t + 1) / nt; ++i) f(i);
        });
    for (auto& x : th) x.join();
}

}  // namespace

extern "C" {

void kt_shard_host_free(kt_shard_host* s) {
    if (!s) return;
    free(s->offsets);
    free(s->heads);
    free(s->ghost_ids);
    free(s->ghost_degree);
    free(s->msg_src);
    free(s->msg_offsets);
    free(s->msg_heads);
    memset(s, 0, sizeof(*s));
}

int kt_shard_make(uint32_t n, const uint64_t* offsets, const uint32_t* heads, uint32_t num_ranks,
                  const uint64_t* bounds, uint32_t rank, uint32_t algorithm, kt_shard_host* out) {
    if (!out || !offsets || (n && !heads) || !bounds || num_ranks == 0 || rank >= num_ranks) return -1;
    memset(out, 0, sizeof(*out));
    const int P = (int)num_ranks;
    if (bounds[0] != 0 || bounds[P] != n) return -1;
    for (int r = 0; r < P; ++r)
        if (bounds[r + 1] < bounds[r]) return -1;
    const uint32_t from = (uint32_t)bounds[rank], to = (uint32_t)bounds[rank + 1];
    const uint32_t nl = to - from;
    out->rank = rank;
    out->num_ranks = num_ranks;
    out->n_local = nl;
    out->algorithm = algorithm;
    // local CSR slice
    const uint64_t base = offsets[from], nnz_l = offsets[to] - base;
    out->offsets = (uint64_t*)malloc(((size_t)nl + 1) * 8);
    out->heads = (uint32_t*)malloc((nnz_l ? nnz_l : 1) * 4);
    if (!out->offsets || !out->heads) {
        kt_shard_host_free(out);
        return -2;
    }
    for (uint32_t i = 0; i <= nl; ++i) out->offsets[i] = offsets[from + i] - base;
    memcpy(out->heads, heads + base, nnz_l * 4);
    // ghosts (find_ghosts, distributed_graph.h:1287-1298) with global degrees
    std::vector<uint32_t> gh;
    gh.reserve(nnz_l / 4 + 16);
    for (uint64_t e = 0; e < nnz_l; ++e) {
        uint32_t u = out->heads[e];
        if (u < from || u >= to) gh.push_back(u);
    }
    std::sort(gh.begin(), gh.end());
    gh.erase(std::unique(gh.begin(), gh.end()), gh.end());
    out->n_ghost = (uint32_t)gh.size();
    out->ghost_ids = (uint32_t*)malloc((gh.size() ? gh.size() : 1) * 4);
    out->ghost_degree = (uint64_t*)malloc((gh.size() ? gh.size() : 1) * 8);
    if (!out->ghost_ids || !out->ghost_degree) {
        kt_shard_host_free(out);
        return -2;
    }
    memcpy(out->ghost_ids, gh.data(), gh.size() * 4);
    for (size_t i = 0; i < gh.size(); ++i) out->ghost_degree[i] = offsets[gh[i] + 1] - offsets[gh[i]];
    // senders: ghosts a with an out-edge a -> u to an owned u
    std::vector<uint8_t> sends(gh.size(), 0);
    pfor(gh.size(), [&](uint64_t i) {
        uint32_t a = gh[i];
        for (uint64_t e = offsets[a]; e < offsets[a + 1]; ++e) {
            uint32_t x = heads[e];
            if (x >= from && x < to && deg_less(offsets, a, x)) {
                sends[i] = 1;
                break;
            }
        }
    });
    std::vector<uint32_t> src;
    for (size_t i = 0; i < gh.size(); ++i)
        if (sends[i]) src.push_back(gh[i]);
    // message payloads: cetric-x N+cut(a) (out-neighbours not owned by
    // owner(a)); patric N+(a) (all out-neighbours); ascending ids
    std::vector<uint64_t> len(src.size() + 1, 0);
    pfor(src.size(), [&](uint64_t i) {
        uint32_t a = src[i];
        int oa = owner_of(a, bounds, P);
        uint64_t c = 0;
        for (uint64_t e = offsets[a]; e < offsets[a + 1]; ++e) {
            uint32_t x = heads[e];
            if (deg_less(offsets, a, x) && (algorithm != 0 || owner_of(x, bounds, P) != oa)) ++c;
        }
        len[i + 1] = c;
    });
    for (size_t i = 0; i < src.size(); ++i) len[i + 1] += len[i];
    out->n_msg = (uint32_t)src.size();
    out->msg_src = (uint32_t*)malloc((src.size() ? src.size() : 1) * 4);
    out->msg_offsets = (uint64_t*)malloc((src.size() + 1) * 8);
    out->msg_heads = (uint32_t*)malloc((len.back() ? len.back() : 1) * 4);
    if (!out->msg_src || !out->msg_offsets || !out->msg_heads) {
        kt_shard_host_free(out);
        return -2;
    }
    memcpy(out->msg_src, src.data(), src.size() * 4);
    memcpy(out->msg_offsets, len.data(), len.size() * 8);
    pfor(src.size(), [&](uint64_t i) {
        uint32_t a = src[i];
        int oa = owner_of(a, bounds, P);
        uint64_t p = len[i];
        for (uint64_t e = offsets[a]; e < offsets[a + 1]; ++e) {
            uint32_t x = heads[e];
            if (deg_less(offsets, a, x) && (algorithm != 0 || owner_of(x, bounds, P) != oa))
                out->msg_heads[p++] = x;
        }
        std::sort(out->msg_heads + len[i], out->msg_heads + len[i + 1]);
    });
    return 0;
}

}  // extern "C"
