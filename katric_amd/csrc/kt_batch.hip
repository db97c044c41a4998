// Batched counting: thousands of independent graphs in one pipeline.
//
// The batch is one disjoint-union CSR (global ids, graph g owning
// [vertex_base[g], vertex_base[g+1])).  Orientation, binning and the
// persistent intersection kernel run once over the union -- every kernel of
// the plain path, unchanged -- and per-graph totals are a segmented sum of the
// per-vertex counts divided by 3 (each triangle adds 1 to each of its three
// vertices, all inside one graph).
//
// The pipeline has no host synchronisation between its kernels: the oriented
// heads buffer holds nnz entries (an asymmetric input can therefore never
// overflow it), validation flags are read after the run, and an input found
// unsorted is redone on the eager path with its oriented lists sorted on the
// device (in place per list, or by the radix sort when a list is longer than
// the in-place sort takes).  That makes
// the whole device side capturable: it is recorded once per batch shape into
// a HIP graph and replayed on later calls with the same shape and buffers.
#include <algorithm>
#include <cstring>

#include "../../include/katric.h"
#include "kt_ctx.h"
#include "kt_device.h"

namespace kt {

struct BatchState {
    hipGraphExec_t exec = nullptr;
    // capture key
    uint32_t G = 0;
    uint64_t N = 0, nnz = 0;
    const void* off = nullptr;
    const void* heads = nullptr;
    const void* pv = nullptr;
    void* arena = nullptr;
    uint64_t arena_bytes = 0;
    TcParams prm{};
    bool host_in = false;
    // pinned per-graph totals
    uint64_t* h_totals = nullptr;
    uint64_t h_cap = 0;
    uint64_t replays = 0;  // launches of captured pipelines (cumulative)
};

// Per-graph totals = (sum of the graph's per-vertex counts) / 3, as a
// segmented sum that does not depend on how the vertices split into graphs:
// each wave sums a chunk of SEG_CHUNK consecutive vertices, one overlap with a
// graph at a time (the first graph found by binary search in vbase), and adds
// each non-zero partial sum to acc[g] (zeroed before) with one atomic.
constexpr uint32_t SEG_CHUNK = 512;

__global__ void k_seg_partial(uint32_t G, uint64_t N, const uint64_t* __restrict__ vbase,
                              const uint64_t* __restrict__ pv, uint64_t* __restrict__ acc) {
    const uint64_t wid = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) / KT_WAVE;
    const uint64_t nw = (uint64_t)gridDim.x * blockDim.x / KT_WAVE;
    for (uint64_t c0 = wid * SEG_CHUNK; c0 < N; c0 += nw * SEG_CHUNK) {
        const uint64_t c1 = min(c0 + SEG_CHUNK, N);
        uint32_t lo = 0, hi = G - 1;  // last graph g with vbase[g] <= c0
        while (lo < hi) {
            const uint32_t mid = lo + (hi - lo + 1) / 2;
            if (vbase[mid] <= c0)
                lo = mid;
            else
                hi = mid - 1;
        }
        for (uint32_t g = lo; g < G; ++g) {
            const uint64_t a = max(vbase[g], c0), b = min(vbase[g + 1], c1);
            if (a >= c1) break;
            uint64_t s = 0;
            for (uint64_t v = a + lane_id(); v < b; v += KT_WAVE) s += pv[v];
            s = wave_sum(s);
            if (lane_id() == 0 && s) atomic_add_u64(&acc[g], s);
        }
    }
}

__global__ void k_seg_final(uint32_t G, uint64_t* __restrict__ totals, uint32_t* __restrict__ err) {
    for (uint32_t g = blockIdx.x * blockDim.x + threadIdx.x; g < G; g += gridDim.x * blockDim.x) {
        const uint64_t t = totals[g];
        if (t % 3) atomicOr(err, KT_ERR_SEGMENT);
        totals[g] = t / 3;
    }
}

uint64_t batch_layout(void* base, uint32_t G, uint64_t N, uint64_t nnz, bool host_in, bool own_pv, bool unsorted,
                      BatchBufs& B) {
    Carver cv{base, 0};
    B.small = static_cast<uint64_t*>(cv.take(64 * 8));
    B.off = nullptr;
    B.heads = nullptr;
    if (host_in) {
        B.off = static_cast<uint64_t*>(cv.take((N + 1) * 8));
        B.heads = static_cast<uint32_t*>(cv.take(std::max<uint64_t>(nnz, 1) * 4));
    }
    B.vbase = static_cast<uint64_t*>(cv.take(((uint64_t)G + 1) * 8));
    B.deg = static_cast<uint32_t*>(cv.take(std::max<uint64_t>(N, 1) * 4));
    B.ooff = static_cast<uint64_t*>(cv.take((N + 1) * 8));
    B.oheads = static_cast<uint32_t*>(cv.take((nnz + 4) * 4));  // capacity nnz: no overflow without a host check
    B.ws_orient = static_cast<uint64_t*>(cv.take(orient_ws_words((uint32_t)N, nnz) * 8));
    B.ws_tc = static_cast<uint64_t*>(cv.take(tc_ws_words((uint32_t)N) * 8));
    B.pv = own_pv ? static_cast<uint64_t*>(cv.take(std::max<uint64_t>(N, 1) * 8)) : nullptr;
    B.totals = static_cast<uint64_t*>(cv.take(std::max<uint64_t>(G, 1) * 8));
    B.keys = B.alt = B.ws_radix = nullptr;
    if (unsorted) {
        const uint64_t m = nnz / 2;
        B.keys = static_cast<uint64_t*>(cv.take(std::max<uint64_t>(m, 1) * 8));
        B.alt = static_cast<uint64_t*>(cv.take(std::max<uint64_t>(m, 1) * 8));
        B.ws_radix = static_cast<uint64_t*>(cv.take(radix_tmp_words(m) * 8));
    }
    return cv.used;
}

static int bits_for(uint64_t x) {
    int b = 0;
    while (b < 64 && (x >> b)) ++b;
    return b;
}

hipError_t batch_sort_mode(hipStream_t st, const BatchBufs& B, uint64_t N, uint64_t* h_small, uint32_t& mode) {
    // ws_tc (tc_ws_words >= its queue header) holds the reduction's partials
    max_len_launch((uint32_t)N, B.ooff, B.ws_tc, B.small + 8, st);
    hipError_t e = hipMemcpyAsync(h_small + 8, B.small + 8, 8, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return e;
    mode = h_small[8] <= segsort_max_len() ? SORT_SEGMENTED : SORT_RADIX;
    return hipSuccess;
}

// Enqueues the device pipeline on `st` (no host synchronisation): results land
// in h_totals (per graph) and h_small ([0] flags, [1] kept edges, [2..4] stats).
hipError_t batch_pipeline(hipStream_t st, int num_cus, const TcParams& prm, const BatchBufs& B, uint32_t G, uint64_t N,
                          uint64_t nnz, uint64_t* pv, uint64_t* h_totals, uint64_t* h_small, uint32_t sort,
                          const uint64_t sym_key[2]) {
    uint32_t* err = reinterpret_cast<uint32_t*>(B.small);
    fill_u64(B.small, 64, 0, st);
    // heads must stay inside their row's graph: a union with edges between
    // graphs is not a batch of independent graphs
    OrientCheck chk{{sym_key[0], sym_key[1]}, B.vbase, G};
    orient_count_launch((uint32_t)N, nnz, B.off, B.heads, B.deg, B.ooff, B.ws_orient, err, chk, st);
    orient_write_launch((uint32_t)N, nnz, B.heads, B.oheads, B.ws_orient, st);
    if (sort == SORT_SEGMENTED) {
        segsort_launch((uint32_t)N, B.ooff, B.oheads, err, segsort_max_len(), num_cus, st);
    } else if (sort == SORT_RADIX) {
        const uint64_t m = nnz / 2;
        make_keys((uint32_t)N, m, B.ooff, B.oheads, B.keys, st);
        radix_sort_u64(B.keys, B.alt, m, 32 + bits_for(N), B.ws_radix, st);
        keys_to_heads(m, B.keys, B.oheads, err, st);
    }
    fill_u64(pv, N, 0, st);
    TcView v{};
    v.X = (uint32_t)N;
    v.s_off = B.ooff;
    v.sheads = B.oheads;
    if (prm.min_side) {
        v.in_off = B.off;
        v.in_heads = B.heads;
        v.in_keep = orient_keep_bits((uint32_t)N, nnz, B.ws_orient);
    }
    TcOut o{pv, B.small + 2};
    const hipError_t le = tc_launch(v, o, prm, B.ws_tc, num_cus, st, nullptr, nullptr);
    if (le != hipSuccess) return le;
    fill_u64(B.totals, G, 0, st);
    k_seg_partial<<<(unsigned)std::clamp<uint64_t>((N + SEG_CHUNK * 4 - 1) / (SEG_CHUNK * 4), 1, 16384), 256, 0, st>>>(
        G, N, B.vbase, pv, B.totals);
    k_seg_final<<<std::min<uint32_t>((G + 255) / 256, 1024), 256, 0, st>>>(G, B.totals, err);
    (void)hipMemcpyAsync(h_totals, B.totals, (uint64_t)G * 8, hipMemcpyDeviceToHost, st);
    (void)hipMemcpyAsync(h_small, B.small, 8, hipMemcpyDeviceToHost, st);
    (void)hipMemcpyAsync(h_small + 1, B.ooff + N, 8, hipMemcpyDeviceToHost, st);
    (void)hipMemcpyAsync(h_small + 2, tc_stats_ptr(B.ws_tc), 32, hipMemcpyDeviceToHost, st);
    return hipGetLastError();
}

}  // namespace kt

using namespace kt;

void kt_batch_release(kt_ctx* c) {
    if (!c->batch) return;
    if (c->batch->exec) (void)hipGraphExecDestroy(c->batch->exec);
    if (c->batch->h_totals) (void)hipHostFree(c->batch->h_totals);
    delete c->batch;
    c->batch = nullptr;
}

extern "C" int kt_batch_count(kt_ctx* c, uint32_t G, const uint64_t* vertex_base, const uint64_t* offsets,
                              const uint32_t* heads, uint32_t flags, uint64_t* totals, uint64_t* per_vertex) {
    if (!c) return KT_EINVAL;
    std::lock_guard<std::mutex> lk(c->mu);
    int rc = c->begin_call(true);
    if (rc) return rc;
    if (!vertex_base || !offsets || !totals) return c->fail(KT_EINVAL, "null argument");
    if (G == 0) return KT_OK;
    if (vertex_base[0] != 0) return c->fail(KT_EINVAL, "vertex_base[0] != 0");
    for (uint32_t g = 0; g < G; ++g)
        if (vertex_base[g + 1] < vertex_base[g]) return c->fail(KT_EINVAL, "vertex_base not ascending");
    const uint64_t N = vertex_base[G];
    if (N >= 0xffffffffull) return c->fail(KT_EINVAL, "batch too large for uint32 ids");
    memset(totals, 0, (uint64_t)G * 8);
    if (N == 0) return KT_OK;
    const bool host_in = !(flags & KT_DEVICE_INPUT);
    const bool want_pv = per_vertex != nullptr;
    const bool pv_dev = want_pv && (flags & KT_DEVICE_OUTPUT);
    uint64_t nnz = 0;
    if (host_in) {
        if (offsets[0] != 0) return c->fail(KT_EINVAL, "offsets[0] != 0");
        nnz = offsets[N];
        if (nnz && !heads) return c->fail(KT_EINVAL, "null heads");
    } else {
        rc = c->read_u64(offsets + N, &nnz, "D2H nnz");
        if (rc) return rc;
    }
    if (nnz & 1) return c->fail(KT_EINVAL, "odd number of CSR entries: not a symmetric graph");
    if (!c->batch) c->batch = new BatchState();
    BatchState* S = c->batch;
    if (S->h_cap < G) {
        if (S->h_totals) (void)hipHostFree(S->h_totals);
        S->h_totals = nullptr;
        S->h_cap = 0;
        rc = c->hip_check(hipHostMalloc(reinterpret_cast<void**>(&S->h_totals), (uint64_t)G * 8, hipHostMallocDefault),
                          "hipHostMalloc totals");
        if (rc) return rc;
        S->h_cap = G;
    }
    hipStream_t st = c->stream;
    uint32_t sort_mode = SORT_NONE;
    for (int attempt = 0; attempt < 2; ++attempt) {
        const bool sort = attempt == 1;
        const bool radix = sort_mode == SORT_RADIX;
        BatchBufs B;
        const uint64_t need = batch_layout(nullptr, G, N, nnz, host_in, !pv_dev, radix, B);
        rc = c->reserve(need);
        if (rc) return rc;
        batch_layout(c->arena, G, N, nnz, host_in, !pv_dev, radix, B);
        if (!host_in) {
            B.off = offsets;
            B.heads = heads;
        }
        uint64_t* pv = pv_dev ? per_vertex : B.pv;
        (void)hipEventRecord(c->ev[EV_START], st);
        if (host_in) {
            rc = c->hip_check(hipMemcpyAsync(const_cast<uint64_t*>(B.off), offsets, (N + 1) * 8, hipMemcpyHostToDevice,
                                             st),
                              "H2D offsets");
            if (!rc && nnz)
                rc = c->hip_check(
                    hipMemcpyAsync(const_cast<uint32_t*>(B.heads), heads, nnz * 4, hipMemcpyHostToDevice, st),
                    "H2D heads");
            if (rc) return rc;
        }
        rc = c->hip_check(hipMemcpyAsync(B.vbase, vertex_base, ((uint64_t)G + 1) * 8, hipMemcpyHostToDevice, st),
                          "H2D vertex_base");
        if (rc) return rc;
        (void)hipEventRecord(c->ev[EV_H2D], st);
        if (!sort) {
            const bool same = S->exec && S->G == G && S->N == N && S->nnz == nnz && S->off == B.off &&
                              S->heads == B.heads && S->pv == pv && S->arena == c->arena &&
                              S->arena_bytes == c->arena_bytes && S->host_in == host_in &&
                              memcmp(&S->prm, &c->prm, sizeof(TcParams)) == 0;
            if (!same) {
                if (S->exec) (void)hipGraphExecDestroy(S->exec);
                S->exec = nullptr;
                hipGraph_t graph = nullptr;
                rc = c->hip_check(hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed), "begin capture");
                if (rc) return rc;
                const hipError_t le = batch_pipeline(st, c->num_cus, c->prm, B, G, N, nnz, pv, S->h_totals,
                                                     c->host_small, SORT_NONE, c->sym_key);
                rc = c->hip_check(hipStreamEndCapture(st, &graph), "end capture");
                if (!rc && le != hipSuccess) {
                    (void)hipGraphDestroy(graph);
                    rc = c->hip_check(le, "batch pipeline launch");
                }
                if (rc) return rc;
                rc = c->hip_check(hipGraphInstantiate(&S->exec, graph, nullptr, nullptr, 0), "instantiate");
                (void)hipGraphDestroy(graph);
                if (rc) return rc;
                S->G = G;
                S->N = N;
                S->nnz = nnz;
                S->off = B.off;
                S->heads = B.heads;
                S->pv = pv;
                S->arena = c->arena;
                S->arena_bytes = c->arena_bytes;
                S->host_in = host_in;
                S->prm = c->prm;
                trace("batch: captured pipeline for G=%u N=%llu nnz=%llu", G, (unsigned long long)N,
                      (unsigned long long)nnz);
            }
            (void)hipEventRecord(c->ev[EV_PREP], st);
            rc = c->hip_check(hipGraphLaunch(S->exec, st), "hipGraphLaunch");
            if (rc) return rc;
            S->replays++;
        } else {
            (void)hipEventRecord(c->ev[EV_PREP], st);
            rc = c->hip_check(batch_pipeline(st, c->num_cus, c->prm, B, G, N, nnz, pv, S->h_totals, c->host_small,
                                             sort_mode, c->sym_key),
                              "batch pipeline launch");
            if (rc) return rc;
        }
        (void)hipEventRecord(c->ev[EV_COUNT], st);
        if (want_pv && !pv_dev) {
            rc = c->hip_check(hipMemcpyAsync(per_vertex, pv, N * 8, hipMemcpyDeviceToHost, st), "D2H per-vertex");
            if (rc) return rc;
        }
        (void)hipEventRecord(c->ev[EV_END], st);
        rc = c->hip_check(hipStreamSynchronize(st), "batch");
        if (rc) return rc;
        const uint32_t fl = (uint32_t)c->host_small[0];
        const uint64_t kept = c->host_small[1];
        if ((fl & KT_ERR_MASK) || 2 * kept != nnz) {
            memset(totals, 0, (uint64_t)G * 8);
            return c->fail(KT_EINVAL, "malformed batch CSR (flags 0x%x, %llu oriented edges for %llu slots)", fl,
                           (unsigned long long)kept, (unsigned long long)nnz);
        }
        if ((fl & KT_FLAG_UNSORTED) && !sort) {
            // redo with the oriented lists sorted: in place, or by the radix
            // sort when a list is longer than the in-place sort takes
            rc = c->hip_check(batch_sort_mode(st, B, N, c->host_small, sort_mode), "longest list");
            if (rc) return rc;
            continue;
        }
        memcpy(totals, S->h_totals, (uint64_t)G * 8);
        uint64_t tot = 0;
        for (uint32_t g = 0; g < G; ++g) tot += totals[g];
        c->stats.local_triangles = tot;
        c->stats.wedge_checks = c->host_small[2];
        c->stats.intersection_size_local = c->host_small[3];
        c->stats.skipped_nodes = c->host_small[4];
        c->stats.streamed_ids = c->host_small[5];
        c->stats.oriented_edges = kept;
        c->stats.vertices = N;
        c->stats.input_slots = nnz;
        c->timing.h2d_ms = c->elapsed(EV_START, EV_H2D);
        c->timing.count_ms = c->elapsed(EV_PREP, EV_COUNT);
        c->timing.d2h_ms = c->elapsed(EV_COUNT, EV_END);
        c->timing.total_ms = c->elapsed(EV_START, EV_END);
        return KT_OK;
    }
    return c->fail(KT_EHIP, "batch: unreachable");
}

extern "C" uint64_t kt_batch_replays(const kt_ctx* c) { return (c && c->batch) ? c->batch->replays : 0; }
