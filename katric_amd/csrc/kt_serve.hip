// Serving loop: continuous batching of count requests on one GPU.
//
// A worker thread owns two device slots (own stream, own grow-only arena,
// pinned result buffers).  Requests are admitted FIFO into the next batch
// while the batch fits the slot's share of the HBM budget and the batch
// limits; the batch's CSRs are uploaded into the slot (device-resident for
// the batch's lifetime) and rebased into one disjoint-union CSR, then counted
// by the batch pipeline.  While one slot computes, the other uploads and
// launches the next batch, so requests arriving during a batch are admitted
// as soon as a slot frees up, and their work overlaps on the device.
//
// Every result equals the one-graph count: a batch whose validation fails is
// re-run one request per batch so that only the malformed request fails, and
// a batch with unsorted lists is re-run with its oriented lists sorted on the
// device.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/katric_serve.h"
#include "kt_ctx.h"
#include "kt_device.h"

namespace kt {

static double now_ms() {
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

// One request of a device batch: where its CSR was copied to in the batch
// arena, its place in the batch, and its first work chunk.
struct GatherDesc {
    const uint64_t* off;
    const uint32_t* heads;
    uint64_t n, nnz;
    uint64_t vb, eb;  // vertex / edge base in the batch
    uint64_t c0;      // first chunk (offsets chunks, then heads chunks)
    uint64_t pad;
};
constexpr uint32_t GATHER_CHUNK = 4096;  // ids per chunk
constexpr size_t SERVE_LOOKAHEAD = 4096; // queued requests scanned when forming a batch

__host__ __device__ inline uint64_t gather_chunks(uint64_t n, uint64_t nnz) {
    return (n + GATHER_CHUNK - 1) / GATHER_CHUNK + (nnz + GATHER_CHUNK - 1) / GATHER_CHUNK;
}

// Rebases the copied CSRs in place into the batch's disjoint-union CSR in one
// launch: off[vb_g + i] = off_g[i] + eb_g, heads[eb_g + k] = heads_g[k] + vb_g
// (a head outside its own graph becomes an invalid id, flagged by the
// orientation pass), off[N] = nnz.  Blocks take chunks of one request's
// offsets or heads (every element is read and written by the same thread).
__global__ void __launch_bounds__(256) k_serve_gather(uint32_t G, const GatherDesc* __restrict__ d,
                                                      uint64_t chunks, uint64_t* __restrict__ off,
                                                      uint32_t* __restrict__ heads, uint64_t N, uint64_t nnz) {
    for (uint64_t c = blockIdx.x; c < chunks; c += gridDim.x) {
        uint32_t lo = 0, hi = G - 1;  // last request with c0 <= c
        while (lo < hi) {
            const uint32_t mid = (lo + hi + 1) / 2;
            if (d[mid].c0 <= c)
                lo = mid;
            else
                hi = mid - 1;
        }
        const GatherDesc r = d[lo];
        const uint64_t oc = (r.n + GATHER_CHUNK - 1) / GATHER_CHUNK;
        const uint64_t k = c - r.c0;
        if (k < oc) {
            const uint64_t i0 = k * GATHER_CHUNK, i1 = min(r.n, i0 + GATHER_CHUNK);
            for (uint64_t i = i0 + threadIdx.x; i < i1; i += blockDim.x) off[r.vb + i] = r.off[i] + r.eb;
        } else {
            const uint64_t e0 = (k - oc) * GATHER_CHUNK, e1 = min(r.nnz, e0 + GATHER_CHUNK);
            for (uint64_t e = e0 + threadIdx.x; e < e1; e += blockDim.x) {
                const uint32_t h = r.heads[e];
                heads[r.eb + e] = (h < r.n) ? (uint32_t)(h + r.vb) : 0xffffffffu;
            }
        }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) off[N] = nnz;
}

struct ServeReq {
    uint64_t ticket;
    uint32_t n;
    uint64_t nnz;
    const uint64_t* off;
    const uint32_t* heads;
    uint64_t* pv_out;
    double t_submit = 0, t_launch = 0, t_done = 0;
    int status = KT_ENOTREADY;
    uint64_t total = 0;
    uint64_t batch_id = 0;
    uint32_t batch_graphs = 0;
    bool isolate = false;
    bool cancelled = false;
    bool done = false;
};

struct ServeSlot {
    hipStream_t stream = nullptr;
    hipEvent_t ev = nullptr;
    void* arena = nullptr;
    uint64_t arena_bytes = 0;
    uint64_t* h_small = nullptr;  // pinned, 64 words
    uint64_t* h_meta = nullptr;   // pinned: vbase (cap + 1), then the gather descriptors (cap)
    uint64_t* h_totals = nullptr; // pinned, cap
    uint64_t meta_cap = 0;
    std::vector<ServeReq*> reqs;
    bool busy = false;
    bool sorted = false;
    bool large = false;  // the batch is one large request
    uint64_t batch_id = 0;
    uint64_t N = 0, nnz = 0;
};

struct ServeState {
    kt_serve_config cfg{};
    int device = 0;
    int num_cus = 0;
    TcParams prm{};
    uint64_t sym_key[2] = {0, 0};
    uint64_t slot_cap = 0;
    uint64_t large_edges = 0;  // requests above this many edges run alone
    std::thread worker;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<ServeReq*> pending;
    std::unordered_map<uint64_t, ServeReq*> all;
    uint64_t next_ticket = 1;
    uint64_t next_batch = 1;
    bool stop = false;
    ServeSlot slots[2];
    kt_serve_stats st{};
    char err[256] = {0};
};

// Device bytes of a slot's batch: the batch pipeline's buffers (with the
// radix-sort scratch for a relaunch of unsorted lists that needs it) and the
// gather descriptors after them.
static uint64_t slot_bytes(uint32_t G, uint64_t N, uint64_t nnz, bool radix) {
    BatchBufs B;
    return ((batch_layout(nullptr, G, N, nnz, true, true, radix, B) + 255) & ~255ull) + (uint64_t)G * sizeof(GatherDesc);
}

static uint64_t req_bytes(const ServeReq* r) { return slot_bytes(1, r->n, r->nnz, false); }

static int slot_grow(ServeSlot& s, uint64_t bytes, uint32_t G) {
    if (bytes > s.arena_bytes) {
        if (s.arena) (void)hipFree(s.arena);
        s.arena = nullptr;
        s.arena_bytes = 0;
        if (hipMalloc(&s.arena, bytes) != hipSuccess) {
            (void)hipGetLastError();
            s.arena = nullptr;
            return KT_ENOMEM;
        }
        s.arena_bytes = bytes;
    }
    if (G + 1 > s.meta_cap) {
        if (s.h_meta) (void)hipHostFree(s.h_meta);
        if (s.h_totals) (void)hipHostFree(s.h_totals);
        s.h_meta = s.h_totals = nullptr;
        s.meta_cap = 0;
        uint64_t cap = std::max<uint64_t>(2 * (uint64_t)G + 2, 64);
        if (hipHostMalloc(reinterpret_cast<void**>(&s.h_meta), (cap + 1) * 8 + cap * sizeof(GatherDesc),
                          hipHostMallocDefault) != hipSuccess ||
            hipHostMalloc(reinterpret_cast<void**>(&s.h_totals), cap * 8, hipHostMallocDefault) != hipSuccess) {
            (void)hipGetLastError();
            return KT_ENOMEM;
        }
        s.meta_cap = cap;
    }
    return KT_OK;
}

// Uploads, rebases and enqueues one batch on the slot (worker thread, no lock).
static int launch(ServeState* S, ServeSlot& sl, uint32_t sort) {
    const uint32_t G = (uint32_t)sl.reqs.size();
    uint64_t N = 0, nnz = 0;
    for (auto* r : sl.reqs) {
        N += r->n;
        nnz += r->nnz;
    }
    BatchBufs B;
    const bool radix = sort == SORT_RADIX;
    const uint64_t used = (batch_layout(nullptr, G, N, nnz, true, true, radix, B) + 255) & ~255ull;
    // a slot never holds more than its share of the budget (a relaunch with
    // the radix-sort scratch can need more than admission counted)
    if (slot_bytes(G, N, nnz, radix) > S->slot_cap) return KT_ENOMEM;
    int rc = slot_grow(sl, used + (uint64_t)G * sizeof(GatherDesc), G);
    if (rc) return rc;
    batch_layout(sl.arena, G, N, nnz, true, true, radix, B);
    GatherDesc* d_desc = reinterpret_cast<GatherDesc*>(static_cast<char*>(sl.arena) + used);
    uint64_t* vb = sl.h_meta;
    GatherDesc* desc = reinterpret_cast<GatherDesc*>(sl.h_meta + (sl.meta_cap + 1));
    hipStream_t st = sl.stream;
    uint64_t* d_off = const_cast<uint64_t*>(B.off);
    uint32_t* d_heads = const_cast<uint32_t*>(B.heads);
    const double t = now_ms();
    vb[0] = 0;
    uint64_t e = 0, chunks = 0;
    for (uint32_t g = 0; g < G; ++g) {
        ServeReq* r = sl.reqs[g];
        r->t_launch = t;
        GatherDesc& dg = desc[g];
        dg.n = r->n;
        dg.nnz = r->nnz;
        dg.vb = vb[g];
        dg.eb = e;
        dg.c0 = chunks;
        // the request's arrays are copied into their place in the batch
        // (copy engines: from page-locked memory these run beside the other
        // slot's kernels) and rebased in place by k_serve_gather
        if (r->n && hipMemcpyAsync(d_off + vb[g], r->off, (uint64_t)r->n * 8, hipMemcpyHostToDevice, st) != hipSuccess)
            return KT_EHIP;
        if (r->nnz && hipMemcpyAsync(d_heads + e, r->heads, r->nnz * 4, hipMemcpyHostToDevice, st) != hipSuccess)
            return KT_EHIP;
        dg.off = d_off + vb[g];
        dg.heads = d_heads + e;
        chunks += gather_chunks(r->n, r->nnz);
        vb[g + 1] = vb[g] + r->n;
        e += r->nnz;
    }
    if (hipMemcpyAsync(B.vbase, vb, ((uint64_t)G + 1) * 8, hipMemcpyHostToDevice, st) != hipSuccess) return KT_EHIP;
    if (hipMemcpyAsync(d_desc, desc, (uint64_t)G * sizeof(GatherDesc), hipMemcpyHostToDevice, st) != hipSuccess)
        return KT_EHIP;
    const uint32_t grid = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(chunks, (uint64_t)S->num_cus * 4));
    k_serve_gather<<<grid, 256, 0, st>>>(G, d_desc, chunks, d_off, d_heads, N, nnz);
    if (batch_pipeline(st, S->num_cus, S->prm, B, G, N, nnz, B.pv, sl.h_totals, sl.h_small, sort, S->sym_key) !=
        hipSuccess)
        return KT_EHIP;
    for (uint32_t g = 0; g < G; ++g) {
        ServeReq* r = sl.reqs[g];
        if (r->pv_out && r->n &&
            hipMemcpyAsync(r->pv_out, B.pv + vb[g], (uint64_t)r->n * 8, hipMemcpyDeviceToHost, st) != hipSuccess)
            return KT_EHIP;
    }
    if (hipEventRecord(sl.ev, st) != hipSuccess) return KT_EHIP;
    sl.N = N;
    sl.nnz = nnz;
    sl.sorted = sort != SORT_NONE;
    sl.busy = true;
    return KT_OK;
}

// Completion of a slot's batch (worker thread, lock held by the caller).
static void finish(ServeState* S, ServeSlot& sl) {
    const uint32_t fl = (uint32_t)sl.h_small[0];
    const uint64_t kept = sl.h_small[1];
    const uint32_t G = (uint32_t)sl.reqs.size();
    const bool bad = (fl & KT_ERR_MASK) || 2 * kept != sl.nnz;
    if (bad && G > 1) {
        // isolate: every request of the batch runs alone, at the queue front
        for (int g = (int)G - 1; g >= 0; --g) {
            sl.reqs[g]->isolate = true;
            S->pending.push_front(sl.reqs[g]);
        }
        S->st.isolation_reruns++;
    } else {
        const double t = now_ms();
        for (uint32_t g = 0; g < G; ++g) {
            ServeReq* r = sl.reqs[g];
            r->status = bad ? KT_EINVAL : (r->cancelled ? KT_ECANCELED : KT_OK);
            r->total = bad ? 0 : sl.h_totals[g];
            r->t_done = t;
            r->batch_id = sl.batch_id;
            r->batch_graphs = G;
            r->done = true;
        }
        if (!bad) {
            S->st.batches++;
            S->st.graphs += G;
            S->st.edges += sl.nnz / 2;
        }
    }
    sl.reqs.clear();
    sl.busy = false;
    S->cv_done.notify_all();
}

// A batch that could not be launched (lock held by the caller): a batch of
// several requests that ran out of device memory is re-run one request at a
// time, so that only a request too large on its own fails; any other error
// fails the batch's requests.
static void launch_failed(ServeState* S, ServeSlot& sl, int rc) {
    const uint32_t G = (uint32_t)sl.reqs.size();
    if (rc == KT_ENOMEM && G > 1) {
        for (int g = (int)G - 1; g >= 0; --g) {
            sl.reqs[g]->isolate = true;
            S->pending.push_front(sl.reqs[g]);
        }
        S->st.isolation_reruns++;
    } else {
        for (auto* r : sl.reqs) {
            r->status = rc;
            r->done = true;
        }
    }
    sl.reqs.clear();
    sl.busy = false;
    S->cv_done.notify_all();
}

static void worker_main(ServeState* S) {
    (void)hipSetDevice(S->device);
    std::unique_lock<std::mutex> lk(S->mu);
    while (true) {
        // publish finished batches
        for (auto& sl : S->slots) {
            if (!sl.busy) continue;
            if (hipEventQuery(sl.ev) != hipSuccess) continue;
            (void)hipGetLastError();
            const uint32_t fl = (uint32_t)sl.h_small[0];
            if (!(fl & KT_ERR_MASK) && (fl & KT_FLAG_UNSORTED) && !sl.sorted && 2 * sl.h_small[1] == sl.nnz) {
                lk.unlock();
                // same requests, with the oriented lists sorted (the sort
                // chosen from the longest list of the run that found them)
                BatchBufs B0;
                batch_layout(sl.arena, (uint32_t)sl.reqs.size(), sl.N, sl.nnz, true, true, false, B0);
                uint32_t mode = SORT_RADIX;
                int rc = batch_sort_mode(sl.stream, B0, sl.N, sl.h_small, mode) == hipSuccess ? KT_OK : KT_EHIP;
                if (rc == KT_OK) rc = launch(S, sl, mode);
                lk.lock();
                if (rc != KT_OK) launch_failed(S, sl, rc);
                continue;
            }
            finish(S, sl);
        }
        ServeSlot* free_slot = nullptr;
        for (auto& sl : S->slots)
            if (!sl.busy) {
                free_slot = &sl;
                break;
            }
        // drop cancelled requests at the queue front
        while (!S->pending.empty() && S->pending.front()->cancelled) {
            ServeReq* r = S->pending.front();
            S->pending.pop_front();
            r->status = KT_ECANCELED;
            r->done = true;
            S->cv_done.notify_all();
        }
        if (free_slot && !S->pending.empty()) {
            // Form the next batch.  A large request (more than large_edges
            // edges) or an isolated one runs alone, and at most one slot runs a
            // large request at a time: while the other slot holds one, a large
            // request at the queue front waits and the small requests behind
            // it form this slot's batch, so that they are not held up behind
            // it.  Small requests are taken in arrival order while the batch
            // fits the slot and the limits.
            ServeSlot& other = (free_slot == &S->slots[0]) ? S->slots[1] : S->slots[0];
            const bool other_large = other.busy && other.large;
            std::vector<ServeReq*> batch;
            bool blocked_large = false;
            uint64_t N = 0, nnz = 0;
            for (size_t i = 0; i < S->pending.size() && i < SERVE_LOOKAHEAD;) {
                ServeReq* r = S->pending[i];
                if (r->cancelled) {
                    S->pending.erase(S->pending.begin() + (ptrdiff_t)i);
                    r->status = KT_ECANCELED;
                    r->done = true;
                    S->cv_done.notify_all();
                    continue;
                }
                const bool large = r->nnz / 2 > S->large_edges;
                if (large || r->isolate) {
                    if (batch.empty() && !(large && other_large)) {
                        S->pending.erase(S->pending.begin() + (ptrdiff_t)i);
                        if (req_bytes(r) > S->slot_cap) {
                            r->status = KT_ENOMEM;
                            r->done = true;
                            S->cv_done.notify_all();
                            continue;
                        }
                        batch.push_back(r);
                        blocked_large = large;
                        break;
                    }
                    ++i;  // keeps its place; smaller requests behind it go ahead
                    continue;
                }
                const uint64_t bytes = slot_bytes((uint32_t)batch.size() + 1, N + r->n, nnz + r->nnz, false);
                const bool limits = batch.size() + 1 <= S->cfg.max_batch_graphs &&
                                    (nnz + r->nnz) / 2 <= S->cfg.max_batch_edges;
                if (!batch.empty() && (bytes > S->slot_cap || !limits)) break;
                S->pending.erase(S->pending.begin() + (ptrdiff_t)i);
                if (bytes > S->slot_cap) {  // alone it does not fit the budget share
                    r->status = KT_ENOMEM;
                    r->done = true;
                    S->cv_done.notify_all();
                    continue;
                }
                batch.push_back(r);
                N += r->n;
                nnz += r->nnz;
            }
            if (!batch.empty()) {
                free_slot->reqs = batch;
                free_slot->large = blocked_large;
                free_slot->batch_id = S->next_batch++;
                lk.unlock();
                int rc = launch(S, *free_slot, SORT_NONE);
                lk.lock();
                if (rc != KT_OK) launch_failed(S, *free_slot, rc);
                continue;
            }
            // only a large request waits for the other slot's large one: poll below
        }
        bool any_busy = S->slots[0].busy || S->slots[1].busy;
        if (S->stop && S->pending.empty() && !any_busy) break;
        if (!any_busy) {
            S->cv_work.wait(lk);
        } else if (free_slot) {
            // a slot is free but nothing is pending: wake on a submission or
            // poll the busy slot again shortly
            S->cv_work.wait_for(lk, std::chrono::microseconds(100));
        } else {
            // both slots busy: wait for the earliest-launched one without the lock
            ServeSlot* w = nullptr;
            for (auto& sl : S->slots)
                if (sl.busy && (!w || sl.batch_id < w->batch_id)) w = &sl;
            hipEvent_t ev = w->ev;
            lk.unlock();
            (void)hipEventSynchronize(ev);
            lk.lock();
        }
    }
}

}  // namespace kt

using namespace kt;

static int serve_stop_impl(kt_ctx* c) {
    ServeState* S = c->serve;
    if (!S) return KT_OK;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->stop = true;
    }
    S->cv_work.notify_all();
    if (S->worker.joinable()) S->worker.join();
    for (auto& sl : S->slots) {
        if (sl.stream) (void)hipStreamSynchronize(sl.stream);
        if (sl.arena) (void)hipFree(sl.arena);
        if (sl.h_small) (void)hipHostFree(sl.h_small);
        if (sl.h_meta) (void)hipHostFree(sl.h_meta);
        if (sl.h_totals) (void)hipHostFree(sl.h_totals);
        if (sl.ev) (void)hipEventDestroy(sl.ev);
        if (sl.stream) (void)hipStreamDestroy(sl.stream);
    }
    for (auto& kv : S->all) delete kv.second;
    delete S;
    c->serve = nullptr;
    return KT_OK;
}

void kt_serve_release(kt_ctx* c) { serve_stop_impl(c); }

extern "C" int kt_serve_start(kt_ctx* c, const kt_serve_config* cfg) {
    if (!c) return KT_EINVAL;
    std::lock_guard<std::mutex> lk(c->mu);
    int rc = c->begin_call(false);
    if (rc) return rc;
    if (c->serve) return c->fail(KT_ESTATE, "serving loop already running");
    ServeState* S = new ServeState();
    if (cfg) S->cfg = *cfg;
    if (S->cfg.max_batch_graphs == 0) S->cfg.max_batch_graphs = 4096;
    if (S->cfg.max_batch_edges == 0) S->cfg.max_batch_edges = 1ull << 25;
    S->large_edges = S->cfg.large_request_edges ? S->cfg.large_request_edges : S->cfg.max_batch_edges / 4;
    uint64_t budget = S->cfg.hbm_budget_bytes ? S->cfg.hbm_budget_bytes : c->budget;
    if (budget == 0) {
        size_t fr = 0, tot = 0;
        if (hipMemGetInfo(&fr, &tot) != hipSuccess) {
            (void)hipGetLastError();
            fr = 16ull << 30;
        }
        budget = fr / 2;
    }
    S->slot_cap = budget / 2;
    S->device = c->device;
    S->num_cus = c->num_cus;
    S->prm = c->prm;
    S->sym_key[0] = c->sym_key[0];
    S->sym_key[1] = c->sym_key[1];
    for (auto& sl : S->slots) {
        if (hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&sl.ev, hipEventDisableTiming) != hipSuccess ||
            hipHostMalloc(reinterpret_cast<void**>(&sl.h_small), 64 * 8, hipHostMallocDefault) != hipSuccess) {
            (void)hipGetLastError();
            c->serve = S;
            serve_stop_impl(c);
            return c->fail(KT_EHIP, "serving loop: stream/event/pinned allocation failed");
        }
    }
    c->serve = S;
    S->worker = std::thread(worker_main, S);
    trace("serve: started, slot budget %llu bytes", (unsigned long long)S->slot_cap);
    return KT_OK;
}

extern "C" int kt_serve_stop(kt_ctx* c) {
    if (!c) return KT_EINVAL;
    std::lock_guard<std::mutex> lk(c->mu);
    return serve_stop_impl(c);
}

extern "C" int kt_serve_submit(kt_ctx* c, uint32_t n, const uint64_t* offsets, const uint32_t* heads,
                               uint64_t* per_vertex_out, uint64_t* ticket) {
    if (!c || !ticket || !offsets) return KT_EINVAL;
    ServeState* S = c->serve;
    if (!S) return KT_ESTATE;
    if (offsets[0] != 0) return KT_EINVAL;
    const uint64_t nnz = offsets[n];
    if ((nnz & 1) || (nnz && !heads)) return KT_EINVAL;
    ServeReq* r = new ServeReq();
    r->n = n;
    r->nnz = nnz;
    r->off = offsets;
    r->heads = heads;
    r->pv_out = per_vertex_out;
    r->t_submit = now_ms();
    {
        std::lock_guard<std::mutex> lk(S->mu);
        if (S->stop) {
            delete r;
            return KT_ESTATE;
        }
        r->ticket = S->next_ticket++;
        S->all[r->ticket] = r;
        if (req_bytes(r) > S->slot_cap) {
            r->status = KT_ENOMEM;
            r->done = true;
        } else {
            S->pending.push_back(r);
        }
        S->st.submitted++;
        *ticket = r->ticket;
    }
    S->cv_work.notify_all();
    return KT_OK;
}

static void fill_result(const ServeReq* r, kt_serve_result* out) {
    out->ticket = r->ticket;
    out->status = r->status;
    out->total = r->total;
    out->queue_ms = r->t_launch > 0 ? r->t_launch - r->t_submit : 0;
    out->latency_ms = r->t_done - r->t_submit;
    out->batch_id = r->batch_id;
    out->batch_graphs = r->batch_graphs;
}

extern "C" int kt_serve_poll(kt_ctx* c, uint64_t ticket, kt_serve_result* out) {
    if (!c || !out) return KT_EINVAL;
    ServeState* S = c->serve;
    if (!S) return KT_ESTATE;
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = S->all.find(ticket);
    if (it == S->all.end()) return KT_ENOTFOUND;
    if (!it->second->done) return KT_ENOTREADY;
    fill_result(it->second, out);
    delete it->second;
    S->all.erase(it);
    return KT_OK;
}

extern "C" int kt_serve_wait(kt_ctx* c, uint64_t ticket, double timeout_ms, kt_serve_result* out) {
    if (!c || !out) return KT_EINVAL;
    ServeState* S = c->serve;
    if (!S) return KT_ESTATE;
    std::unique_lock<std::mutex> lk(S->mu);
    auto it = S->all.find(ticket);
    if (it == S->all.end()) return KT_ENOTFOUND;
    ServeReq* r = it->second;
    auto pred = [&] { return r->done; };
    if (timeout_ms < 0)
        S->cv_done.wait(lk, pred);
    else if (!S->cv_done.wait_for(lk, std::chrono::duration<double, std::milli>(timeout_ms), pred))
        return KT_ENOTREADY;
    fill_result(r, out);
    delete r;
    S->all.erase(ticket);
    return KT_OK;
}

extern "C" int kt_serve_cancel(kt_ctx* c, uint64_t ticket) {
    if (!c) return KT_EINVAL;
    ServeState* S = c->serve;
    if (!S) return KT_ESTATE;
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = S->all.find(ticket);
    if (it == S->all.end()) return KT_ENOTFOUND;
    if (it->second->done) return KT_ESTATE;
    it->second->cancelled = true;
    S->cv_work.notify_all();
    return KT_OK;
}

extern "C" int kt_serve_get_stats(kt_ctx* c, kt_serve_stats* out) {
    if (!c || !out) return KT_EINVAL;
    ServeState* S = c->serve;
    if (!S) return KT_ESTATE;
    std::lock_guard<std::mutex> lk(S->mu);
    *out = S->st;
    out->pending = S->pending.size();
    return KT_OK;
}
