// C ABI of katric_amd: context, device workspace, plain count / orient.
#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <new>

#include "../../include/katric.h"
#include "kt_ctx.h"

using namespace kt;

bool kt::trace_enabled() {
    static int on = -1;
    if (on < 0) {
        const char* e = getenv("KATRIC_TRACE");
        on = (e && e[0] && e[0] != '0') ? 1 : 0;
    }
    return on == 1;
}

void kt::trace(const char* fmt, ...) {
    if (!trace_enabled()) return;
    va_list ap;
    va_start(ap, fmt);
    fprintf(stderr, "[katric] ");
    vfprintf(stderr, fmt, ap);
    fputc('\n', stderr);
    fflush(stderr);
    va_end(ap);
}

// ---------------------------------------------------------------------------
// context helpers
// ---------------------------------------------------------------------------
int kt_ctx::fail(int status, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(err, sizeof(err), fmt, ap);
    va_end(ap);
    return status;
}

int kt_ctx::hip_check(hipError_t e, const char* what) {
    if (e == hipSuccess) return KT_OK;
    (void)hipGetLastError();
    return fail(KT_EHIP, "%s: %s", what, hipGetErrorString(e));
}

int kt_ctx::reserve(uint64_t bytes) {
    bytes = (bytes + 255) & ~255ull;
    if (bytes <= arena_bytes) return KT_OK;
    uint64_t want = std::max<uint64_t>(bytes, arena_bytes + arena_bytes / 4);
    if (budget && pinned_other + want > budget) want = bytes;
    if (budget && pinned_other + want > budget)
        return fail(KT_ENOMEM, "workspace of %llu bytes exceeds the HBM budget of %llu bytes",
                    (unsigned long long)want, (unsigned long long)budget);
    if (arena) {
        hipError_t e = hipStreamSynchronize(stream);
        if (e != hipSuccess) return hip_check(e, "hipStreamSynchronize");
        (void)hipFree(arena);
        arena = nullptr;
        arena_bytes = 0;
    }
    hipError_t e = hipMalloc(&arena, want);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        arena = nullptr;
        return fail(KT_ENOMEM, "hipMalloc(%llu) failed: %s", (unsigned long long)want, hipGetErrorString(e));
    }
    arena_bytes = want;
    return KT_OK;
}

void* Carver::take(uint64_t bytes) {
    uint64_t at = (used + 255) & ~255ull;
    used = at + ((bytes + 255) & ~255ull);
    return base ? static_cast<char*>(base) + at : nullptr;
}

int kt_ctx::begin_call(bool counting) {
    hipError_t e = hipSetDevice(device);
    if (e != hipSuccess) return hip_check(e, "hipSetDevice");
    err[0] = 0;
    if (counting) {
        memset(&stats, 0, sizeof(stats));
        memset(&timing, 0, sizeof(timing));
    }
    return KT_OK;
}

int kt_ctx::reserve_aux(int i, uint64_t bytes) {
    bytes = (std::max<uint64_t>(bytes, 256) + 255) & ~255ull;
    if (bytes <= aux_bytes[i]) return KT_OK;
    uint64_t want = std::max<uint64_t>(bytes, aux_bytes[i] + aux_bytes[i] / 4);
    const uint64_t others = arena_bytes + pinned_other - aux_bytes[i];
    if (budget && others + want > budget) want = bytes;
    if (budget && others + want > budget)
        return fail(KT_ENOMEM, "%llu bytes of working memory exceed the HBM budget of %llu bytes",
                    (unsigned long long)want, (unsigned long long)budget);
    if (aux[i]) {
        int rc = hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize");
        if (rc) return rc;
        (void)hipFree(aux[i]);
        pinned_other -= aux_bytes[i];
        aux[i] = nullptr;
        aux_bytes[i] = 0;
    }
    hipError_t e = hipMalloc(&aux[i], want);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        aux[i] = nullptr;
        return fail(KT_ENOMEM, "hipMalloc(%llu) failed", (unsigned long long)want);
    }
    aux_bytes[i] = want;
    pinned_other += want;
    return KT_OK;
}

void kt_ctx::release_aux() {
    for (int i = 0; i < 2; ++i) {
        if (aux[i]) {
            (void)hipFree(aux[i]);
            pinned_other -= aux_bytes[i];
        }
        aux[i] = nullptr;
        aux_bytes[i] = 0;
    }
}

int kt_ctx::read_u64(const uint64_t* dev, uint64_t* out, const char* what) {
    int rc = hip_check(hipMemcpyAsync(host_small + 63, dev, 8, hipMemcpyDeviceToHost, stream), what);
    if (!rc) rc = hip_check(hipStreamSynchronize(stream), what);
    if (!rc) *out = host_small[63];
    return rc;
}

float kt_ctx::elapsed(int a, int b) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, ev[a], ev[b]) != hipSuccess) {
        (void)hipGetLastError();
        return 0.f;
    }
    return ms;
}

// ---------------------------------------------------------------------------
// lifecycle
// ---------------------------------------------------------------------------
extern "C" int kt_create(kt_ctx** out, int device, uint64_t hbm_budget_bytes) {
    if (!out) return KT_EINVAL;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) {
        (void)hipGetLastError();
        return KT_ENODEV;
    }
    if (device < 0 || device >= count) return KT_ENODEV;
    if (hipSetDevice(device) != hipSuccess) {
        (void)hipGetLastError();
        return KT_ENODEV;
    }
    kt_ctx* c = new (std::nothrow) kt_ctx();
    if (!c) return KT_ENOMEM;
    c->device = device;
    c->budget = hbm_budget_bytes;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) {
        (void)hipGetLastError();
        delete c;
        return KT_EHIP;
    }
    c->num_cus = prop.multiProcessorCount;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        (void)hipGetLastError();
        delete c;
        return KT_EHIP;
    }
    for (int i = 0; i < KT_NUM_EVENTS; ++i) {
        if (hipEventCreate(&c->ev[i]) != hipSuccess) {
            (void)hipGetLastError();
            kt_destroy(c);
            return KT_EHIP;
        }
    }
    if (hipHostMalloc(reinterpret_cast<void**>(&c->host_small), 64 * sizeof(uint64_t), hipHostMallocDefault) !=
        hipSuccess) {
        (void)hipGetLastError();
        c->host_small = nullptr;
        kt_destroy(c);
        return KT_EHIP;
    }
    random_keys(c->sym_key);
    // the no-allocation entry's workspace address on this device (retried by
    // that entry if it fails here)
    (void)noalloc_init(device);
    c->prm.wave_max_deg = 128;
    c->prm.group_max_deg = 2048;
    c->prm.team_chunk = 8;
    c->prm.grid_per_cu = 4;
    c->prm.strategy = TC_STREAMED;
    c->prm.order = TC_BY_DEGREE;
    trace("context on device %d: %d CUs, stream %p", device, c->num_cus, (void*)c->stream);
    *out = c;
    return KT_OK;
}

extern "C" void kt_destroy(kt_ctx* c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    c->destroy_extensions();
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->arena) (void)hipFree(c->arena);
    if (c->host_small) (void)hipHostFree(c->host_small);
    for (int i = 0; i < KT_NUM_EVENTS; ++i)
        if (c->ev[i]) (void)hipEventDestroy(c->ev[i]);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    (void)hipGetLastError();
    delete c;
}

extern "C" const char* kt_strerror(int s) {
    switch (s) {
        case KT_OK: return "ok";
        case KT_EINVAL: return "invalid argument or malformed graph";
        case KT_ENOMEM: return "device memory budget exhausted";
        case KT_EHIP: return "HIP runtime error";
        case KT_ENODEV: return "no usable GPU device";
        case KT_ENOTREADY: return "request not finished";
        case KT_ECANCELED: return "request cancelled";
        case KT_ENOTFOUND: return "unknown request";
        case KT_ESTATE: return "call not valid in the current state";
        case KT_ERANGE: return "output buffer too small";
        default: return "unknown status";
    }
}

extern "C" const char* kt_last_error(const kt_ctx* c) { return c ? c->err : "no context"; }

extern "C" int kt_set_params(kt_ctx* c, const kt_params* p) {
    if (!c || !p) return KT_EINVAL;
    if (p->wave_max_deg < 2 || p->wave_max_deg > 256 || p->group_max_deg < p->wave_max_deg ||
        p->group_max_deg > 2048 || p->grid_per_cu < 1 || p->grid_per_cu > 32 || p->strategy > KT_STRATEGY_GROUPED2 ||
        p->order > KT_ORDER_VERTEX || p->bsearch_min_deg < 1 || p->bsearch_min_deg > 8192 || p->pseudo2core > 1 ||
        p->min_side > 1)
        return KT_EINVAL;
    std::lock_guard<std::mutex> lk(c->mu);
    c->prm.wave_max_deg = p->wave_max_deg;
    c->prm.group_max_deg = p->group_max_deg;
    c->prm.grid_per_cu = p->grid_per_cu;
    c->prm.strategy = p->strategy;
    c->prm.order = p->order;
    c->prm.bsearch_min_deg = p->bsearch_min_deg;
    c->prm.pseudo2core = p->pseudo2core;
    c->prm.min_side = p->min_side;
    return KT_OK;
}

extern "C" int kt_get_params(const kt_ctx* c, kt_params* p) {
    if (!c || !p) return KT_EINVAL;
    p->wave_max_deg = c->prm.wave_max_deg;
    p->group_max_deg = c->prm.group_max_deg;
    p->grid_per_cu = c->prm.grid_per_cu;
    p->strategy = c->prm.strategy;
    p->order = c->prm.order;
    p->bsearch_min_deg = c->prm.bsearch_min_deg;
    p->pseudo2core = c->prm.pseudo2core;
    p->min_side = c->prm.min_side;
    return KT_OK;
}

extern "C" int kt_last_stats(const kt_ctx* c, kt_stats* s) {
    if (!c || !s) return KT_EINVAL;
    *s = c->stats;
    return KT_OK;
}

extern "C" int kt_last_timing(const kt_ctx* c, kt_timing* t) {
    if (!c || !t) return KT_EINVAL;
    *t = c->timing;
    return KT_OK;
}

extern "C" uint64_t kt_device_bytes(const kt_ctx* c) { return c ? c->arena_bytes + c->pinned_other : 0; }

// ---------------------------------------------------------------------------
// plain path: orientation + intersection
// ---------------------------------------------------------------------------
namespace {

struct PlainBufs {
    const uint64_t* off;
    const uint32_t* heads;
    uint32_t* deg;
    uint64_t* ooff;
    uint32_t* oheads;
    uint64_t* keys;
    uint64_t* alt;
    uint64_t* ws_orient;
    uint64_t* ws_tc;
    uint64_t* ws_radix;
    uint64_t* pv;
    uint64_t* small;  // [0] = err flags (u32), [1] = total, [2..] = tc stats
};

// Lays out every buffer of the plain path; with base == nullptr only sizes.
uint64_t plain_layout(void* base, uint32_t n, uint64_t nnz, bool host_in, bool want_pv_dev, bool unsorted,
                      PlainBufs& B) {
    Carver cv{base, 0};
    const uint64_t m = nnz / 2;
    B.small = static_cast<uint64_t*>(cv.take(64 * sizeof(uint64_t)));
    B.off = nullptr;
    B.heads = nullptr;
    if (host_in) {
        B.off = static_cast<uint64_t*>(cv.take((uint64_t(n) + 1) * 8));
        B.heads = static_cast<uint32_t*>(cv.take(std::max<uint64_t>(nnz, 1) * 4));
    }
    B.deg = static_cast<uint32_t*>(cv.take(std::max<uint64_t>(n, 1) * 4));
    B.ooff = static_cast<uint64_t*>(cv.take((uint64_t(n) + 1) * 8));
    B.oheads = static_cast<uint32_t*>(cv.take((m + 4) * 4));  // +4: 16-byte block reads past the end
    B.ws_orient = static_cast<uint64_t*>(cv.take(orient_ws_words(n, nnz) * 8));
    // also the scratch of the longest-list reduction of unsorted input
    B.ws_tc = static_cast<uint64_t*>(cv.take(std::max(tc_ws_words(n), segsort_tmp_words()) * 8));
    B.pv = want_pv_dev ? static_cast<uint64_t*>(cv.take(std::max<uint64_t>(n, 1) * 8)) : nullptr;
    B.keys = B.alt = B.ws_radix = nullptr;
    if (unsorted) {
        B.keys = static_cast<uint64_t*>(cv.take(std::max<uint64_t>(m, 1) * 8));
        B.alt = static_cast<uint64_t*>(cv.take(std::max<uint64_t>(m, 1) * 8));
        B.ws_radix = static_cast<uint64_t*>(cv.take(radix_tmp_words(m) * 8));
    }
    return cv.used;
}

int bits_for(uint64_t x) {
    int b = 0;
    while (b < 64 && (x >> b)) ++b;
    return b;
}

// Uploads host input (if any) and runs the count pass of the orientation.
int upload_and_count(kt_ctx* c, uint32_t n, uint64_t nnz, const uint64_t* user_off, const uint32_t* user_heads,
                     bool host_in, PlainBufs& B) {
    hipStream_t st = c->stream;
    int rc;
    if (host_in) {
        rc = c->hip_check(hipMemcpyAsync(const_cast<uint64_t*>(B.off), user_off, (uint64_t(n) + 1) * 8,
                                         hipMemcpyHostToDevice, st),
                          "H2D offsets");
        if (rc) return rc;
        if (nnz) {
            rc = c->hip_check(hipMemcpyAsync(const_cast<uint32_t*>(B.heads), user_heads, nnz * 4,
                                             hipMemcpyHostToDevice, st),
                              "H2D heads");
            if (rc) return rc;
        }
    } else {
        B.off = user_off;
        B.heads = user_heads;
    }
    (void)hipEventRecord(c->ev[EV_H2D], st);
    trace("upload done, launching count pass");
    fill_u64(B.small, 64, 0, st);
    OrientCheck chk{{c->sym_key[0], c->sym_key[1]}};
    orient_count_launch(n, nnz, B.off, B.heads, B.deg, B.ooff, B.ws_orient, reinterpret_cast<uint32_t*>(B.small), chk,
                        st);
    rc = c->hip_check(hipMemcpyAsync(c->host_small, B.small, 8, hipMemcpyDeviceToHost, st), "D2H flags");
    if (rc) return rc;
    rc = c->hip_check(hipMemcpyAsync(c->host_small + 1, B.ooff + n, 8, hipMemcpyDeviceToHost, st), "D2H kept");
    if (rc) return rc;
    return c->hip_check(hipStreamSynchronize(st), "orientation count");
}

// Builds the oriented CSR in B.ooff / B.oheads.  The count pass is checked on
// the host (one stream sync) before anything is written, so malformed or
// asymmetric input never reaches the write pass or the intersection.
// The device workspace of a plain call: the context's grow-only arena, or a
// fixed region given by the caller (never grown: a call that does not fit
// fails with KT_ENOMEM).
struct Workspace {
    void* fixed_base = nullptr;
    uint64_t fixed_cap = 0;
    void* base(const kt_ctx* c) const { return fixed_base ? fixed_base : c->arena; }
    uint64_t cap(const kt_ctx* c) const { return fixed_base ? fixed_cap : c->arena_bytes; }
    int ensure(kt_ctx* c, uint64_t need) const {
        if (!fixed_base) return c->reserve(need);
        if (need <= fixed_cap) return KT_OK;
        return c->fail(KT_ENOMEM, "graph needs %llu bytes of workspace, the fixed workspace holds %llu",
                       (unsigned long long)need, (unsigned long long)fixed_cap);
    }
};

int run_orientation(kt_ctx* c, uint32_t n, uint64_t nnz, const uint64_t* user_off, const uint32_t* user_heads,
                    bool host_in, bool pv_dev_needed, PlainBufs& B, bool& unsorted, const Workspace& ws = {}) {
    hipStream_t st = c->stream;
    unsorted = false;
    uint64_t need = plain_layout(nullptr, n, nnz, host_in, pv_dev_needed, false, B);
    trace("orientation: n=%u nnz=%llu workspace=%llu", n, (unsigned long long)nnz, (unsigned long long)need);
    int rc = ws.ensure(c, need);
    if (rc) return rc;
    plain_layout(ws.base(c), n, nnz, host_in, pv_dev_needed, false, B);
    trace("workspace=%p bytes=%llu", ws.base(c), (unsigned long long)ws.cap(c));
    (void)hipEventRecord(c->ev[EV_START], st);
    rc = upload_and_count(c, n, nnz, user_off, user_heads, host_in, B);
    if (rc) return rc;
    uint32_t flags = static_cast<uint32_t>(c->host_small[0]);
    uint64_t kept = c->host_small[1];
    trace("count pass: flags=0x%x kept=%llu", flags, (unsigned long long)kept);
    if (flags & (KT_ERR_MASK & ~KT_ERR_ASYMMETRIC)) {
        return c->fail(KT_EINVAL, "malformed CSR:%s%s%s%s", (flags & KT_ERR_OFFSETS) ? " bad offsets" : "",
                       (flags & KT_ERR_HEAD_RANGE) ? " head id >= n" : "",
                       (flags & KT_ERR_SELF_LOOP) ? " self-loop" : "",
                       (flags & KT_ERR_DUPLICATE) ? " duplicate edge" : "");
    }
    if (flags & KT_ERR_ASYMMETRIC)
        return c->fail(KT_EINVAL, "CSR is not symmetric (some slot v -> u has no slot u -> v)");
    if (2 * kept != nnz) {
        return c->fail(KT_EINVAL, "CSR is not symmetric (%llu oriented edges for %llu slots)",
                       (unsigned long long)kept, (unsigned long long)nnz);
    }
    const uint64_t m = nnz / 2;
    uint64_t max_len = 0;
    if (flags & KT_FLAG_UNSORTED) {
        // lists of up to segsort_max_len() ids are sorted in place, per list;
        // a longer one needs the global radix sort (and its buffers)
        max_len_launch(n, B.ooff, B.ws_tc, B.small + 8, st);
        rc = c->hip_check(hipMemcpyAsync(c->host_small + 8, B.small + 8, 8, hipMemcpyDeviceToHost, st),
                          "D2H longest list");
        if (rc) return rc;
        rc = c->hip_check(hipStreamSynchronize(st), "longest list");
        if (rc) return rc;
        max_len = c->host_small[8];
        trace("unsorted lists, longest oriented list %llu", (unsigned long long)max_len);
    }
    if ((flags & KT_FLAG_UNSORTED) && max_len <= segsort_max_len()) {
        unsorted = true;
        orient_write_launch(n, nnz, B.heads, B.oheads, B.ws_orient, st);
        segsort_launch(n, B.ooff, B.oheads, reinterpret_cast<uint32_t*>(B.small), max_len, c->num_cus, st);
        rc = c->hip_check(hipMemcpyAsync(c->host_small, B.small, 8, hipMemcpyDeviceToHost, st), "D2H flags");
        if (rc) return rc;
        rc = c->hip_check(hipStreamSynchronize(st), "sort");
        if (rc) return rc;
        flags = static_cast<uint32_t>(c->host_small[0]);
        if (flags & KT_ERR_MASK) return c->fail(KT_EINVAL, "malformed CSR: duplicate edge");
    } else if (flags & KT_FLAG_UNSORTED) {
        unsorted = true;
        need = plain_layout(nullptr, n, nnz, host_in, pv_dev_needed, true, B);
        if (need > ws.cap(c)) {
            // growing the arena drops its contents: upload and count again
            rc = ws.ensure(c, need);
            if (rc) return rc;
            plain_layout(ws.base(c), n, nnz, host_in, pv_dev_needed, true, B);
            rc = upload_and_count(c, n, nnz, user_off, user_heads, host_in, B);
            if (rc) return rc;
        } else {
            plain_layout(ws.base(c), n, nnz, host_in, pv_dev_needed, true, B);
            if (!host_in) {
                B.off = user_off;
                B.heads = user_heads;
            }
        }
        orient_write_launch(n, nnz, B.heads, B.oheads, B.ws_orient, st);
        make_keys(n, m, B.ooff, B.oheads, B.keys, st);
        radix_sort_u64(B.keys, B.alt, m, 32 + bits_for(n), B.ws_radix, st);
        keys_to_heads(m, B.keys, B.oheads, reinterpret_cast<uint32_t*>(B.small), st);
        rc = c->hip_check(hipMemcpyAsync(c->host_small, B.small, 8, hipMemcpyDeviceToHost, st), "D2H flags");
        if (rc) return rc;
        rc = c->hip_check(hipStreamSynchronize(st), "sort");
        if (rc) return rc;
        flags = static_cast<uint32_t>(c->host_small[0]);
        if (flags & KT_ERR_MASK) return c->fail(KT_EINVAL, "malformed CSR: duplicate edge");
    } else {
        orient_write_launch(n, nnz, B.heads, B.oheads, B.ws_orient, st);
    }
    c->stats.oriented_edges = kept;
    c->stats.vertices = n;
    c->stats.input_slots = nnz;
    return KT_OK;
}

int read_nnz(kt_ctx* c, uint32_t n, const uint64_t* offsets, uint32_t flags, uint64_t& nnz) {
    if (flags & KT_DEVICE_INPUT) {
        uint64_t v[2] = {0, 0};
        int rc = c->read_u64(offsets, v, "D2H offsets[0]");
        if (rc) return rc;
        rc = c->read_u64(offsets + n, v + 1, "D2H offsets[n]");
        if (rc) return rc;
        if (v[0] != 0) return c->fail(KT_EINVAL, "offsets[0] != 0");
        nnz = v[1];
    } else {
        if (offsets[0] != 0) return c->fail(KT_EINVAL, "offsets[0] != 0");
        nnz = offsets[n];
    }
    if (nnz & 1) return c->fail(KT_EINVAL, "odd number of CSR entries: not a symmetric graph");
    if (nnz / 2 > 0xffffffffull * 2) return c->fail(KT_EINVAL, "graph too large");
    return KT_OK;
}

}  // namespace

// The body of kt_count on a given workspace (the caller holds c->mu and has
// begun the call).
int kt::count_locked(kt_ctx* c, uint32_t n, const uint64_t* offsets, const uint32_t* heads, uint32_t flags,
                     uint64_t* total_out, uint64_t* per_vertex_out, void* fixed_ws, uint64_t fixed_ws_bytes) {
    Workspace ws;
    ws.fixed_base = fixed_ws;
    ws.fixed_cap = fixed_ws_bytes;
    int rc;
    if (!total_out || !offsets || (n > 0 && !heads && !(flags & KT_DEVICE_INPUT)))
        return c->fail(KT_EINVAL, "null argument");
    *total_out = 0;
    if (n == 0) return KT_OK;
    uint64_t nnz = 0;
    rc = read_nnz(c, n, offsets, flags, nnz);
    if (rc) return rc;
    const bool host_in = !(flags & KT_DEVICE_INPUT);
    const bool want_pv = per_vertex_out != nullptr;
    const bool pv_dev = want_pv && (flags & KT_DEVICE_OUTPUT);
    PlainBufs B;
    bool unsorted;
    rc = run_orientation(c, n, nnz, offsets, heads, host_in, want_pv && !pv_dev, B, unsorted, ws);
    if (rc) return rc;
    hipStream_t st = c->stream;
    (void)hipEventRecord(c->ev[EV_PREP], st);
    uint64_t* pv = want_pv ? (pv_dev ? per_vertex_out : B.pv) : nullptr;
    if (pv) fill_u64(pv, n, 0, st);
    uint64_t* totals = B.small + 1;
    TcView view{};
    view.X = n;
    view.s_off = B.ooff;
    view.sheads = B.oheads;
    if (c->prm.min_side) {
        view.in_off = B.off;
        view.in_heads = B.heads;
        view.in_keep = orient_keep_bits(n, nnz, B.ws_orient);
    }
    TcOut out{pv, totals};
    trace("intersection: X=%u pv=%p", n, (void*)pv);
    // wave-bin tails per queue access for one graph: each grab costs a queue
    // atomic and one round trip for the tails' bounds, and a larger grab
    // leaves more uneven work at the end of the launch.  With the min-side
    // walk: RMAT-22 k_tc_main 17.0 ms with 4 tails per grab, 13.9 with 8, 12.6
    // with 16 or 24, 12.6 with 32, 12.9 with 64; G(n,m) 2^16 0.207 ms with 8,
    // 0.170 with 16, 0.218 with 32.  (The forward walk was fastest with 8.)
    // The shard and batch views, whose tails are many and tiny, take 32.
    TcParams pp = c->prm;
    pp.wave_grab = 16;
    rc = c->hip_check(tc_launch(view, out, pp, B.ws_tc, c->num_cus, st, c->ev[EV_TC0], c->ev[EV_TC1]),
                      "intersection launch");
    if (rc) return rc;
    (void)hipEventRecord(c->ev[EV_COUNT], st);
    uint64_t* tstats = tc_stats_ptr(B.ws_tc);
    rc = c->hip_check(hipMemcpyAsync(c->host_small + 1, totals, 8, hipMemcpyDeviceToHost, st), "D2H total");
    if (rc) return rc;
    rc = c->hip_check(hipMemcpyAsync(c->host_small + 2, tstats, 32, hipMemcpyDeviceToHost, st), "D2H stats");
    if (rc) return rc;
    if (want_pv && !pv_dev) {
        rc = c->hip_check(hipMemcpyAsync(per_vertex_out, pv, uint64_t(n) * 8, hipMemcpyDeviceToHost, st),
                          "D2H per-vertex");
        if (rc) return rc;
    }
    (void)hipEventRecord(c->ev[EV_END], st);
    rc = c->hip_check(hipStreamSynchronize(st), "count");
    if (rc) return rc;
    trace("count done: total=%llu", (unsigned long long)c->host_small[1]);
    *total_out = c->host_small[1];
    c->stats.local_triangles = c->host_small[1];
    c->stats.wedge_checks = c->host_small[2];
    c->stats.intersection_size_local = c->host_small[3];
    c->stats.skipped_nodes = c->host_small[4];
    c->stats.streamed_ids = c->host_small[5];
    c->timing.h2d_ms = c->elapsed(EV_START, EV_H2D);
    c->timing.prep_ms = c->elapsed(EV_H2D, EV_PREP);
    c->timing.count_ms = c->elapsed(EV_PREP, EV_COUNT);
    c->timing.intersect_ms = c->elapsed(EV_TC0, EV_TC1);
    c->timing.d2h_ms = c->elapsed(EV_COUNT, EV_END);
    c->timing.total_ms = c->elapsed(EV_START, EV_END);
    return KT_OK;
}

extern "C" int kt_count(kt_ctx* c, uint32_t n, const uint64_t* offsets, const uint32_t* heads, uint32_t flags,
                        uint64_t* total_out, uint64_t* per_vertex_out) {
    if (!c) return KT_EINVAL;
    std::lock_guard<std::mutex> lk(c->mu);
    int rc = c->begin_call(true);
    if (rc) return rc;
    return count_locked(c, n, offsets, heads, flags, total_out, per_vertex_out, nullptr, 0);
}

extern "C" int kt_list_triangles(kt_ctx* c, uint32_t n, const uint64_t* offsets, const uint32_t* heads,
                                 uint32_t flags, uint32_t* triangles, uint64_t capacity, uint64_t* total_out) {
    if (!c) return KT_EINVAL;
    std::lock_guard<std::mutex> lk(c->mu);
    int rc = c->begin_call(true);
    if (rc) return rc;
    if (!total_out || !offsets || (n > 0 && !heads && !(flags & KT_DEVICE_INPUT)))
        return c->fail(KT_EINVAL, "null argument");
    *total_out = 0;
    if (n == 0) return KT_OK;
    uint64_t nnz = 0;
    rc = read_nnz(c, n, offsets, flags, nnz);
    if (rc) return rc;
    const bool host_in = !(flags & KT_DEVICE_INPUT);
    const bool dev_out = (flags & KT_DEVICE_OUTPUT) != 0;
    PlainBufs B;
    bool unsorted;
    rc = run_orientation(c, n, nnz, offsets, heads, host_in, false, B, unsorted);
    if (rc) return rc;
    // per-tail triangle counts, scanned in place into output offsets
    Carver cv{nullptr, 0};
    cv.take(((uint64_t)n + 1) * 8);
    cv.take(scan_tmp_words((uint64_t)n + 1) * 8);
    rc = c->reserve_aux(0, cv.used);
    if (rc) return rc;
    cv = Carver{c->aux[0], 0};
    uint64_t* tcnt = static_cast<uint64_t*>(cv.take(((uint64_t)n + 1) * 8));
    uint64_t* stmp = static_cast<uint64_t*>(cv.take(scan_tmp_words((uint64_t)n + 1) * 8));
    hipStream_t st = c->stream;
    (void)hipEventRecord(c->ev[EV_PREP], st);
    fill_u64(tcnt, (uint64_t)n + 1, 0, st);
    uint64_t* totals = B.small + 1;
    TcView view{};
    view.X = n;
    view.s_off = B.ooff;
    view.sheads = B.oheads;
    TcOut cnt_out{nullptr, totals};
    cnt_out.tail_count = tcnt;
    rc = c->hip_check(tc_launch(view, cnt_out, c->prm, B.ws_tc, c->num_cus, st, c->ev[EV_TC0], c->ev[EV_TC1]),
                      "intersection launch");
    if (rc) return rc;
    scan_u64(tcnt, tcnt, n, stmp, st);
    rc = c->hip_check(hipMemcpyAsync(c->host_small + 2, tc_stats_ptr(B.ws_tc), 32, hipMemcpyDeviceToHost, st),
                      "D2H stats");
    if (rc) return rc;
    uint64_t total = 0;
    rc = c->read_u64(tcnt + n, &total, "D2H total");
    if (rc) return rc;
    *total_out = total;
    c->stats.local_triangles = total;
    c->stats.wedge_checks = c->host_small[2];
    c->stats.intersection_size_local = c->host_small[3];
    c->stats.skipped_nodes = c->host_small[4];
    c->stats.streamed_ids = c->host_small[5];
    if (total > capacity)
        return c->fail(KT_ERANGE, "%llu triangles do not fit a buffer of %llu", (unsigned long long)total,
                       (unsigned long long)capacity);
    if (total && !triangles) return c->fail(KT_EINVAL, "null triangle buffer");
    if (total) {
        uint32_t* dtri = triangles;
        if (!dev_out) {
            rc = c->reserve_aux(1, total * 12);
            if (rc) return rc;
            dtri = static_cast<uint32_t*>(c->aux[1]);
        }
        TcOut list_out{nullptr, totals};
        list_out.tail_off = tcnt;
        list_out.tri = dtri;
        rc = c->hip_check(tc_launch(view, list_out, c->prm, B.ws_tc, c->num_cus, st, nullptr, nullptr),
                          "listing launch");
        if (rc) return rc;
        (void)hipEventRecord(c->ev[EV_COUNT], st);
        if (!dev_out) {
            rc = c->hip_check(hipMemcpyAsync(triangles, dtri, total * 12, hipMemcpyDeviceToHost, st), "D2H triangles");
            if (rc) return rc;
        }
    } else {
        (void)hipEventRecord(c->ev[EV_COUNT], st);
    }
    (void)hipEventRecord(c->ev[EV_END], st);
    rc = c->hip_check(hipStreamSynchronize(st), "listing");
    if (rc) return rc;
    c->timing.h2d_ms = c->elapsed(EV_START, EV_H2D);
    c->timing.prep_ms = c->elapsed(EV_H2D, EV_PREP);
    c->timing.count_ms = c->elapsed(EV_PREP, EV_COUNT);
    c->timing.intersect_ms = c->elapsed(EV_TC0, EV_TC1);
    c->timing.d2h_ms = c->elapsed(EV_COUNT, EV_END);
    c->timing.total_ms = c->elapsed(EV_START, EV_END);
    return KT_OK;
}

extern "C" int kt_orient(kt_ctx* c, uint32_t n, const uint64_t* offsets, const uint32_t* heads, uint32_t flags,
                         uint64_t* out_offsets, uint32_t* out_heads) {
    if (!c) return KT_EINVAL;
    std::lock_guard<std::mutex> lk(c->mu);
    int rc = c->begin_call(true);
    if (rc) return rc;
    if (!offsets || !out_offsets) return c->fail(KT_EINVAL, "null argument");
    if (n == 0) {
        if (flags & KT_DEVICE_OUTPUT)
            return c->hip_check(hipMemset(out_offsets, 0, 8), "memset");
        out_offsets[0] = 0;
        return KT_OK;
    }
    uint64_t nnz = 0;
    rc = read_nnz(c, n, offsets, flags, nnz);
    if (rc) return rc;
    PlainBufs B;
    bool unsorted;
    rc = run_orientation(c, n, nnz, offsets, heads, !(flags & KT_DEVICE_INPUT), false, B, unsorted);
    if (rc) return rc;
    hipStream_t st = c->stream;
    hipMemcpyKind kind = (flags & KT_DEVICE_OUTPUT) ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
    rc = c->hip_check(hipMemcpyAsync(out_offsets, B.ooff, (uint64_t(n) + 1) * 8, kind, st), "copy offsets");
    if (rc) return rc;
    if (nnz && out_heads) {
        rc = c->hip_check(hipMemcpyAsync(out_heads, B.oheads, nnz / 2 * 4, kind, st), "copy heads");
        if (rc) return rc;
    }
    return c->hip_check(hipStreamSynchronize(st), "orient");
}

// ---------------------------------------------------------------------------
// device buffers
// ---------------------------------------------------------------------------
extern "C" int kt_buffer_alloc(kt_ctx* c, uint64_t bytes, void** dptr) {
    if (!c || !dptr) return KT_EINVAL;
    std::lock_guard<std::mutex> lk(c->mu);
    int rc = c->begin_call(false);
    if (rc) return rc;
    *dptr = nullptr;
    bytes = std::max<uint64_t>(bytes, 1);
    if (c->budget && c->arena_bytes + c->pinned_other + bytes > c->budget)
        return c->fail(KT_ENOMEM, "buffer of %llu bytes exceeds the HBM budget", (unsigned long long)bytes);
    hipError_t e = hipMalloc(dptr, bytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        *dptr = nullptr;
        return c->fail(KT_ENOMEM, "hipMalloc(%llu): %s", (unsigned long long)bytes, hipGetErrorString(e));
    }
    c->buffers[*dptr] = bytes;
    c->pinned_other += bytes;
    return KT_OK;
}

extern "C" int kt_buffer_free(kt_ctx* c, void* dptr) {
    if (!c) return KT_EINVAL;
    if (!dptr) return KT_OK;
    std::lock_guard<std::mutex> lk(c->mu);
    int rc = c->begin_call(false);
    if (rc) return rc;
    auto it = c->buffers.find(dptr);
    if (it == c->buffers.end()) return c->fail(KT_EINVAL, "pointer was not allocated by kt_buffer_alloc");
    rc = c->hip_check(hipStreamSynchronize(c->stream), "sync");
    if (rc) return rc;
    c->pinned_other -= it->second;
    c->buffers.erase(it);
    return c->hip_check(hipFree(dptr), "hipFree");
}

extern "C" int kt_buffer_upload(kt_ctx* c, void* dst, const void* src, uint64_t bytes) {
    if (!c || (bytes && (!dst || !src))) return KT_EINVAL;
    std::lock_guard<std::mutex> lk(c->mu);
    int rc = c->begin_call(false);
    if (rc) return rc;
    if (!bytes) return KT_OK;
    rc = c->hip_check(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream), "upload");
    if (rc) return rc;
    return c->hip_check(hipStreamSynchronize(c->stream), "upload sync");
}

extern "C" int kt_buffer_download(kt_ctx* c, void* dst, const void* src, uint64_t bytes) {
    if (!c || (bytes && (!dst || !src))) return KT_EINVAL;
    std::lock_guard<std::mutex> lk(c->mu);
    int rc = c->begin_call(false);
    if (rc) return rc;
    if (!bytes) return KT_OK;
    rc = c->hip_check(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream), "download");
    if (rc) return rc;
    return c->hip_check(hipStreamSynchronize(c->stream), "download sync");
}

extern "C" int kt_host_alloc(uint64_t bytes, void** host_ptr) {
    if (!host_ptr) return KT_EINVAL;
    *host_ptr = nullptr;
    if (hipHostMalloc(host_ptr, std::max<uint64_t>(bytes, 1), hipHostMallocDefault) != hipSuccess) {
        (void)hipGetLastError();
        *host_ptr = nullptr;
        return KT_ENOMEM;
    }
    return KT_OK;
}

extern "C" int kt_host_free(void* host_ptr) {
    if (!host_ptr) return KT_OK;
    if (hipHostFree(host_ptr) != hipSuccess) {
        (void)hipGetLastError();
        return KT_EHIP;
    }
    return KT_OK;
}

extern "C" int kt_synchronize(kt_ctx* c) {
    if (!c) return KT_EINVAL;
    std::lock_guard<std::mutex> lk(c->mu);
    int rc = c->begin_call(false);
    if (rc) return rc;
    return c->hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
}

void kt_ctx::destroy_extensions() {
    kt_serve_release(this);
    kt_batch_release(this);
    release_aux();
    for (auto& kv : buffers) (void)hipFree(kv.first);
    buffers.clear();
}
