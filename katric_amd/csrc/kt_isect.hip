// Batched intersection of caller-given sorted id lists (kt_intersect).
//
// Replaces cetric::intersection (counters/intersection.h:40-79) for callers
// that intersect lists of their own, as tests/intersection_test.cpp:35-173
// does: the common ids of every pair are counted and, optionally, written in
// ascending order -- the order merge_intersection (:81-96) emits them in.
//
// One wave per pair (grid-stride over pairs).  The shorter list is read 64
// ids at a time, one per lane, and each id is looked up by binary search in
// the longer list in global memory; the lanes that hit are ranked with a
// ballot, so a pair's matches are written in list order.  Both lists are
// checked to be strictly increasing (the reference's debug precondition,
// intersection.h:50-54): the check reads every id once.
#include <algorithm>

#include "../../include/katric.h"
#include "kt_ctx.h"
#include "kt_device.h"

namespace kt {

constexpr int IS_THREADS = 256;
#define KT_ERR_ISECT_ORDER 0x1u  // a list is not strictly increasing
#define KT_ERR_ISECT_OFFSETS 0x2u

// Index of the first entry >= w in l[0, n).
__device__ __forceinline__ uint64_t lower_bound_u32(const uint32_t* __restrict__ l, uint64_t n, uint32_t w) {
    uint64_t lo = 0, hi = n;
    while (lo < hi) {
        const uint64_t mid = (lo + hi) >> 1;
        if (l[mid] < w)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}

// WRITE = false: counts[p] = |A_p cap B_p| and order / offset validation.
// WRITE = true: pair p's common ids, ascending, from out + out_off[p] on.
// A pair whose bounds are not ascending or leave [0, na) / [0, nb) (na, nb:
// the id arrays' lengths, the offsets' last entries) is flagged and skipped,
// so no read leaves the id arrays whatever the offsets hold.
template <bool WRITE>
__global__ void __launch_bounds__(IS_THREADS)
    k_isect(uint64_t P, const uint64_t* __restrict__ aoff, const uint32_t* __restrict__ a, uint64_t na,
            const uint64_t* __restrict__ boff, const uint32_t* __restrict__ b, uint64_t nb,
            uint64_t* __restrict__ counts, const uint64_t* __restrict__ out_off, uint32_t* __restrict__ out,
            uint32_t* __restrict__ err) {
    const uint32_t lane = lane_id();
    const uint64_t nw = (uint64_t)gridDim.x * (IS_THREADS / KT_WAVE);
    uint32_t flags = 0;
    for (uint64_t p = (uint64_t)blockIdx.x * (IS_THREADS / KT_WAVE) + threadIdx.x / KT_WAVE; p < P; p += nw) {
        const uint64_t a0 = aoff[p], a1 = aoff[p + 1], b0 = boff[p], b1 = boff[p + 1];
        if (a1 < a0 || b1 < b0 || a1 > na || b1 > nb) {
            flags |= KT_ERR_ISECT_OFFSETS;
            if (!WRITE && lane == 0) counts[p] = 0;
            continue;
        }
        const bool a_short = a1 - a0 <= b1 - b0;
        const uint32_t* s = a_short ? a + a0 : b + b0;
        const uint32_t* l = a_short ? b + b0 : a + a0;
        const uint64_t ns = a_short ? a1 - a0 : b1 - b0;
        const uint64_t nl = a_short ? b1 - b0 : a1 - a0;
        if (!WRITE) {
            for (uint64_t i = lane; i + 1 < nl; i += KT_WAVE)
                if (l[i] >= l[i + 1]) flags |= KT_ERR_ISECT_ORDER;
        }
        uint64_t hits = 0;
        uint64_t dst = WRITE ? out_off[p] : 0;
        for (uint64_t j = 0; j < ns; j += KT_WAVE) {
            const uint64_t i = j + lane;
            bool hit = false;
            uint32_t w = 0;
            if (i < ns) {
                w = s[i];
                if (!WRITE && i + 1 < ns && w >= s[i + 1]) flags |= KT_ERR_ISECT_ORDER;
                const uint64_t k = lower_bound_u32(l, nl, w);
                hit = k < nl && l[k] == w;
            }
            const uint64_t m = __ballot(hit);
            if (WRITE && hit) out[dst + __popcll(m & lanemask_lt())] = w;
            dst += (uint64_t)__popcll(m);
            hits += (uint64_t)__popcll(m);
        }
        if (!WRITE && lane == 0) counts[p] = hits;
    }
    or_flags(err, flags);
}

}  // namespace kt

using namespace kt;

extern "C" int kt_intersect(kt_ctx* c, uint64_t num_pairs, const uint64_t* a_offsets, const uint32_t* a_ids,
                            const uint64_t* b_offsets, const uint32_t* b_ids, uint32_t flags, uint64_t* counts,
                            uint32_t* matches, uint64_t capacity, uint64_t* total_out) {
    if (!c) return KT_EINVAL;
    std::lock_guard<std::mutex> lk(c->mu);
    int rc = c->begin_call(true);
    if (rc) return rc;
    if (!a_offsets || !b_offsets || !counts || !total_out) return c->fail(KT_EINVAL, "null argument");
    *total_out = 0;
    if (num_pairs == 0) return KT_OK;
    if (num_pairs >= (1ull << 40)) return c->fail(KT_EINVAL, "too many pairs");
    const bool host_in = !(flags & KT_DEVICE_INPUT);
    const bool dev_out = (flags & KT_DEVICE_OUTPUT) != 0;
    hipStream_t st = c->stream;
    // list lengths of the two id arrays
    uint64_t na = 0, nb = 0;
    if (host_in) {
        if (a_offsets[0] != 0 || b_offsets[0] != 0) return c->fail(KT_EINVAL, "offsets[0] != 0");
        na = a_offsets[num_pairs];
        nb = b_offsets[num_pairs];
    } else {
        uint64_t v[4];
        const uint64_t* src[4] = {a_offsets, a_offsets + num_pairs, b_offsets, b_offsets + num_pairs};
        for (int i = 0; i < 4; ++i) {
            rc = c->read_u64(src[i], &v[i], "D2H offsets");
            if (rc) return rc;
        }
        if (v[0] != 0 || v[2] != 0) return c->fail(KT_EINVAL, "offsets[0] != 0");
        na = v[1];
        nb = v[3];
    }
    if ((na && !a_ids) || (nb && !b_ids)) return c->fail(KT_EINVAL, "null id array");
    const uint64_t P = num_pairs;
    Carver cv{nullptr, 0};
    auto layout = [&](Carver& k, uint64_t*& small, uint64_t*& aoff, uint64_t*& boff, uint32_t*& a, uint32_t*& b,
                      uint64_t*& cnt, uint64_t*& coff, uint64_t*& stmp) {
        small = static_cast<uint64_t*>(k.take(64 * 8));
        aoff = host_in ? static_cast<uint64_t*>(k.take((P + 1) * 8)) : nullptr;
        boff = host_in ? static_cast<uint64_t*>(k.take((P + 1) * 8)) : nullptr;
        a = host_in ? static_cast<uint32_t*>(k.take(std::max<uint64_t>(na, 1) * 4)) : nullptr;
        b = host_in ? static_cast<uint32_t*>(k.take(std::max<uint64_t>(nb, 1) * 4)) : nullptr;
        cnt = static_cast<uint64_t*>(k.take((P + 1) * 8));
        coff = static_cast<uint64_t*>(k.take((P + 1) * 8));
        stmp = static_cast<uint64_t*>(k.take(scan_tmp_words(P + 1) * 8));
    };
    uint64_t *small, *aoff, *boff, *cnt, *coff, *stmp;
    uint32_t *a, *b;
    layout(cv, small, aoff, boff, a, b, cnt, coff, stmp);
    rc = c->reserve(cv.used);
    if (rc) return rc;
    cv = Carver{c->arena, 0};
    layout(cv, small, aoff, boff, a, b, cnt, coff, stmp);
    (void)hipEventRecord(c->ev[EV_START], st);
    if (host_in) {
        rc = c->hip_check(hipMemcpyAsync(aoff, a_offsets, (P + 1) * 8, hipMemcpyHostToDevice, st), "H2D");
        if (!rc) rc = c->hip_check(hipMemcpyAsync(boff, b_offsets, (P + 1) * 8, hipMemcpyHostToDevice, st), "H2D");
        if (!rc && na) rc = c->hip_check(hipMemcpyAsync(a, a_ids, na * 4, hipMemcpyHostToDevice, st), "H2D");
        if (!rc && nb) rc = c->hip_check(hipMemcpyAsync(b, b_ids, nb * 4, hipMemcpyHostToDevice, st), "H2D");
        if (rc) return rc;
    } else {
        aoff = const_cast<uint64_t*>(a_offsets);
        boff = const_cast<uint64_t*>(b_offsets);
        a = const_cast<uint32_t*>(a_ids);
        b = const_cast<uint32_t*>(b_ids);
    }
    (void)hipEventRecord(c->ev[EV_H2D], st);
    uint32_t* err = reinterpret_cast<uint32_t*>(small);
    fill_u64(small, 64, 0, st);
    const unsigned grid =
        (unsigned)std::min<uint64_t>((P + IS_THREADS / KT_WAVE - 1) / (IS_THREADS / KT_WAVE), (uint64_t)c->num_cus * 16);
    k_isect<false><<<grid, IS_THREADS, 0, st>>>(P, aoff, a, na, boff, b, nb, cnt, nullptr, nullptr, err);
    scan_u64(cnt, coff, P, stmp, st);
    rc = c->hip_check(hipMemcpyAsync(c->host_small, small, 8, hipMemcpyDeviceToHost, st), "D2H flags");
    if (!rc) rc = c->hip_check(hipMemcpyAsync(c->host_small + 1, coff + P, 8, hipMemcpyDeviceToHost, st), "D2H total");
    if (!rc) rc = c->hip_check(hipStreamSynchronize(st), "intersect count");
    if (rc) return rc;
    const uint32_t fl = (uint32_t)c->host_small[0];
    const uint64_t total = c->host_small[1];
    if (fl & KT_ERR_ISECT_OFFSETS) return c->fail(KT_EINVAL, "pair offsets not ascending");
    if (fl & KT_ERR_ISECT_ORDER) return c->fail(KT_EINVAL, "a list is not strictly increasing");
    *total_out = total;
    c->stats.local_triangles = 0;
    c->stats.wedge_checks = 0;
    c->stats.intersection_size_local = na + nb;  // ids of both lists, as the reference counts per intersection
    rc = c->hip_check(hipMemcpyAsync(counts, cnt, P * 8, dev_out ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, st),
                      "copy counts");
    if (rc) return rc;
    if (matches && total > capacity) {
        rc = c->hip_check(hipStreamSynchronize(st), "intersect");
        if (rc) return rc;
        return c->fail(KT_ERANGE, "%llu matches do not fit a buffer of %llu", (unsigned long long)total,
                       (unsigned long long)capacity);
    }
    if (matches && total) {
        uint32_t* dm = matches;
        if (!dev_out) {
            rc = c->reserve_aux(1, total * 4);
            if (rc) return rc;
            dm = static_cast<uint32_t*>(c->aux[1]);
        }
        k_isect<true><<<grid, IS_THREADS, 0, st>>>(P, aoff, a, na, boff, b, nb, nullptr, coff, dm, err);
        if (!dev_out) {
            rc = c->hip_check(hipMemcpyAsync(matches, dm, total * 4, hipMemcpyDeviceToHost, st), "D2H matches");
            if (rc) return rc;
        }
    }
    (void)hipEventRecord(c->ev[EV_END], st);
    rc = c->hip_check(hipStreamSynchronize(st), "intersect");
    if (rc) return rc;
    c->timing.h2d_ms = c->elapsed(EV_START, EV_H2D);
    c->timing.count_ms = c->elapsed(EV_H2D, EV_END);
    c->timing.total_ms = c->elapsed(EV_START, EV_END);
    return KT_OK;
}
