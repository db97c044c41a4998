// Device-side producer of one rank's shard inputs (kt_count_shard_graph).
//
// From a full symmetric CSR resident in HBM the rank's kt_shard_in is built
// on the device -- what the reference's MPI exchange delivers to the rank --
// and then counted by the kt_count_shard pipeline:
//   1. global degrees deg[v] of every vertex (offsets validated);
//   2. one pass over the rank's slots [off[from], off[to]): a bit per
//      non-owned head (its ghosts, find_ghosts, distributed_graph.h:1287-1298)
//      and a bit per ghost a with (deg a, a) < (deg x, x) for an owned
//      neighbour x (a has an out-edge into the rank, so a's owner sends
//      [a, list] to it, cetric_edge_iterator.h:1328-1379, 1632-1661);
//   3. both bitmaps compacted in id order (popcount per word, scan, write):
//      ghost ids ascending with their global degrees (DegreeCommunicator,
//      graph_communicator.h:48-266), senders ascending;
//   4. the senders' full lists, concatenated in sender order, form a virtual
//      slot space (virtual offsets = scan of the senders' degrees); tiles of
//      4096 virtual slots resolve each slot's sender in LDS and keep x with
//      (deg a, a) < (deg x, x) -- cetric-x also drops the x owned by a's owner
//      (N+cut(a)), patric keeps all of N+(a).  A counting pass gives each
//      tile's kept count; after their scan the same pass writes every kept id
//      at (tile base + rank in the tile), which is its position in the
//      concatenated message lists, and each sender's first virtual slot gives
//      its message offset.  Lists that are not ascending are put in order by
//      one key sort over (message, id).
// The rank's slice is not copied: its offsets are rebased into a new array and
// its heads are the full graph's heads from off[from] on.  The senders are
// found from the rank's own slots; for a symmetric CSR that is the same set
// as walking every ghost's list.
#include <algorithm>

#include "../../include/katric_shard.h"
#include "kt_ctx.h"
#include "kt_device.h"

namespace kt {

constexpr int PD_THREADS = 256;
constexpr int PD_ITEMS = 16;
constexpr int PD_TILE = PD_THREADS * PD_ITEMS;

__device__ __forceinline__ bool before(uint32_t da, uint32_t a, uint32_t db, uint32_t b) {
    return da < db || (da == db && a < b);
}

__device__ __forceinline__ void set_bit(uint32_t* bits, uint32_t u) {
    const uint32_t w = u >> 5, b = 1u << (u & 31);
    if (!(bits[w] & b)) atomicOr(&bits[w], b);  // a stale read only costs a redundant atomic
}

__global__ void k_pd_degree(uint32_t n, uint64_t nnz, const uint64_t* __restrict__ off, uint32_t* __restrict__ deg,
                            uint32_t* __restrict__ err) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += stride) {
        if (v == 0 && (off[0] != 0 || off[n] != nnz)) atomicOr(err, KT_ERR_OFFSETS);
        const uint64_t a = off[v], b = off[v + 1];
        if (b < a || b - a > 0xffffffffull || b > nnz) {
            atomicOr(err, KT_ERR_OFFSETS);
            deg[v] = 0;
        } else {
            deg[v] = (uint32_t)(b - a);
        }
    }
}

// Slots of the rank's rows [from, from + nl): loff = rebased offsets (loff[0]
// = 0), lheads = the rows' heads, tbnd = rows of every tile's first and last
// slot.
__global__ void __launch_bounds__(PD_THREADS)
    k_pd_mark(const uint64_t* __restrict__ loff, uint32_t nl, uint32_t from, uint32_t n, uint64_t nnz_l,
              const uint32_t* __restrict__ lheads, const uint32_t* __restrict__ tbnd, const uint32_t* __restrict__ deg,
              uint32_t* __restrict__ gbits, uint32_t* __restrict__ sbits, uint32_t* __restrict__ err) {
    __shared__ uint32_t marks[PD_TILE];
    __shared__ uint32_t wmax[PD_THREADS / KT_WAVE];
    __shared__ uint32_t bnd[2];
    const uint64_t tile0 = (uint64_t)blockIdx.x * PD_TILE;
    const uint64_t tend = min(tile0 + PD_TILE, nnz_l);
    uint32_t lo, hi;
    tile_rows<PD_THREADS, PD_ITEMS>(loff, nl, tile0, tend, marks, bnd, wmax, lo, hi, tbnd);
    const uint32_t to = from + nl;
    uint32_t us[PD_ITEMS];
#pragma unroll
    for (int k = 0; k < PD_ITEMS; ++k) {
        const uint64_t sl = tile0 + k * PD_THREADS + threadIdx.x;
        us[k] = sl < tend ? lheads[sl] : from;
    }
    uint32_t flags = 0;
#pragma unroll
    for (int k = 0; k < PD_ITEMS; ++k) {
        const unsigned i = k * PD_THREADS + threadIdx.x;
        const uint32_t u = us[k];
        if (tile0 + i >= tend) continue;
        if (u >= n) {
            flags |= KT_ERR_HEAD_RANGE;
        } else if (u < from || u >= to) {
            const uint32_t x = from + lo + marks[i];
            set_bit(gbits, u);
            if (before(deg[u], u, deg[x], x)) set_bit(sbits, u);
        }
    }
    or_flags(err, flags);
}

__global__ void k_pd_popc(uint64_t words, const uint32_t* __restrict__ gbits, const uint32_t* __restrict__ sbits,
                          uint64_t* __restrict__ gcnt, uint64_t* __restrict__ scnt) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; w < words; w += stride) {
        gcnt[w] = __popc(gbits[w]);
        scnt[w] = __popc(sbits[w]);
    }
}

__global__ void k_pd_compact(uint64_t words, const uint32_t* __restrict__ gbits, const uint64_t* __restrict__ goff,
                             const uint32_t* __restrict__ sbits, const uint64_t* __restrict__ soff,
                             const uint64_t* __restrict__ off, uint32_t* __restrict__ ghost_ids,
                             uint64_t* __restrict__ ghost_deg, uint32_t* __restrict__ msg_src,
                             uint64_t* __restrict__ vlen) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; w < words; w += stride) {
        uint32_t bits = gbits[w];
        uint64_t p = goff[w];
        while (bits) {
            const uint32_t id = (uint32_t)(w * 32) + (__ffs(bits) - 1);
            bits &= bits - 1;
            ghost_ids[p] = id;
            ghost_deg[p] = off[id + 1] - off[id];
            ++p;
        }
        bits = sbits[w];
        p = soff[w];
        while (bits) {
            const uint32_t id = (uint32_t)(w * 32) + (__ffs(bits) - 1);
            bits &= bits - 1;
            msg_src[p] = id;
            vlen[p] = off[id + 1] - off[id];
            ++p;
        }
    }
}

__global__ void k_pd_rebase(uint32_t nl, const uint64_t* __restrict__ soff, uint64_t* __restrict__ loff) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i <= nl; i += stride) loff[i] = soff[i] - soff[0];
}

// [lo, hi): the id range of the owner of a, the first rank r with a < bd[r+1]
// (as the reference's rank lookup over range ends, util/helpers.h:83-101).
__device__ __forceinline__ void owner_range(const uint64_t* __restrict__ bd, uint32_t P, uint32_t a, uint64_t& lo,
                                            uint64_t& hi) {
    uint32_t l = 0, h = P;
    while (l < h) {
        const uint32_t mid = (l + h) >> 1;
        if (a < bd[mid + 1])
            h = mid;
        else
            l = mid + 1;
    }
    const uint32_t o = min(l, P - 1);
    lo = bd[o];
    hi = bd[o + 1];
}

// What a virtual slot needs of its sender.
struct SenderInfo {
    uint64_t offa;    // off[a]: a's first slot in the full CSR
    uint64_t vbeg;    // the message's first virtual slot
    uint32_t a, da;   // sender id and its global degree
    uint32_t lo, hi;  // ids excluded from the message: cetric-x those owned by a's owner, patric none
};

__global__ void k_pd_sinfo(uint32_t nmsg, const uint32_t* __restrict__ src, const uint64_t* __restrict__ voff,
                           const uint64_t* __restrict__ off, const uint32_t* __restrict__ deg,
                           const uint64_t* __restrict__ bd, uint32_t P, int cetric, SenderInfo* __restrict__ si) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nmsg; i += gridDim.x * blockDim.x) {
        const uint32_t a = src[i];
        uint64_t lo = 0, hi = 0;
        if (cetric) owner_range(bd, P, a, lo, hi);
        SenderInfo s;
        s.offa = off[a];
        s.vbeg = voff[i];
        s.a = a;
        s.da = deg[a];
        s.lo = (uint32_t)lo;
        s.hi = (uint32_t)hi;
        si[i] = s;
    }
}

// Tiles of the virtual slot space [0, V).  Counting pass (!EMIT): keep bits
// (one word per 64 slots, to vkeep), per-tile kept counts, validation and
// order flags.  EMIT: kept ids, by the counting pass's bits (no degree
// gathers), to mheads at tile_off[tile] + rank in the tile, and moff[i] =
// the output position of sender i's first virtual slot.
template <bool EMIT>
__global__ void __launch_bounds__(PD_THREADS)
    k_pd_msgs(uint64_t V, uint32_t nmsg, const uint64_t* __restrict__ voff, const uint32_t* __restrict__ tbnd,
              const SenderInfo* __restrict__ si, const uint32_t* __restrict__ heads, const uint32_t* __restrict__ deg,
              uint32_t n, uint64_t* __restrict__ vkeep, uint64_t* __restrict__ tile_cnt,
              const uint64_t* __restrict__ tile_off, uint32_t* __restrict__ mheads, uint64_t* __restrict__ moff,
              uint32_t* __restrict__ err) {
    constexpr int NW = PD_THREADS / KT_WAVE;
    __shared__ uint32_t marks[PD_TILE];
    __shared__ uint32_t wmax[NW];
    __shared__ uint32_t bnd[2];
    __shared__ uint32_t wcnt[PD_ITEMS * NW];
    const uint64_t tile0 = (uint64_t)blockIdx.x * PD_TILE;
    const uint64_t tend = min(tile0 + PD_TILE, V);
    uint32_t lo, hi;
    tile_rows<PD_THREADS, PD_ITEMS>(voff, nmsg, tile0, tend, marks, bnd, wmax, lo, hi, tbnd);
    const unsigned tid = threadIdx.x, wave = tid / KT_WAVE;
    // every item's list slot and head first, then the degree gathers
    uint32_t xs[PD_ITEMS], dx[PD_ITEMS];
    uint64_t ps[PD_ITEMS];
#pragma unroll
    for (int k = 0; k < PD_ITEMS; ++k) {
        const unsigned idx = k * PD_THREADS + tid;
        const uint64_t j = tile0 + idx;
        const SenderInfo* s = si + (lo + marks[idx]);
        ps[k] = s->offa + (j - s->vbeg);
        xs[k] = j < tend ? heads[ps[k]] : 0u;
    }
    uint32_t flags = 0;
    uint64_t bits[PD_ITEMS];
    if constexpr (!EMIT) {
#pragma unroll
        for (int k = 0; k < PD_ITEMS; ++k)
            dx[k] = (tile0 + k * PD_THREADS + tid < tend && xs[k] < n) ? deg[xs[k]] : 0u;
    }
#pragma unroll
    for (int k = 0; k < PD_ITEMS; ++k) {
        const unsigned idx = k * PD_THREADS + tid;
        const uint64_t j = tile0 + idx;
        const uint64_t w0 = tile0 + k * PD_THREADS + (tid & ~63u);  // first slot of this wave's word
        if constexpr (EMIT) {
            bits[k] = w0 < tend ? vkeep[w0 / 64] : 0;
        } else {
            bool keep = false;
            if (j < tend) {
                const SenderInfo* s = si + (lo + marks[idx]);
                const uint32_t a = s->a, x = xs[k];
                if (x >= n) {
                    flags |= KT_ERR_HEAD_RANGE;
                } else {
                    keep = before(s->da, a, dx[k], x) && !(x >= s->lo && x < s->hi);
                    if (ps[k] + 1 < s->offa + s->da && heads[ps[k] + 1] <= x) flags |= KT_FLAG_UNSORTED;
                }
            }
            bits[k] = __ballot(keep);
            if (lane_id() == 0 && w0 < tend) vkeep[w0 / 64] = bits[k];
        }
        if (lane_id() == 0) wcnt[k * NW + wave] = (uint32_t)__popcll(bits[k]);
    }
    if (!EMIT) or_flags(err, flags);
    __syncthreads();
    if (!EMIT) {
        if (tid == 0) {
            uint64_t t = 0;
            for (int q = 0; q < PD_ITEMS * NW; ++q) t += wcnt[q];
            tile_cnt[blockIdx.x] = t;
        }
        return;
    }
    uint64_t base = tile_off[blockIdx.x];
#pragma unroll
    for (int k = 0; k < PD_ITEMS; ++k) {
        uint32_t before_w = 0, round = 0;
#pragma unroll
        for (int w = 0; w < NW; ++w) {
            const uint32_t c = wcnt[k * NW + w];
            before_w += (w < (int)wave) ? c : 0u;
            round += c;
        }
        const unsigned idx = k * PD_THREADS + tid;
        const uint64_t j = tile0 + idx;
        const uint64_t pos = base + before_w + __popcll(bits[k] & lanemask_lt());
        if ((bits[k] >> lane_id()) & 1) mheads[pos] = xs[k];
        if (j < tend) {
            const uint32_t i = lo + marks[idx];
            if (j == si[i].vbeg) moff[i] = pos;
        }
        base += round;
    }
}

// keys[p] = (message << 32) | id over every message entry
__global__ void k_pd_keys(uint32_t nmsg, const uint64_t* __restrict__ moff, const uint32_t* __restrict__ mheads,
                          uint64_t* __restrict__ keys) {
    const uint32_t wid = (blockIdx.x * blockDim.x + threadIdx.x) / KT_WAVE;
    const uint32_t nw = gridDim.x * blockDim.x / KT_WAVE;
    for (uint32_t i = wid; i < nmsg; i += nw)
        for (uint64_t p = moff[i] + lane_id(); p < moff[i + 1]; p += KT_WAVE) keys[p] = ((uint64_t)i << 32) | mheads[p];
}

__global__ void k_pd_keys_low(uint64_t len, const uint64_t* __restrict__ keys, uint32_t* __restrict__ mheads) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; p < len; p += stride)
        mheads[p] = (uint32_t)keys[p];
}

// Per-vertex counts of the rank's combined index space (owned, then ghosts)
// added to the global-id array.
__global__ void k_pd_scatter(uint32_t nl, uint32_t ng, uint32_t from, const uint32_t* __restrict__ ghost_ids,
                             const uint64_t* __restrict__ pvx, uint64_t* __restrict__ out) {
    const uint64_t X = (uint64_t)nl + ng;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; c < X; c += stride) {
        const uint32_t g = c < nl ? from + (uint32_t)c : ghost_ids[c - nl];
        out[g] += pvx[c];
    }
}

}  // namespace kt

using namespace kt;

namespace {

unsigned grid_for(uint64_t items, unsigned per_block = 256, unsigned cap = 16384) {
    return (unsigned)std::max<uint64_t>(1, std::min<uint64_t>((items + per_block - 1) / per_block, cap));
}

int bits_of(uint64_t x) {
    int b = 0;
    while (b < 64 && (x >> b)) ++b;
    return b;
}

struct ProdBufs {
    uint64_t* small;
    uint64_t* bd;
    uint32_t* deg;
    uint32_t* bits;  // ghost bits, then sender bits
    uint64_t* gcnt;
    uint64_t* scnt;
    uint64_t* stmp;
    uint32_t* tbnd;
    uint64_t* tcnt;  // per-tile kept counts of the virtual slot space, then their scan
    uint64_t* vkeep; // keep bits of the virtual slot space, one word per 64 slots
    uint64_t* loff;
    uint32_t* ghost_ids;
    uint64_t* ghost_deg;
    uint32_t* msg_src;
    uint64_t* voff;  // virtual offsets of the senders' full lists
    uint64_t* moff;
    SenderInfo* sinfo;
    uint64_t* pvx;
    uint64_t* gpv;
    uint64_t* hoff;
    uint32_t* hheads;
    uint64_t* symws;  // symmetry check of the whole graph
};

uint64_t prod_layout(void* base, uint32_t n, uint64_t nnz, uint32_t P, uint32_t nl, uint64_t ng_max, bool want_pv,
                     bool host_pv, bool host_in, ProdBufs& B) {
    Carver cv{base, 0};
    const uint64_t W = ((uint64_t)n + 31) / 32;
    const uint64_t tiles = nnz / PD_TILE + 1;  // bounds the tiles of the slice and of the virtual slot space
    B.small = static_cast<uint64_t*>(cv.take(64 * 8));
    B.bd = static_cast<uint64_t*>(cv.take(((uint64_t)P + 1) * 8));
    B.deg = static_cast<uint32_t*>(cv.take((uint64_t)n * 4));
    B.bits = static_cast<uint32_t*>(cv.take(2 * W * 4));
    B.gcnt = static_cast<uint64_t*>(cv.take((W + 1) * 8));
    B.scnt = static_cast<uint64_t*>(cv.take((W + 1) * 8));
    B.stmp = static_cast<uint64_t*>(cv.take(scan_tmp_words(std::max(std::max<uint64_t>(W, ng_max), tiles) + 1) * 8));
    B.tbnd = static_cast<uint32_t*>(cv.take(2 * tiles * 4));
    B.tcnt = static_cast<uint64_t*>(cv.take((tiles + 1) * 8));
    B.vkeep = static_cast<uint64_t*>(cv.take(tiles * (PD_TILE / 64) * 8));
    B.loff = static_cast<uint64_t*>(cv.take(((uint64_t)nl + 1) * 8));
    B.ghost_ids = static_cast<uint32_t*>(cv.take(std::max<uint64_t>(ng_max, 1) * 4));
    B.ghost_deg = static_cast<uint64_t*>(cv.take(std::max<uint64_t>(ng_max, 1) * 8));
    B.msg_src = static_cast<uint32_t*>(cv.take(std::max<uint64_t>(ng_max, 1) * 4));
    B.voff = static_cast<uint64_t*>(cv.take((ng_max + 1) * 8));
    B.moff = static_cast<uint64_t*>(cv.take((ng_max + 1) * 8));
    B.sinfo = static_cast<SenderInfo*>(cv.take(std::max<uint64_t>(ng_max, 1) * sizeof(SenderInfo)));
    B.pvx = want_pv ? static_cast<uint64_t*>(cv.take(std::max<uint64_t>(nl + ng_max, 1) * 8)) : nullptr;
    B.gpv = host_pv ? static_cast<uint64_t*>(cv.take((uint64_t)n * 8)) : nullptr;
    B.hoff = host_in ? static_cast<uint64_t*>(cv.take(((uint64_t)n + 1) * 8)) : nullptr;
    B.hheads = host_in ? static_cast<uint32_t*>(cv.take(std::max<uint64_t>(nnz, 1) * 4)) : nullptr;
    B.symws = static_cast<uint64_t*>(cv.take(sym_ws_words(n, nnz) * 8));
    return cv.used;
}

int read_small(kt_ctx* c, const uint64_t* small, int words, const char* what) {
    int rc = c->hip_check(hipMemcpyAsync(c->host_small, small, words * 8, hipMemcpyDeviceToHost, c->stream), what);
    if (rc) return rc;
    return c->hip_check(hipStreamSynchronize(c->stream), what);
}

int count_shard_graph_locked(kt_ctx* c, uint32_t n, const uint64_t* offsets, const uint32_t* heads, uint32_t flags,
                             uint32_t algorithm, uint32_t P, const uint64_t* bd, uint32_t rank, kt_shard_out* out) {
    if (!out || !offsets || !bd || P == 0 || rank >= P || algorithm > KT_SHARD_DITRIC)
        return c->fail(KT_EINVAL, "bad shard arguments");
    if (bd[0] != 0 || bd[P] != n) return c->fail(KT_EINVAL, "range_bounds must cover [0, n)");
    if (c->prm.pseudo2core) return c->fail(KT_EINVAL, "pseudo2core applies to the shared-memory path only");
    for (uint32_t r = 0; r < P; ++r)
        if (bd[r + 1] < bd[r]) return c->fail(KT_EINVAL, "range bounds not ascending");
    out->local_triangles = out->type3_triangles = 0;
    if (n == 0) return KT_OK;
    const bool host_in = !(flags & KT_DEVICE_INPUT);
    const bool want_pv = out->per_vertex != nullptr;
    const bool pv_dev = want_pv && (flags & KT_DEVICE_OUTPUT);
    const bool acc = (flags & KT_SHARD_ACCUMULATE) != 0;
    const bool cetric = algorithm == KT_SHARD_CETRIC;
    const uint32_t from = (uint32_t)bd[rank], to = (uint32_t)bd[rank + 1], nl = to - from;
    hipStream_t st = c->stream;
    int rc;
    // sizes of the graph and of the rank's slice
    uint64_t nnz, base, end;
    if (host_in) {
        nnz = offsets[n];
        base = offsets[from];
        end = offsets[to];
    } else {
        const uint64_t* src[3] = {offsets + n, offsets + from, offsets + to};
        for (int i = 0; i < 3; ++i) {
            rc = c->hip_check(hipMemcpyAsync(c->host_small + i, src[i], 8, hipMemcpyDeviceToHost, st), "D2H sizes");
            if (rc) return rc;
        }
        rc = c->hip_check(hipStreamSynchronize(st), "D2H sizes");
        if (rc) return rc;
        nnz = c->host_small[0];
        base = c->host_small[1];
        end = c->host_small[2];
    }
    if (base > end || end > nnz || (nnz >> 40)) return c->fail(KT_EINVAL, "malformed CSR: bad offsets");
    if (nnz && !heads) return c->fail(KT_EINVAL, "null heads");
    const uint64_t nnz_l = end - base;
    const uint64_t ng_max = std::min<uint64_t>((uint64_t)n - nl, nnz_l);
    ProdBufs B;
    const uint64_t need = prod_layout(nullptr, n, nnz, P, nl, ng_max, want_pv, want_pv && !pv_dev, host_in, B);
    rc = c->reserve_aux(0, need);
    if (rc) return rc;
    prod_layout(c->aux[0], n, nnz, P, nl, ng_max, want_pv, want_pv && !pv_dev, host_in, B);
    (void)hipEventRecord(c->ev[EV_X0], st);
    const uint64_t* off = offsets;
    const uint32_t* hd = heads;
    if (host_in) {
        rc = c->hip_check(hipMemcpyAsync(B.hoff, offsets, ((uint64_t)n + 1) * 8, hipMemcpyHostToDevice, st), "H2D");
        if (!rc && nnz) rc = c->hip_check(hipMemcpyAsync(B.hheads, heads, nnz * 4, hipMemcpyHostToDevice, st), "H2D");
        if (rc) return rc;
        off = B.hoff;
        hd = B.hheads;
    }
    rc = c->hip_check(hipMemcpyAsync(B.bd, bd, ((uint64_t)P + 1) * 8, hipMemcpyHostToDevice, st), "H2D bounds");
    if (rc) return rc;
    uint32_t* err = reinterpret_cast<uint32_t*>(B.small);
    const uint64_t W = ((uint64_t)n + 31) / 32;
    fill_u64(B.small, 64, 0, st);
    fill_u32(B.bits, 2 * W, 0, st);
    fill_u64(B.voff, ng_max + 1, 0, st);
    k_pd_degree<<<grid_for(n), 256, 0, st>>>(n, nnz, off, B.deg, err);
    // the whole graph must be symmetric, as every rank's view of it assumes
    // (one pass over all slots: pair checksums, self-loops, duplicates)
    const OrientCheck chk{{c->sym_key[0], c->sym_key[1]}};
    sym_check_launch(n, nnz, off, hd, B.symws, err, chk, st);
    // The kernels up to the first flag check stay in bounds whatever the
    // offsets hold: slots are bounded by the host-checked slice [base, end),
    // rows by tile_rows' clamping, ids by n.
    k_pd_rebase<<<grid_for((uint64_t)nl + 1), 256, 0, st>>>(nl, off + from, B.loff);
    const uint64_t tiles_l = (nnz_l + PD_TILE - 1) / PD_TILE;
    if (nl && tiles_l) {
        tile_bounds_launch(nl, nnz_l, PD_TILE, B.loff, B.tbnd, st);
        k_pd_mark<<<(unsigned)tiles_l, PD_THREADS, 0, st>>>(B.loff, nl, from, n, nnz_l, hd + base, B.tbnd, B.deg,
                                                            B.bits, B.bits + W, err);
    }
    // ghosts and senders, in id order; the senders' degrees are the lengths
    // of the virtual lists (entries past the last sender stay zero)
    k_pd_popc<<<grid_for(W), 256, 0, st>>>(W, B.bits, B.bits + W, B.gcnt, B.scnt);
    scan_u64(B.gcnt, B.gcnt, W, B.stmp, st);
    scan_u64(B.scnt, B.scnt, W, B.stmp, st);
    k_pd_compact<<<grid_for(W), 256, 0, st>>>(W, B.bits, B.gcnt, B.bits + W, B.scnt, off, B.ghost_ids, B.ghost_deg,
                                               B.msg_src, B.voff);
    scan_u64(B.voff, B.voff, ng_max, B.stmp, st);
    rc = c->hip_check(hipMemcpyAsync(B.small + 1, B.gcnt + W, 8, hipMemcpyDeviceToDevice, st), "D2D counts");
    if (!rc) rc = c->hip_check(hipMemcpyAsync(B.small + 2, B.scnt + W, 8, hipMemcpyDeviceToDevice, st), "D2D counts");
    if (!rc) rc = c->hip_check(hipMemcpyAsync(B.small + 3, B.voff + ng_max, 8, hipMemcpyDeviceToDevice, st), "D2D V");
    if (!rc) rc = read_small(c, B.small, 4, "ghosts");
    if (rc) return rc;
    uint32_t fl = (uint32_t)c->host_small[0];
    const uint64_t ng = c->host_small[1], nmsg = c->host_small[2], V = c->host_small[3];
    if (fl & KT_ERR_MASK)
        return c->fail(KT_EINVAL, "malformed CSR:%s%s%s%s%s", (fl & KT_ERR_OFFSETS) ? " bad offsets" : "",
                       (fl & KT_ERR_HEAD_RANGE) ? " head id >= n" : "", (fl & KT_ERR_SELF_LOOP) ? " self-loop" : "",
                       (fl & KT_ERR_DUPLICATE) ? " duplicate edge" : "",
                       (fl & KT_ERR_ASYMMETRIC) ? " not symmetric" : "");
    if (ng > ng_max || nmsg > ng || V > nnz) return c->fail(KT_EINVAL, "shard producer: inconsistent ghost counts");
    // message lists: count, scan, emit over the virtual slot space
    uint64_t mtot = 0;
    uint32_t* mheads = nullptr;
    if (nmsg) {
        const uint64_t tiles_v = (V + PD_TILE - 1) / PD_TILE;
        k_pd_sinfo<<<grid_for(nmsg), 256, 0, st>>>((uint32_t)nmsg, B.msg_src, B.voff, off, B.deg, B.bd, P,
                                                   cetric ? 1 : 0, B.sinfo);
        tile_bounds_launch((uint32_t)nmsg, V, PD_TILE, B.voff, B.tbnd, st);
        k_pd_msgs<false><<<(unsigned)tiles_v, PD_THREADS, 0, st>>>(V, (uint32_t)nmsg, B.voff, B.tbnd, B.sinfo, hd,
                                                                   B.deg, n, B.vkeep, B.tcnt, nullptr, nullptr,
                                                                   nullptr, err);
        scan_u64(B.tcnt, B.tcnt, tiles_v, B.stmp, st);
        rc = c->hip_check(hipMemcpyAsync(B.small + 1, B.tcnt + tiles_v, 8, hipMemcpyDeviceToDevice, st), "D2D total");
        if (!rc) rc = read_small(c, B.small, 2, "message lengths");
        if (rc) return rc;
        fl = (uint32_t)c->host_small[0];
        mtot = c->host_small[1];
        if (fl & KT_ERR_MASK)
            return c->fail(KT_EINVAL, "malformed CSR:%s", (fl & KT_ERR_HEAD_RANGE) ? " head id >= n" : "");
        const bool unsorted = (fl & KT_FLAG_UNSORTED) != 0;
        Carver cv{nullptr, 0};
        cv.take(std::max<uint64_t>(mtot, 1) * 4);
        if (unsorted) {
            cv.take(mtot * 8);
            cv.take(mtot * 8);
            cv.take(radix_tmp_words(mtot) * 8);
        }
        rc = c->reserve_aux(1, cv.used);
        if (rc) return rc;
        cv = Carver{c->aux[1], 0};
        mheads = static_cast<uint32_t*>(cv.take(std::max<uint64_t>(mtot, 1) * 4));
        k_pd_msgs<true><<<(unsigned)tiles_v, PD_THREADS, 0, st>>>(V, (uint32_t)nmsg, B.voff, B.tbnd, B.sinfo, hd,
                                                                  B.deg, n, B.vkeep, nullptr, B.tcnt, mheads, B.moff,
                                                                  err);
        fill_u64(B.moff + nmsg, 1, mtot, st);
        if (unsorted && mtot) {
            uint64_t* keys = static_cast<uint64_t*>(cv.take(mtot * 8));
            uint64_t* alt = static_cast<uint64_t*>(cv.take(mtot * 8));
            uint64_t* rtmp = static_cast<uint64_t*>(cv.take(radix_tmp_words(mtot) * 8));
            k_pd_keys<<<grid_for(nmsg * KT_WAVE, 256, 8192), 256, 0, st>>>((uint32_t)nmsg, B.moff, mheads, keys);
            radix_sort_u64(keys, alt, mtot, 32 + bits_of(nmsg), rtmp, st);
            k_pd_keys_low<<<grid_for(mtot), 256, 0, st>>>(mtot, keys, mheads);
        }
    }
    (void)hipEventRecord(c->ev[EV_X2], st);
    // the rank's count on the produced inputs
    kt_shard_in in{};
    in.algorithm = algorithm;
    in.rank = rank;
    in.num_ranks = P;
    in.flags = KT_DEVICE_INPUT;
    in.range_bounds = bd;
    in.offsets = B.loff;
    in.heads = hd + base;
    in.n_ghost = (uint32_t)ng;
    in.ghost_ids = B.ghost_ids;
    in.ghost_degree = B.ghost_deg;
    in.n_msg = (uint32_t)nmsg;
    in.msg_src = B.msg_src;
    in.msg_offsets = B.moff;
    in.msg_heads = mheads;
    kt_shard_out o{0, 0, nullptr};
    if (want_pv) {
        in.flags |= KT_DEVICE_OUTPUT;
        o.per_vertex = B.pvx;
        fill_u64(B.pvx, (uint64_t)nl + ng, 0, st);  // stays zero if the rank sees nothing
    }
    rc = count_shard_locked(c, &in, &o);
    if (rc) return rc;
    if (want_pv) {
        uint64_t* dst = pv_dev ? out->per_vertex : B.gpv;
        if (!acc)
            fill_u64(dst, n, 0, st);
        else if (!pv_dev) {
            rc = c->hip_check(hipMemcpyAsync(dst, out->per_vertex, (uint64_t)n * 8, hipMemcpyHostToDevice, st),
                              "H2D per-vertex");
            if (rc) return rc;
        }
        k_pd_scatter<<<grid_for((uint64_t)nl + ng), 256, 0, st>>>(nl, (uint32_t)ng, from, B.ghost_ids, B.pvx, dst);
        if (!pv_dev) {
            rc = c->hip_check(hipMemcpyAsync(out->per_vertex, dst, (uint64_t)n * 8, hipMemcpyDeviceToHost, st),
                              "D2H per-vertex");
            if (rc) return rc;
        }
    }
    (void)hipEventRecord(c->ev[EV_X1], st);
    rc = c->hip_check(hipStreamSynchronize(st), "shard graph");
    if (rc) return rc;
    out->local_triangles = o.local_triangles;
    out->type3_triangles = o.type3_triangles;
    // the producer counts as preprocessing (it stands for the exchange of
    // the rank's inputs); total spans producer, count and per-vertex scatter
    c->timing.prep_ms += c->elapsed(EV_X0, EV_X2);
    c->timing.total_ms = c->elapsed(EV_X0, EV_X1);
    return KT_OK;
}

}  // namespace

extern "C" int kt_count_shard_graph(kt_ctx* c, uint32_t n, const uint64_t* offsets, const uint32_t* heads,
                                    uint32_t flags, uint32_t algorithm, uint32_t num_ranks,
                                    const uint64_t* range_bounds, uint32_t rank, kt_shard_out* out) {
    if (!c) return KT_EINVAL;
    std::lock_guard<std::mutex> lk(c->mu);
    int rc = c->begin_call(true);
    if (rc) return rc;
    return count_shard_graph_locked(c, n, offsets, heads, flags, algorithm, num_ranks, range_bounds, rank, out);
}
