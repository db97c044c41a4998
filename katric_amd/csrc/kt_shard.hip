// One rank's shard of CeTriC ("cetric-x") / DiTriC ("patric") on one GPU.
//
// Vertex space: a combined index c over the rank's owned vertices
// [0, n_local) (c = id - from) followed by its ghosts [n_local, n_local +
// n_ghost) in ascending id order.  Sorting any list by c therefore equals the
// reference's node_ordering::id_outward(rank) (owned first, then by id,
// distributed_graph.h:102-112), and the ghost part of a list is its suffix.
//
// Preprocessing (one radix sort):
//   every owned slot (v, u) with c = map(u) contributes
//     (v, c)  if (deg v, v) < (deg u, u)           -- N+(v), global degrees
//     (c, v)  if not, c is a ghost and cetric-x     -- ghost_adj(c)
//   (remove_in_edges_and_expand_ghosts, distributed_graph.h:1057-1169, with
//   node_ordering::degree on exchanged ghost degrees, :113-146).  The sorted
//   (tail << b | head) keys form a CSR L over the combined space.
// Local phase (same intersection kernel as the plain path):
//   cetric-x: every tail x (owned or ghost), S = R = L(x), T(u) = L(u)
//             (run_local_sequential_with_ghosts, cetric_edge_iterator.h:300-388);
//   patric:   owned tails, S = L(v), R = owned part of L(v), T(u) = L(u)
//             (run_local_sequential with ghost heads skipped, :229-298).
// Contracted / distributed phase: message i = [a, list]; its entries split
// into owned receivers R and ghosts G (entries unknown to this rank can match
// nothing).  cetric-x: S = G, T(u) = N+cut(u) = ghost suffix of L(u)
// (contraction, distributed_graph.h:679-764, then intersect_from_message,
// cetric_edge_iterator.h:1797-1889); patric: S = R ++ G, T(u) = L(u).  Every
// hit w emits Triangle{a, u, w}.
#include <algorithm>

#include "../../include/katric_shard.h"
#include "kt_ctx.h"
#include "kt_device.h"

namespace kt {

constexpr int SH_THREADS = 256;
constexpr int SH_ITEMS = 16;
constexpr int SH_TILE = SH_THREADS * SH_ITEMS;

// Global id -> combined index (0xffffffff: unknown to this rank).  Ghosts are
// ranked in a bitmap over all global ids: rs[w] = {ghost bits of ids 32w ..
// 32w+31, ghosts below 32w}, so a lookup is one 8-byte gather and a popcount
// (n_all / 4 bytes in all, a table the caches hold far better than one index
// per id).
struct ShardMap {
    uint32_t from, to, nl, ng;
    const uint32_t* ghost_ids;
    const uint2* rs;
    uint32_t n_all;
    __device__ __forceinline__ uint32_t map(uint32_t u) const {
        if (u >= from && u < to) return u - from;
        if (u >= n_all) return 0xffffffffu;
        const uint2 e = rs[u >> 5];
        const uint32_t bit = 1u << (u & 31);
        return (e.x & bit) ? nl + e.y + (uint32_t)__popc(e.x & (bit - 1)) : 0xffffffffu;
    }
    __device__ __forceinline__ uint32_t gid(uint32_t c) const { return c < nl ? from + c : ghost_ids[c - nl]; }
};

__global__ void k_shard_prepare(ShardMap M, uint64_t nnz, const uint64_t* __restrict__ off,
                                const uint64_t* __restrict__ ghost_degree, uint32_t* __restrict__ deg,
                                uint32_t* __restrict__ err) {
    const uint64_t X = (uint64_t)M.nl + M.ng;
    for (uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; c < X; c += (uint64_t)gridDim.x * blockDim.x) {
        if (c == 0 && (off[0] != 0 || off[M.nl] != nnz)) atomicOr(err, KT_ERR_OFFSETS);
        if (c < M.nl) {
            uint64_t a = off[c], b = off[c + 1];
            if (b < a || b - a > 0xffffffffull) {
                atomicOr(err, KT_ERR_OFFSETS);
                deg[c] = 0;
            } else {
                deg[c] = (uint32_t)(b - a);
            }
        } else {
            const uint64_t g = c - M.nl;
            const uint32_t id = M.ghost_ids[g];
            if ((id >= M.from && id < M.to) || (g > 0 && M.ghost_ids[g - 1] >= id)) atomicOr(err, KT_ERR_GHOST);
            uint64_t d = ghost_degree[g];
            deg[c] = d > 0xffffffffull ? 0xffffffffu : (uint32_t)d;
        }
    }
}

// Ghost bits of the rank-select table (zeroed before): word w of the table
// is rs32[2w], its prefix count rs32[2w + 1].
__global__ void k_ghost_bits(uint32_t ng, const uint32_t* __restrict__ ghost_ids, uint32_t n_all,
                             uint32_t* __restrict__ rs32, uint32_t* __restrict__ err) {
    for (uint32_t g = blockIdx.x * blockDim.x + threadIdx.x; g < ng; g += gridDim.x * blockDim.x) {
        const uint32_t id = ghost_ids[g];
        if (id < n_all)
            atomicOr(&rs32[2 * (id >> 5)], 1u << (id & 31));
        else
            atomicOr(err, KT_ERR_GHOST);
    }
}

__global__ void k_rs_count(uint64_t words, const uint2* __restrict__ rs, uint64_t* __restrict__ cnt) {
    for (uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; w < words; w += (uint64_t)gridDim.x * blockDim.x)
        cnt[w] = __popc(rs[w].x);
}

__global__ void k_rs_prefix(uint64_t words, const uint64_t* __restrict__ pre, uint2* __restrict__ rs) {
    for (uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; w < words; w += (uint64_t)gridDim.x * blockDim.x)
        rs[w].y = (uint32_t)pre[w];
}

__device__ __forceinline__ bool c_less(const ShardMap& M, const uint32_t* deg, uint32_t a, uint32_t b) {
    const uint32_t da = deg[a], db = deg[b];
    return da < db || (da == db && M.gid(a) < M.gid(b));
}

// Owned slots -> (tail, head) keys of N+ lists and (cetric-x) ghost lists,
// appended in any order (they are sorted next): every wave counts its keys
// over the tile's 16 slot rounds, and one atomic per workgroup reserves the
// workgroup's output range.
__global__ void __launch_bounds__(SH_THREADS)
    k_shard_entries(ShardMap M, uint64_t nnz, const uint64_t* __restrict__ off, const uint32_t* __restrict__ tbnd,
                    const uint32_t* __restrict__ heads, const uint32_t* __restrict__ deg, int bits, int ghost_lists,
                    uint64_t* __restrict__ keys, uint64_t* __restrict__ nkeys, uint32_t* __restrict__ err) {
    constexpr int NW = SH_THREADS / KT_WAVE;
    __shared__ uint32_t marks[SH_TILE];
    __shared__ uint32_t wmax[NW];
    __shared__ uint32_t bnd[2];
    __shared__ uint32_t wtot[NW];
    __shared__ uint64_t bbase;
    const uint64_t tile0 = (uint64_t)blockIdx.x * SH_TILE;
    const uint64_t tend = min(tile0 + SH_TILE, nnz);
    uint32_t lo, hi;
    tile_rows<SH_THREADS, SH_ITEMS>(off, M.nl, tile0, tend, marks, bnd, wmax, lo, hi, tbnd);
    const unsigned tid = threadIdx.x, wave = tid / KT_WAVE;
    uint32_t cs[SH_ITEMS];
#pragma unroll
    for (int k = 0; k < SH_ITEMS; ++k) {
        const uint64_t sl = tile0 + k * SH_THREADS + tid;
        cs[k] = sl < tend ? M.map(heads[sl]) : 0u;
    }
    uint32_t flags = 0, count = 0;
    uint64_t key[SH_ITEMS], pm[SH_ITEMS];
#pragma unroll
    for (int k = 0; k < SH_ITEMS; ++k) {
        const unsigned i = k * SH_THREADS + tid;
        bool out_e = false, in_e = false;
        const uint32_t v = lo + marks[i], c = cs[k];
        if (tile0 + i < tend) {
            if (c == 0xffffffffu) {
                flags |= KT_ERR_GHOST;
            } else if (c == v) {
                flags |= KT_ERR_SELF_LOOP;
            } else {
                out_e = c_less(M, deg, v, c);
                in_e = ghost_lists && !out_e && c >= M.nl;
            }
        }
        key[k] = out_e ? (((uint64_t)v << bits) | c) : (((uint64_t)c << bits) | v);
        pm[k] = __ballot(out_e || in_e);
        count += (uint32_t)__popcll(pm[k]);
    }
    or_flags(err, flags);
    if (lane_id() == 0) wtot[wave] = count;
    __syncthreads();
    if (tid == 0) {
        uint64_t t = 0;
        for (int w = 0; w < NW; ++w) t += wtot[w];
        bbase = t ? atomic_add_u64(nkeys, t) : 0;
    }
    __syncthreads();
    uint64_t base = bbase;
    for (unsigned w = 0; w < wave; ++w) base += wtot[w];
#pragma unroll
    for (int k = 0; k < SH_ITEMS; ++k) {
        if ((pm[k] >> lane_id()) & 1) keys[base + __popcll(pm[k] & lanemask_lt())] = key[k];
        base += __popcll(pm[k]);
    }
}

// Sorted keys -> heads of the CSR over X tails, and the list length of every
// tail in cnt[0, X) (zeroed by the caller; an exclusive scan makes it the
// offsets).  Runs of equal tails inside a wave add their length with one
// atomic from the run's last lane.
__global__ void k_keys_csr(uint64_t M, int bits, const uint64_t* __restrict__ keys, uint64_t* __restrict__ cnt,
                           uint32_t* __restrict__ cheads, uint32_t* __restrict__ err) {
    const uint64_t mask = (1ull << bits) - 1;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i0 = (uint64_t)blockIdx.x * blockDim.x; i0 < M; i0 += stride) {  // wave-uniform trip count
        const uint64_t i = i0 + threadIdx.x;
        const bool in = i < M;
        const uint64_t k = in ? keys[i] : ~0ull;
        const uint64_t t = k >> bits;
        if (in) {
            cheads[i] = (uint32_t)(k & mask);
            if (i > 0 && keys[i - 1] == k) atomicOr(err, KT_ERR_DUPLICATE);
        }
        const uint64_t tn = __shfl_down(t, 1, 64);
        const unsigned lane = lane_id();
        const bool end = in && (lane == KT_WAVE - 1 || i + 1 == M || tn != t);
        const uint64_t ends = __ballot(end);
        if (end) {
            const uint64_t below = ends & lanemask_lt();
            const int prev = below ? 63 - __clzll((long long)below) : -1;
            atomic_add_u64(&cnt[t], (uint64_t)((int)lane - prev));
        }
    }
}

// First position of a ghost head (combined index >= nl) in each owned list.
__global__ void k_cut_begin(uint32_t nl, const uint64_t* __restrict__ coff, const uint32_t* __restrict__ cheads,
                            uint64_t* __restrict__ cut) {
    uint32_t u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= nl) return;
    uint64_t lo = coff[u], hi = coff[u + 1];
    while (lo < hi) {
        uint64_t mid = (lo + hi) >> 1;
        if (cheads[mid] < nl)
            lo = mid + 1;
        else
            hi = mid;
    }
    cut[u] = lo;
}

// Messages are walked by groups of MSG_G lanes (one message per group; most
// messages are short), each lane taking MSG_E consecutive entries per step so
// that a message of up to 64 ids takes one step, its loads and rank-select
// gathers all in flight together.
constexpr unsigned MSG_G = 8;
constexpr unsigned MSG_E = 8;

__device__ __forceinline__ uint32_t group_sum(uint32_t x) {
#pragma unroll
    for (unsigned d = MSG_G / 2; d >= 1; d >>= 1) x += __shfl_xor(x, d, MSG_G);
    return x;
}

// One step of a group over entries [p0, p0 + MSG_G * MSG_E) of a message
// ending at e: this lane's entries p0 + gl + MSG_G * k (round k covers
// MSG_G consecutive entries, one per lane, so every load and store of a
// round is one contiguous run) and their combined indices (0xffffffff past
// the end or unknown to the rank).
struct MsgStep {
    uint32_t x[MSG_E], c[MSG_E];
    __device__ __forceinline__ static uint64_t at(uint64_t p0, unsigned gl, unsigned k) {
        return p0 + gl + MSG_G * k;
    }
    __device__ __forceinline__ void load(const ShardMap& M, const uint32_t* __restrict__ mheads, uint64_t p0,
                                         unsigned gl, uint64_t e) {
#pragma unroll
        for (unsigned k = 0; k < MSG_E; ++k) x[k] = (at(p0, gl, k) < e) ? mheads[at(p0, gl, k)] : 0u;
#pragma unroll
        for (unsigned k = 0; k < MSG_E; ++k) c[k] = (at(p0, gl, k) < e) ? M.map(x[k]) : 0xffffffffu;
    }
    // the combined indices stored by a previous pass at the entries' positions
    __device__ __forceinline__ void load_mapped(const uint32_t* __restrict__ mapped, uint64_t p0, unsigned gl,
                                                uint64_t e) {
#pragma unroll
        for (unsigned k = 0; k < MSG_E; ++k) c[k] = (at(p0, gl, k) < e) ? mapped[at(p0, gl, k)] : 0xffffffffu;
    }
};

// This group's bits of a wave ballot (lane gl of the group = bit gl).
__device__ __forceinline__ uint32_t group_bits(uint64_t ballot) {
    return (uint32_t)(ballot >> (lane_id() & ~(MSG_G - 1))) & ((1u << MSG_G) - 1);
}

// Per message: owned / ghost entry counts; sender -> combined index.  The
// combined index of every entry is stored at the entry's position in `mapped`
// (one streaming write instead of a second round of rank-select gathers).
__global__ void k_msg_count(ShardMap M, uint32_t nmsg, uint64_t mtot, const uint32_t* __restrict__ src,
                            const uint64_t* __restrict__ moff, const uint32_t* __restrict__ mheads,
                            uint64_t* __restrict__ kept, uint32_t* __restrict__ nown, uint32_t* __restrict__ tail_id,
                            uint32_t* __restrict__ mapped, uint32_t* __restrict__ err) {
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t groups = (uint64_t)gridDim.x * blockDim.x / MSG_G;
    const unsigned gl = lane_id() & (MSG_G - 1);
    uint32_t flags = 0;
    for (uint64_t i = t / MSG_G; i < nmsg; i += groups) {
        const uint64_t b = moff[i], e = moff[i + 1];
        const bool bad = e < b || e > mtot;
        uint32_t own = 0, gh = 0;  // this lane's share
        if (!bad) {
            for (uint64_t p0 = b; p0 < e; p0 += MSG_G * MSG_E) {  // uniform over the group
                MsgStep s;
                s.load(M, mheads, p0, gl, e);
                // the entry after this step's last one (read by the group's last lane)
                const uint64_t pn = p0 + MSG_G * MSG_E;
                const uint32_t xn = (gl == MSG_G - 1 && pn < e) ? mheads[pn] : 0u;
#pragma unroll
                for (unsigned k = 0; k < MSG_E; ++k) {
                    // successor of entry (k, gl): lane gl + 1 of round k, or lane 0
                    // of round k + 1 (shuffles outside any branch: the group is whole)
                    const uint32_t up = __shfl_down(s.x[k], 1, MSG_G);
                    const uint32_t first = __shfl(s.x[k + 1 < MSG_E ? k + 1 : k], 0, MSG_G);
                    const uint32_t nx = gl + 1 < MSG_G ? up : (k + 1 < MSG_E ? first : xn);
                    const uint64_t q = MsgStep::at(p0, gl, k);
                    if (q >= e) continue;
                    mapped[q] = s.c[k];
                    if (q + 1 < e && nx <= s.x[k]) flags |= KT_ERR_MESSAGE;
                    own += s.c[k] < M.nl ? 1u : 0u;
                    gh += (s.c[k] != 0xffffffffu && s.c[k] >= M.nl) ? 1u : 0u;
                }
            }
        }
        own = group_sum(own);
        gh = group_sum(gh);
        if (gl == 0) {
            const uint32_t a = M.map(src[i]);
            if (bad || a == 0xffffffffu || a < M.nl) flags |= KT_ERR_MESSAGE;
            tail_id[i] = (a == 0xffffffffu) ? 0u : a;
            kept[i] = own + gh;
            nown[i] = own;
        }
    }
    or_flags(err, flags);
}

// Per message: owned entries then ghost entries, each in message order.
__global__ void k_msg_write(uint32_t nl, uint32_t nmsg, uint64_t mtot, const uint64_t* __restrict__ moff,
                            const uint32_t* __restrict__ mapped, const uint64_t* __restrict__ coff_msg,
                            const uint32_t* __restrict__ nown, uint32_t* __restrict__ out,
                            uint64_t* __restrict__ split) {
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t groups = (uint64_t)gridDim.x * blockDim.x / MSG_G;
    const unsigned gl = lane_id() & (MSG_G - 1);
    const uint32_t below = (1u << gl) - 1;  // this lane's predecessors in the group
    for (uint64_t i = t / MSG_G; i < nmsg; i += groups) {
        const uint64_t b = moff[i], e = moff[i + 1];
        uint64_t po = coff_msg[i], pg = coff_msg[i] + nown[i];
        if (gl == 0) split[i] = pg;
        if (e < b || e > mtot) continue;  // flagged by k_msg_count; nothing was counted for it (uniform per group)
        for (uint64_t p0 = b; p0 < e; p0 += MSG_G * MSG_E) {
            MsgStep s;
            s.load_mapped(mapped, p0, gl, e);
#pragma unroll
            for (unsigned k = 0; k < MSG_E; ++k) {
                // round k holds consecutive entries in lane order
                const bool o = s.c[k] < nl, h = s.c[k] != 0xffffffffu && s.c[k] >= nl;
                const uint32_t ob = group_bits(__ballot(o)), hb = group_bits(__ballot(h));
                if (o) out[po + __popc(ob & below)] = s.c[k];
                if (h) out[pg + __popc(hb & below)] = s.c[k];
                po += __popc(ob);
                pg += __popc(hb);
            }
        }
    }
}

}  // namespace kt

using namespace kt;

namespace {

int bits_for(uint64_t x) {
    int b = 0;
    while (b < 64 && (x >> b)) ++b;
    return b;
}

struct ShardBufs {
    uint64_t* small;
    const uint64_t* off;
    const uint32_t* heads;
    const uint32_t* ghost_ids;
    const uint64_t* ghost_degree;
    const uint32_t* msg_src;
    const uint64_t* msg_off;
    const uint32_t* msg_heads;
    uint32_t* deg;
    uint64_t* keys;
    uint64_t* alt;
    uint64_t* radix_tmp;
    uint64_t* coff;
    uint32_t* cheads;
    uint64_t* cut;
    uint64_t* mkept;
    uint64_t* moff;
    uint32_t* mnown;
    uint32_t* mtail;
    uint64_t* msplit;
    uint32_t* mheads2;
    uint32_t* mmapped;
    uint64_t* scan_tmp;
    uint64_t* tc_ws;
    uint64_t* pv;
    uint2* rs;
    uint64_t* rs_cnt;
    uint32_t* tbnd;
};


uint64_t shard_layout(void* base, const kt_shard_in* in, uint64_t nnz, uint64_t mtot, bool host_in, bool pv_dev,
                      ShardBufs& B) {
    Carver cv{base, 0};
    const uint64_t nloc = in->range_bounds[in->rank + 1] - in->range_bounds[in->rank];
    const uint64_t X = nloc + in->n_ghost;
    B.small = static_cast<uint64_t*>(cv.take(64 * 8));
    if (host_in) {
        B.off = static_cast<uint64_t*>(cv.take((nloc + 1) * 8));
        B.heads = static_cast<uint32_t*>(cv.take(std::max<uint64_t>(nnz, 1) * 4));
        B.ghost_ids = static_cast<uint32_t*>(cv.take(std::max<uint64_t>(in->n_ghost, 1) * 4));
        B.ghost_degree = static_cast<uint64_t*>(cv.take(std::max<uint64_t>(in->n_ghost, 1) * 8));
        B.msg_src = static_cast<uint32_t*>(cv.take(std::max<uint64_t>(in->n_msg, 1) * 4));
        B.msg_off = static_cast<uint64_t*>(cv.take(((uint64_t)in->n_msg + 1) * 8));
        B.msg_heads = static_cast<uint32_t*>(cv.take(std::max<uint64_t>(mtot, 1) * 4));
    }
    B.deg = static_cast<uint32_t*>(cv.take(std::max<uint64_t>(X, 1) * 4));
    B.keys = static_cast<uint64_t*>(cv.take(std::max<uint64_t>(nnz, 1) * 8));
    B.alt = static_cast<uint64_t*>(cv.take(std::max<uint64_t>(nnz, 1) * 8));
    B.radix_tmp = static_cast<uint64_t*>(cv.take(radix_tmp_words(nnz) * 8));
    B.coff = static_cast<uint64_t*>(cv.take((X + 1) * 8));
    B.cheads = static_cast<uint32_t*>(cv.take((nnz + 4) * 4));
    B.cut = static_cast<uint64_t*>(cv.take(std::max<uint64_t>(nloc, 1) * 8));
    B.mkept = static_cast<uint64_t*>(cv.take(((uint64_t)in->n_msg + 1) * 8));
    B.moff = static_cast<uint64_t*>(cv.take(((uint64_t)in->n_msg + 1) * 8));
    B.mnown = static_cast<uint32_t*>(cv.take(std::max<uint64_t>(in->n_msg, 1) * 4));
    B.mtail = static_cast<uint32_t*>(cv.take(std::max<uint64_t>(in->n_msg, 1) * 4));
    B.msplit = static_cast<uint64_t*>(cv.take(std::max<uint64_t>(in->n_msg, 1) * 8));
    B.mheads2 = static_cast<uint32_t*>(cv.take((mtot + 4) * 4));
    B.mmapped = static_cast<uint32_t*>(cv.take(std::max<uint64_t>(mtot, 1) * 4));
    const uint64_t n_all = in->range_bounds[in->num_ranks];
    const uint64_t rs_words = (n_all + 31) / 32;
    B.scan_tmp = static_cast<uint64_t*>(
        cv.take(scan_tmp_words(std::max(std::max<uint64_t>(X, in->n_msg), rs_words) + 1) * 8));
    B.tc_ws = static_cast<uint64_t*>(cv.take(tc_ws_words((uint32_t)std::max<uint64_t>(X, in->n_msg)) * 8));
    B.pv = pv_dev ? static_cast<uint64_t*>(cv.take(std::max<uint64_t>(X, 1) * 8)) : nullptr;
    B.rs = static_cast<uint2*>(cv.take((rs_words + 1) * 8));
    B.rs_cnt = static_cast<uint64_t*>(cv.take((rs_words + 1) * 8));
    B.tbnd = static_cast<uint32_t*>(cv.take((2 * ((nnz + SH_TILE - 1) / SH_TILE) + 1) * 4));
    return cv.used;
}

}  // namespace

// The body of kt_count_shard; the caller holds c->mu and has begun the call.
int kt::count_shard_locked(kt_ctx* c, const kt_shard_in* in, kt_shard_out* out) {
    int rc = KT_OK;
    if (!in || !out || !in->range_bounds || !in->offsets || in->num_ranks == 0 || in->rank >= in->num_ranks ||
        in->algorithm > KT_SHARD_DITRIC)
        return c->fail(KT_EINVAL, "bad shard arguments");
    // with pseudo2core the reference's senders also drop skipped interface
    // nodes from the distributed phase (cetric_edge_iterator.h:1247-1271),
    // and the messages are this call's inputs: not supported on shards
    if (c->prm.pseudo2core) return c->fail(KT_EINVAL, "pseudo2core applies to the shared-memory path only");
    const uint32_t P = in->num_ranks;
    const uint64_t* bd = in->range_bounds;
    if (bd[0] != 0) return c->fail(KT_EINVAL, "range_bounds[0] != 0");
    for (uint32_t r = 0; r < P; ++r)
        if (bd[r + 1] < bd[r] || bd[r + 1] > 0xffffffffull) return c->fail(KT_EINVAL, "range bounds not ascending");
    const uint32_t from = (uint32_t)bd[in->rank], to = (uint32_t)bd[in->rank + 1];
    const uint32_t nl = to - from, ng = in->n_ghost;
    const uint64_t X = (uint64_t)nl + ng;
    if (X >= 0xffffffffull) return c->fail(KT_EINVAL, "shard too large");
    out->local_triangles = out->type3_triangles = 0;
    const bool host_in = !(in->flags & KT_DEVICE_INPUT);
    const bool want_pv = out->per_vertex != nullptr;
    const bool pv_dev = want_pv && (in->flags & KT_DEVICE_OUTPUT);
    if (nl == 0 && ng == 0) return KT_OK;
    // sizes
    uint64_t nnz = 0, mtot = 0;
    if (host_in) {
        if (in->offsets[0] != 0) return c->fail(KT_EINVAL, "offsets[0] != 0");
        nnz = in->offsets[nl];
        mtot = in->n_msg ? in->msg_offsets[in->n_msg] : 0;
        if (in->n_msg && in->msg_offsets[0] != 0) return c->fail(KT_EINVAL, "msg_offsets[0] != 0");
    } else {
        rc = c->read_u64(in->offsets + nl, &nnz, "D2H nnz");
        if (rc) return rc;
        if (in->n_msg) {
            rc = c->read_u64(in->msg_offsets + in->n_msg, &mtot, "D2H mtot");
            if (rc) return rc;
        }
    }
    if ((nnz && !in->heads) || (ng && (!in->ghost_ids || !in->ghost_degree)) ||
        (in->n_msg && (!in->msg_src || !in->msg_offsets || (mtot && !in->msg_heads))))
        return c->fail(KT_EINVAL, "null shard array");
    ShardBufs B;
    const uint64_t need = shard_layout(nullptr, in, nnz, mtot, host_in, want_pv && !pv_dev, B);
    rc = c->reserve(need);
    if (rc) return rc;
    shard_layout(c->arena, in, nnz, mtot, host_in, want_pv && !pv_dev, B);
    hipStream_t st = c->stream;
    (void)hipEventRecord(c->ev[EV_START], st);
    if (host_in) {
        struct {
            const void* dst;
            const void* src;
            uint64_t bytes;
        } up[] = {{B.off, in->offsets, ((uint64_t)nl + 1) * 8},
                  {B.heads, in->heads, nnz * 4},
                  {B.ghost_ids, in->ghost_ids, (uint64_t)ng * 4},
                  {B.ghost_degree, in->ghost_degree, (uint64_t)ng * 8},
                  {B.msg_src, in->msg_src, (uint64_t)in->n_msg * 4},
                  {B.msg_off, in->msg_offsets, in->n_msg ? ((uint64_t)in->n_msg + 1) * 8 : 0},
                  {B.msg_heads, in->msg_heads, mtot * 4}};
        for (auto& u : up) {
            if (!u.bytes) continue;
            rc = c->hip_check(hipMemcpyAsync(const_cast<void*>(u.dst), u.src, u.bytes, hipMemcpyHostToDevice, st),
                              "H2D shard");
            if (rc) return rc;
        }
    } else {
        B.off = in->offsets;
        B.heads = in->heads;
        B.ghost_ids = in->ghost_ids;
        B.ghost_degree = in->ghost_degree;
        B.msg_src = in->msg_src;
        B.msg_off = in->msg_offsets;
        B.msg_heads = in->msg_heads;
    }
    (void)hipEventRecord(c->ev[EV_H2D], st);
    uint32_t* err = reinterpret_cast<uint32_t*>(B.small);
    uint64_t* nkeys = B.small + 1;
    fill_u64(B.small, 64, 0, st);
    const uint64_t n_all = bd[P];
    ShardMap M{from, to, nl, ng, B.ghost_ids, B.rs, (uint32_t)std::min<uint64_t>(n_all, 0xffffffffull)};
    {
        const uint64_t rs_words = (n_all + 31) / 32;
        fill_u64(reinterpret_cast<uint64_t*>(B.rs), rs_words, 0, st);
        if (ng)
            k_ghost_bits<<<(unsigned)std::min<uint64_t>((ng + 255) / 256, 8192), 256, 0, st>>>(
                ng, B.ghost_ids, M.n_all, reinterpret_cast<uint32_t*>(B.rs), err);
        const unsigned gw = (unsigned)std::max<uint64_t>(1, std::min<uint64_t>((rs_words + 255) / 256, 16384));
        k_rs_count<<<gw, 256, 0, st>>>(rs_words, B.rs, B.rs_cnt);
        scan_u64(B.rs_cnt, B.rs_cnt, rs_words, B.scan_tmp, st);
        k_rs_prefix<<<gw, 256, 0, st>>>(rs_words, B.rs_cnt, B.rs);
    }
    const int bits = std::max(1, bits_for(X));  // X itself is never a valid index: keys stay below 2^(2 bits)
    const bool cetric = in->algorithm == KT_SHARD_CETRIC;
    k_shard_prepare<<<(unsigned)std::min<uint64_t>((X + 255) / 256, 8192), 256, 0, st>>>(M, nnz, B.off,
                                                                                      B.ghost_degree, B.deg, err);
    const uint64_t tiles = (nnz + SH_TILE - 1) / SH_TILE;
    if (tiles && nl) {
        tile_bounds_launch(nl, nnz, SH_TILE, B.off, B.tbnd, st);
        k_shard_entries<<<(unsigned)tiles, SH_THREADS, 0, st>>>(M, nnz, B.off, B.tbnd, B.heads, B.deg, bits,
                                                                cetric ? 1 : 0, B.keys, nkeys, err);
    }
    rc = c->hip_check(hipMemcpyAsync(c->host_small, B.small, 16, hipMemcpyDeviceToHost, st), "D2H flags");
    if (rc) return rc;
    rc = c->hip_check(hipStreamSynchronize(st), "shard entries");
    if (rc) return rc;
    uint32_t flags = (uint32_t)c->host_small[0];
    const uint64_t Mk = c->host_small[1];
    if (flags & KT_ERR_MASK)
        return c->fail(KT_EINVAL, "malformed shard:%s%s%s%s", (flags & KT_ERR_OFFSETS) ? " bad offsets" : "",
                       (flags & KT_ERR_GHOST) ? " ghost table" : "", (flags & KT_ERR_SELF_LOOP) ? " self-loop" : "",
                       (flags & KT_ERR_DUPLICATE) ? " duplicate" : "");
    radix_sort_u64(B.keys, B.alt, Mk, 2 * bits, B.radix_tmp, st);
    fill_u64(B.coff, X + 1, 0, st);
    if (Mk)
        k_keys_csr<<<(unsigned)std::min<uint64_t>((Mk + 255) / 256, 16384), 256, 0, st>>>(Mk, bits, B.keys, B.coff,
                                                                                         B.cheads, err);
    scan_u64(B.coff, B.coff, X, B.scan_tmp, st);
    if (nl) k_cut_begin<<<(nl + 255) / 256, 256, 0, st>>>(nl, B.coff, B.cheads, B.cut);
    // messages
    if (in->n_msg) {
        k_msg_count<<<(unsigned)std::min<uint64_t>(((uint64_t)in->n_msg * MSG_G + 255) / 256, 8192), 256, 0, st>>>(
            M, in->n_msg, mtot, B.msg_src, B.msg_off, B.msg_heads, B.mkept, B.mnown, B.mtail, B.mmapped, err);
        scan_u64(B.mkept, B.moff, in->n_msg, B.scan_tmp, st);
        k_msg_write<<<(unsigned)std::min<uint64_t>(((uint64_t)in->n_msg * MSG_G + 255) / 256, 8192), 256, 0, st>>>(
            M.nl, in->n_msg, mtot, B.msg_off, B.mmapped, B.moff, B.mnown, B.mheads2, B.msplit);
    }
    rc = c->hip_check(hipMemcpyAsync(c->host_small, B.small, 8, hipMemcpyDeviceToHost, st), "D2H flags");
    if (rc) return rc;
    rc = c->hip_check(hipStreamSynchronize(st), "shard preprocessing");
    if (rc) return rc;
    flags = (uint32_t)c->host_small[0];
    if (flags & KT_ERR_MASK)
        return c->fail(KT_EINVAL, "malformed shard:%s%s", (flags & KT_ERR_DUPLICATE) ? " duplicate edge" : "",
                       (flags & KT_ERR_MESSAGE) ? " bad message" : "");
    (void)hipEventRecord(c->ev[EV_PREP], st);
    uint64_t* pv = want_pv ? (pv_dev ? out->per_vertex : B.pv) : nullptr;
    if (pv) fill_u64(pv, X, 0, st);
    uint64_t* totals = B.small + 4;  // [4] local, [5] type3
    uint64_t* tstats = tc_stats_ptr(B.tc_ws);
    // local phase
    {
        TcView v{};
        v.sheads = B.cheads;
        v.s_off = B.coff;
        if (cetric) {
            v.X = (uint32_t)X;  // plain view over owned + ghost tails
        } else {
            v.X = nl;
            v.r_beg = B.coff;
            v.r_end = B.cut;
            v.rheads = B.cheads;
            v.t_beg = B.coff;
            v.t_end = B.coff + 1;
            v.theads = B.cheads;
        }
        TcOut o{pv, totals};
        rc = c->hip_check(tc_launch(v, o, c->prm, B.tc_ws, c->num_cus, st, c->ev[EV_TC0], c->ev[EV_TC1]),
                          "intersection launch");
        if (rc) return rc;
        rc = c->hip_check(hipMemcpyAsync(c->host_small + 8, tstats, 32, hipMemcpyDeviceToHost, st), "D2H stats");
        if (rc) return rc;
    }
    // contracted / distributed phase
    if (in->n_msg) {
        TcView v{};
        v.X = in->n_msg;
        v.sheads = B.mheads2;
        v.rheads = B.mheads2;
        v.r_beg = B.moff;
        v.r_end = B.msplit;
        if (cetric) {
            v.s_beg = B.msplit;
            v.s_end = B.moff + 1;
            v.t_beg = B.cut;
        } else {
            v.s_off = B.moff;
            v.t_beg = B.coff;
        }
        v.t_end = B.coff + 1;
        v.theads = B.cheads;
        v.tail_id = B.mtail;
        // counters as intersect_from_message keeps them (cetric_edge_iterator.h:1815-1856):
        // per local receiver u, |N+(u)| wedge checks and |message| + |N+(u)|
        // intersected ids; a CeTriC receiver without cut edges was compacted away
        v.stat_off = B.msg_off;
        v.stat_skip_empty = cetric ? 1u : 0u;
        TcOut o{pv, totals + 1};
        // message tails have short owned parts: a whole workgroup on one of
        // them idles, so tails up to the largest wave table stay with one
        // wave (measured on RMAT-22 and RMAT-26 / 16 ranks); the workgroup
        // teams on the remaining (large) message tails share their receivers
        // 4 (CeTriC) or 2 (DiTriC) at a time, since a few long lists per
        // message make coarser grabs leave waves idle at the end (RMAT-26 /
        // 16: CeTriC contracted phase 64.6 ms with 8 per grab, 58.4 ms with 4;
        // DiTriC count 76.7 / 70.3 / 68.9 ms with 8 / 4 / 2)
        TcParams pm = c->prm;
        pm.wave_max_deg = std::max<uint32_t>(pm.wave_max_deg, 256u);
        pm.team_chunk = std::min<uint32_t>(pm.team_chunk, cetric ? 4u : 2u);
        rc = c->hip_check(tc_launch(v, o, pm, B.tc_ws, c->num_cus, st, nullptr, nullptr), "intersection launch");
        if (rc) return rc;
        rc = c->hip_check(hipMemcpyAsync(c->host_small + 12, tstats, 32, hipMemcpyDeviceToHost, st), "D2H stats");
        if (rc) return rc;
    }
    (void)hipEventRecord(c->ev[EV_COUNT], st);
    rc = c->hip_check(hipMemcpyAsync(c->host_small + 4, totals, 16, hipMemcpyDeviceToHost, st), "D2H totals");
    if (rc) return rc;
    if (want_pv && !pv_dev) {
        rc = c->hip_check(hipMemcpyAsync(out->per_vertex, pv, X * 8, hipMemcpyDeviceToHost, st), "D2H per-vertex");
        if (rc) return rc;
    }
    (void)hipEventRecord(c->ev[EV_END], st);
    rc = c->hip_check(hipStreamSynchronize(st), "shard count");
    if (rc) return rc;
    out->local_triangles = c->host_small[4];
    out->type3_triangles = in->n_msg ? c->host_small[5] : 0;
    c->stats.local_triangles = out->local_triangles;
    c->stats.type3_triangles = out->type3_triangles;
    c->stats.wedge_checks = c->host_small[8] + (in->n_msg ? c->host_small[12] : 0);
    c->stats.intersection_size_local = c->host_small[9];
    c->stats.intersection_size_global = in->n_msg ? c->host_small[13] : 0;
    c->stats.streamed_ids = c->host_small[11] + (in->n_msg ? c->host_small[15] : 0);
    c->stats.oriented_edges = Mk;
    c->stats.vertices = X;
    c->stats.input_slots = nnz;
    c->timing.h2d_ms = c->elapsed(EV_START, EV_H2D);
    c->timing.prep_ms = c->elapsed(EV_H2D, EV_PREP);
    c->timing.count_ms = c->elapsed(EV_PREP, EV_COUNT);
    c->timing.intersect_ms = c->elapsed(EV_TC0, EV_TC1);
    c->timing.d2h_ms = c->elapsed(EV_COUNT, EV_END);
    c->timing.total_ms = c->elapsed(EV_START, EV_END);
    return KT_OK;
}

extern "C" int kt_count_shard(kt_ctx* c, const kt_shard_in* in, kt_shard_out* out) {
    if (!c) return KT_EINVAL;
    std::lock_guard<std::mutex> lk(c->mu);
    int rc = c->begin_call(true);
    if (rc) return rc;
    return count_shard_locked(c, in, out);
}
