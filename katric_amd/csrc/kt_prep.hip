// Oriented-CSR construction on the device: degree orientation, prefix-scan
// compaction, LSD radix sort (for unsorted inputs and ghost-list grouping).
//
// Replaces DistributedGraph::orient / sort_neighborhoods
// (include/cetric/datastructures/distributed/distributed_graph.h:537-593) and
// the out-only CSR rewrite of remove_in_edges_and_expand_ghosts (:1057-1169)
// for the shared-memory path (counters/shared_memory_cetric.h:71-103).
//
// Layout: the symmetric input CSR is (off[n+1] u64, heads[nnz] u32).  The
// oriented CSR is (ooff[n+1] u64, oheads[nnz/2] u32): for each v the heads u
// with (deg v, v) < (deg u, u), in input order (ascending ids when the input
// lists are sorted; otherwise they are sorted afterwards by radix sort).
#include <algorithm>
#include <random>

#include "kt_device.h"
#include "kt_internal.h"

namespace kt {

// ---------------------------------------------------------------------------
// generic block exclusive scan over 4096-element tiles (256 threads x 16)
// ---------------------------------------------------------------------------
constexpr int SCAN_THREADS = 256;
constexpr int SCAN_ITEMS = 16;
constexpr int SCAN_TILE = SCAN_THREADS * SCAN_ITEMS;

// Exclusive scan of one value per thread across a 256-thread block.
__device__ __forceinline__ uint64_t block_excl_sum(uint64_t x, uint64_t* lds_wave, uint64_t* total) {
    const unsigned w = threadIdx.x / KT_WAVE;
    uint64_t incl = wave_incl_sum(x);
    if (lane_id() == KT_WAVE - 1) lds_wave[w] = incl;
    __syncthreads();
    uint64_t base = 0, tot = 0;
#pragma unroll
    for (unsigned i = 0; i < SCAN_THREADS / KT_WAVE; ++i) {
        uint64_t t = lds_wave[i];
        if (i < w) base += t;
        tot += t;
    }
    __syncthreads();
    *total = tot;
    return base + incl - x;
}

// The tile's sum: the order of the terms does not matter, so the loads are
// coalesced (element tile0 + i * 256 + tid), unlike the scan's per-thread runs.
__global__ void __launch_bounds__(SCAN_THREADS) k_tile_reduce(const uint64_t* __restrict__ in, uint64_t len,
                                                               uint64_t* __restrict__ sums) {
    __shared__ uint64_t lw[SCAN_THREADS / KT_WAVE];
    uint64_t base = (uint64_t)blockIdx.x * SCAN_TILE + threadIdx.x;
    uint64_t s = 0;
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; ++i)
        if (base + i * SCAN_THREADS < len) s += in[base + i * SCAN_THREADS];
    uint64_t tot;
    block_excl_sum(s, lw, &tot);
    if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

// out[i] = tile_base[tile] + exclusive prefix within the tile; out[len] = grand total.
__global__ void __launch_bounds__(SCAN_THREADS) k_tile_scan(const uint64_t* __restrict__ in, uint64_t len,
                                                             const uint64_t* __restrict__ tile_base,
                                                             uint64_t* __restrict__ out) {
    __shared__ uint64_t lw[SCAN_THREADS / KT_WAVE];
    uint64_t base = (uint64_t)blockIdx.x * SCAN_TILE + (uint64_t)threadIdx.x * SCAN_ITEMS;
    uint64_t v[SCAN_ITEMS];
    uint64_t s = 0;
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; ++i) {
        v[i] = (base + i < len) ? in[base + i] : 0;
        s += v[i];
    }
    uint64_t tot;
    uint64_t run = block_excl_sum(s, lw, &tot) + (tile_base ? tile_base[blockIdx.x] : 0);
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; ++i) {
        if (base + i < len) out[base + i] = run;
        run += v[i];
    }
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0)
        out[len] = (tile_base ? tile_base[blockIdx.x] : 0) + tot;
}

uint64_t scan_tmp_words(uint64_t len) {
    uint64_t words = 0;
    while (len > SCAN_TILE) {
        len = (len + SCAN_TILE - 1) / SCAN_TILE;
        words += 2 * (len + 1);
    }
    return words + 2;
}

// Exclusive scan, out has len+1 entries (out[len] = total).  in may alias out.
void scan_u64(const uint64_t* in, uint64_t* out, uint64_t len, uint64_t* tmp, hipStream_t st) {
    if (len == 0) {
        fill_u64(out, 1, 0, st);
        return;
    }
    uint64_t tiles = (len + SCAN_TILE - 1) / SCAN_TILE;
    if (tiles == 1) {
        k_tile_scan<<<1, SCAN_THREADS, 0, st>>>(in, len, nullptr, out);
        return;
    }
    uint64_t* sums = tmp;
    uint64_t* sums_off = tmp + tiles + 1;
    k_tile_reduce<<<(unsigned)tiles, SCAN_THREADS, 0, st>>>(in, len, sums);
    scan_u64(sums, sums_off, tiles, tmp + 2 * (tiles + 1), st);
    k_tile_scan<<<(unsigned)tiles, SCAN_THREADS, 0, st>>>(in, len, sums_off, out);
}

// ---------------------------------------------------------------------------
// degrees + offset validation
// ---------------------------------------------------------------------------
__global__ void k_degree(uint32_t n, uint64_t nnz, const uint64_t* __restrict__ off, uint32_t* __restrict__ deg,
                         uint8_t* __restrict__ ncls, uint32_t* __restrict__ err) {
    // blockDim is a multiple of 64: vertices 2i and 2i+1 sit in one wave
    const uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v == 0 && (off[0] != 0 || off[n] != nnz)) atomicOr(err, KT_ERR_OFFSETS);
    uint32_t c = 0;
    if (v < n) {
        const uint64_t a = off[v], b = off[v + 1];
        uint32_t d = 0;
        if (b < a || b - a > 0xffffffffull)
            atomicOr(err, KT_ERR_OFFSETS);
        else
            d = (uint32_t)(b - a);
        deg[v] = d;
        c = deg_class4(d);
    }
    const uint32_t c1 = (uint32_t)__shfl_down((int)c, 1, KT_WAVE);
    if (v < n && (v & 1) == 0) ncls[v >> 1] = (uint8_t)(c | (v + 1 < n ? c1 << 4 : 0u));
}

// ---------------------------------------------------------------------------
// orientation over 2048-slot tiles of the symmetric CSR
//
// k_orient_keep: slots are visited coalesced (slot = tile0 + k*256 + tid).
// The row of every slot is resolved in LDS: each row starting inside the tile
// marks its first slot, and an inclusive max-scan over the 2048 marks gives
// the row of every slot.  The keep predicate (deg v, v) < (deg u, u) is
// evaluated once per slot and stored as a bitmask (one u64 word per 64
// slots); per-row kept counts d+(v) and the per-tile kept count are written
// for the scans.  k_orient_emit then copies the kept heads in slot order using
// only the bitmask and the tile bases.
// ---------------------------------------------------------------------------
constexpr int OR_THREADS = 256;
constexpr int OR_ITEMS = 8;
constexpr int OR_TILE = OR_THREADS * OR_ITEMS;

// (deg v, v) < (deg u, u): node_ordering::degree_outward with every vertex
// local (distributed_graph.h:147-163).
__device__ __forceinline__ bool keep_edge(uint32_t v, uint32_t dv, uint32_t u, uint32_t du) {
    return dv < du || (dv == du && v < u);
}

// bnd[2t] / bnd[2t+1] = the rows holding the first / last slot of tile t
// (tiles of `tile` slots), written by the row that holds them, so every tile
// resolves its rows without a dependent search (bnd is pre-filled with 0).
__global__ void k_tile_bounds(uint32_t n, uint64_t nnz, uint64_t tile, const uint64_t* __restrict__ off,
                              uint32_t* __restrict__ bnd) {
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x) {
        const uint64_t a = off[v], b = min(off[v + 1], nnz);
        if (b <= a) continue;  // empty row (or offsets the validation will reject)
        for (uint64_t t = a / tile; t * tile < b; ++t) {
            if (t * tile >= a) bnd[2 * t] = v;
            const uint64_t last = min((t + 1) * tile, nnz) - 1;
            if (last >= a && last < b) bnd[2 * t + 1] = v;
        }
    }
}

// Graph bounds of a tile's rows for the batch range check: the graphs of its
// first and last row (from k_tile_graphs), and the vertex range of the first.
struct TileGraphs {
    uint32_t glo, ghi;
    uint64_t a, b;
};

// A head u of row v lies outside v's graph.
__device__ __forceinline__ bool outside_graph(const uint64_t* __restrict__ vbase, const TileGraphs& t, uint32_t v,
                                              uint32_t u) {
    if (t.glo == t.ghi) return u < t.a || u >= t.b;
    const uint32_t g = graph_of(vbase, t.glo, t.ghi, v);
    return u < vbase[g] || u >= vbase[g + 1];
}

// Sets flag bits in the global error word with one atomic per block, and
// only when some bit is not set yet: every wave ORing into the same word
// serialises the whole grid in L2 (the unsorted flag is set by nearly every
// wave of an unsorted input).
__device__ __forceinline__ void flags_flush(uint32_t flags, uint32_t* lds_flags, uint32_t* __restrict__ err) {
    if (flags) atomicOr(lds_flags, flags);
    __syncthreads();
    if (threadIdx.x == 0) {
        const uint32_t f = *lds_flags;
        if (f && (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & f) != f) atomicOr(err, f);
    }
}

struct CheckArgs {
    uint64_t k0, k1;
    const uint64_t* vbase;
    uint32_t G;
    const uint32_t* tgraph;  // batch: graphs of each tile's first and last row (k_tile_graphs)
};

// Graphs of every tile's first and last row, for the batch range check: one
// thread per tile, so the binary searches over vertex_base run side by side
// instead of inside each tile's block.  Rows clamped as tile_rows clamps them.
__global__ void k_tile_graphs(uint32_t n, uint64_t tiles, const uint32_t* __restrict__ tbnd,
                              const uint64_t* __restrict__ vbase, uint32_t G, uint32_t* __restrict__ tg) {
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= tiles) return;
    const uint32_t a = min(tbnd[2 * t], n - 1);
    const uint32_t b = max(a, min(tbnd[2 * t + 1], n - 1));
    const uint32_t glo = graph_of(vbase, 0, G - 1, a);
    tg[2 * t] = glo;
    tg[2 * t + 1] = graph_of(vbase, glo, G - 1, b);
}

template <bool BATCH>
__global__ void __launch_bounds__(OR_THREADS)
    k_orient_keep(uint32_t n, uint64_t nnz, const uint64_t* __restrict__ off, const uint32_t* __restrict__ heads,
                  const uint32_t* __restrict__ deg, const uint8_t* __restrict__ ncls, const uint32_t* __restrict__ tbnd,
                  uint64_t* __restrict__ keep_bits,
                  uint64_t* __restrict__ dplus, uint64_t* __restrict__ tile_cnt, uint32_t* __restrict__ err,
                  CheckArgs ck, uint64_t* __restrict__ sym_part) {
    __shared__ uint32_t marks[OR_TILE];
    __shared__ uint32_t rowcnt[OR_TILE];
    __shared__ uint32_t wmax[OR_THREADS / KT_WAVE];
    __shared__ uint32_t bnd[2];
    __shared__ uint64_t lw[OR_THREADS / KT_WAVE];
    __shared__ uint32_t bflags;
    const uint64_t tile0 = (uint64_t)blockIdx.x * OR_TILE;
    const uint64_t tend = min(tile0 + OR_TILE, nnz);
    const unsigned tid = threadIdx.x;
    for (int i = tid; i < OR_TILE; i += OR_THREADS) rowcnt[i] = 0;
    if (tid == 0) bflags = 0;
    uint32_t lo, hi;
    tile_rows<OR_THREADS, OR_ITEMS>(off, n, tile0, tend, marks, bnd, wmax, lo, hi, tbnd);
    const uint64_t nrows = (uint64_t)hi - lo + 1;
    const bool lds_rows = nrows <= OR_TILE;
    TileGraphs tg{};
    if (BATCH) {
        const uint32_t glo = ck.tgraph[2 * blockIdx.x];
        tg = TileGraphs{glo, ck.tgraph[2 * blockIdx.x + 1], ck.vbase[glo], ck.vbase[glo + 1]};
    }
    uint32_t flags = 0, cnt = 0;
    // The 64 slots of a wave are consecutive: the next slot's head and row
    // come from the next lane (the last lane loads them), and kept slots are
    // counted per run of equal rows with one LDS atomic from the run's last lane.
    const unsigned lane = lane_id();
    // all loads first (the heads, then one four-bit class gather per head
    // from an n/2-byte table, half the L2 footprint of byte classes, and the
    // rows' degrees), so that a thread has every item's memory traffic in
    // flight at once
    // (and, on a wave's last lane, the head of the slot after it, which the
    // adjacency check of that lane compares against)
    uint32_t hu[OR_ITEMS], hn[OR_ITEMS], cu[OR_ITEMS], dv[OR_ITEMS];
#pragma unroll
    for (int k = 0; k < OR_ITEMS; ++k) {
        const uint64_t sl = tile0 + k * OR_THREADS + tid;
        hu[k] = sl < tend ? heads[sl] : 0u;
        hn[k] = (lane == KT_WAVE - 1 && sl + 1 < nnz) ? heads[sl + 1] : 0u;
    }
#pragma unroll
    for (int k = 0; k < OR_ITEMS; ++k) {
        const unsigned i = k * OR_THREADS + tid;
        const bool in = tile0 + i < tend;
        cu[k] = (in && hu[k] < n) ? class4_of(ncls, hu[k]) : 0u;
        dv[k] = in ? deg[lo + marks[i]] : 0u;
    }
    // the pair checksum while the class gathers are in flight: heads and
    // rows are at hand, and the VALU work hides behind the gathers' latency
    uint64_t s0 = 0;
#pragma unroll
    for (int k = 0; k < OR_ITEMS; ++k) {
        const unsigned i = k * OR_THREADS + tid;
        if (tile0 + i < tend && hu[k] < n) s0 += pair_term(lo + marks[i], hu[k], ck.k0);
    }
    // equal classes: the head's exact degree (a second, rarer gather)
    uint32_t du[OR_ITEMS];
#pragma unroll
    for (int k = 0; k < OR_ITEMS; ++k) {
        const unsigned i = k * OR_THREADS + tid;
        du[k] = (tile0 + i < tend && hu[k] < n && cu[k] == deg_class4(dv[k])) ? deg[hu[k]] : 0u;
    }
#pragma unroll
    for (int k = 0; k < OR_ITEMS; ++k) {
        const unsigned i = k * OR_THREADS + tid;
        const uint64_t sl = tile0 + i;
        const bool in = sl < tend;
        bool keep = false;
        const uint32_t rr = in ? marks[i] : 0xffffffffu;
        const uint32_t v = lo + rr;
        const uint32_t u = hu[k];
        if (in) {
            if (u >= n || (BATCH && outside_graph(ck.vbase, tg, v, u))) {
                flags |= KT_ERR_HEAD_RANGE;
            } else {
                if (u == v) flags |= KT_ERR_SELF_LOOP;
                const uint32_t c = deg_class4(dv[k]);
                keep = c != cu[k] ? c < cu[k] : keep_edge(v, dv[k], u, du[k]);
            }
        }
        // next slot in the same row? (shuffles outside any branch: every lane active)
        uint32_t nx = __shfl_down(u, 1, 64);
        const uint32_t rr_next = __shfl_down(rr, 1, 64);
        bool same = rr_next == rr;
        if (lane == KT_WAVE - 1) {
            same = in && ((i + 1 < OR_TILE) ? (sl + 1 < tend && marks[i + 1] == rr)
                                            : (sl + 1 < nnz && sl + 1 < off[v + 1]));
            nx = hn[k];
        }
        if (in && same) {
            if (nx == u)
                flags |= KT_ERR_DUPLICATE;
            else if (nx < u)
                flags |= KT_FLAG_UNSORTED;
        }
        const uint64_t bits = __ballot(keep);
        cnt += keep ? 1u : 0u;
        // run ends: the last lane, the last slot, or the next lane starts another row
        const bool end = in && (lane == KT_WAVE - 1 || sl + 1 >= tend || rr_next != rr);
        const uint64_t ends = __ballot(end);
        if (end) {
            const uint64_t below = ends & lanemask_lt();
            const uint64_t upto = (lane == KT_WAVE - 1) ? ~0ull : ((2ull << lane) - 1);
            const uint64_t run = upto & (below ? ~((2ull << (63 - __clzll((long long)below))) - 1) : ~0ull);
            const uint32_t c = (uint32_t)__popcll(bits & run);
            if (c) {
                if (lds_rows)
                    atomicAdd(&rowcnt[rr], c);
                else
                    atomic_add_u64(&dplus[v], c);
            }
        }
        if (lane == 0 && tile0 + k * OR_THREADS + (tid & ~63u) < tend)
            keep_bits[(tile0 + k * OR_THREADS + (tid & ~63u)) / 64] = bits;
    }
    flags_flush(flags, &bflags, err);
    if (lds_rows) {
        // rows wholly inside the tile get a plain store, the two boundary rows an atomic
        for (uint64_t rr = tid; rr < nrows; rr += OR_THREADS) {
            const uint32_t c = rowcnt[rr];
            const uint32_t v = lo + (uint32_t)rr;
            if (rr == 0 || rr == nrows - 1) {
                if (c) atomic_add_u64(&dplus[v], c);
            } else {
                dplus[v] = c;
            }
        }
    }
    uint64_t tot, stot;
    block_excl_sum(cnt, lw, &tot);
    block_excl_sum(s0, lw, &stot);
    if (tid == 0) {
        tile_cnt[blockIdx.x] = tot;
        sym_part[blockIdx.x] = stot;  // the tile's checksum, summed by k_sym_final
    }
}

// Folds the per-tile pair checksums (tiles entries) and the kept count into
// the error flags: a symmetric CSR keeps exactly half of its slots and its
// checksum is zero.  One 1024-thread block, eight independent loads in flight
// per thread (a batch of 2^31 slots has half a million tiles).
constexpr int SYM_FINAL_THREADS = 1024;
__global__ void __launch_bounds__(SYM_FINAL_THREADS) k_sym_final(const uint64_t* __restrict__ sym_part,
                                                                 uint64_t tiles, const uint64_t* __restrict__ kept,
                                                                 uint64_t nnz, uint32_t* __restrict__ err) {
    constexpr int U = 8;
    __shared__ uint64_t lw[SYM_FINAL_THREADS / KT_WAVE];
    uint64_t s[U] = {};
    uint64_t i = threadIdx.x;
    for (; i + (U - 1) * SYM_FINAL_THREADS < tiles; i += U * SYM_FINAL_THREADS)
#pragma unroll
        for (int k = 0; k < U; ++k) s[k] += sym_part[i + (uint64_t)k * SYM_FINAL_THREADS];
    for (; i < tiles; i += SYM_FINAL_THREADS) s[0] += sym_part[i];
#pragma unroll
    for (int k = 1; k < U; ++k) s[0] += s[k];
    const uint64_t w = wave_incl_sum(s[0]);
    if (lane_id() == KT_WAVE - 1) lw[threadIdx.x / KT_WAVE] = w;
    __syncthreads();
    if (threadIdx.x != 0) return;
    uint64_t tot = 0;
#pragma unroll
    for (int k = 0; k < SYM_FINAL_THREADS / KT_WAVE; ++k) tot += lw[k];
    bool bad = tot != 0;
    if (kept) bad |= 2 * kept[0] != nnz;
    if (bad) atomicOr(err, KT_ERR_ASYMMETRIC);
}

// Validation of a whole CSR without orienting it (rows resolved per tile as
// in k_orient_keep): head range, self-loops, adjacent duplicates of sorted
// rows, and the pair checksums.
__global__ void __launch_bounds__(OR_THREADS)
    k_sym_tiles(uint32_t n, uint64_t nnz, const uint64_t* __restrict__ off, const uint32_t* __restrict__ heads,
                const uint32_t* __restrict__ tbnd, uint32_t* __restrict__ err, CheckArgs ck,
                uint64_t* __restrict__ sym_part) {
    __shared__ uint32_t marks[OR_TILE];
    __shared__ uint32_t wmax[OR_THREADS / KT_WAVE];
    __shared__ uint32_t bnd[2];
    __shared__ uint64_t lw[OR_THREADS / KT_WAVE];
    __shared__ uint32_t bflags;
    const uint64_t tile0 = (uint64_t)blockIdx.x * OR_TILE;
    const uint64_t tend = min(tile0 + OR_TILE, nnz);
    const unsigned tid = threadIdx.x, lane = lane_id();
    if (tid == 0) bflags = 0;
    uint32_t lo, hi;
    tile_rows<OR_THREADS, OR_ITEMS>(off, n, tile0, tend, marks, bnd, wmax, lo, hi, tbnd);
    uint32_t flags = 0;
    uint64_t s0 = 0;
#pragma unroll 4
    for (int k = 0; k < OR_ITEMS; ++k) {
        const unsigned i = k * OR_THREADS + tid;
        const uint64_t sl = tile0 + i;
        const bool in = sl < tend;
        const uint32_t rr = in ? marks[i] : 0xffffffffu;
        const uint32_t v = lo + rr;
        const uint32_t u = in ? heads[sl] : 0u;
        if (in) {
            if (u >= n) {
                flags |= KT_ERR_HEAD_RANGE;
            } else {
                if (u == v) flags |= KT_ERR_SELF_LOOP;
                s0 += pair_term(v, u, ck.k0);
            }
        }
        uint32_t nx = __shfl_down(u, 1, 64);
        bool same = __shfl_down(rr, 1, 64) == rr;
        if (lane == KT_WAVE - 1) {
            same = in && ((i + 1 < OR_TILE) ? (sl + 1 < tend && marks[i + 1] == rr)
                                            : (sl + 1 < nnz && sl + 1 < off[v + 1]));
            if (same) nx = heads[sl + 1];
        }
        if (in && same) {
            if (nx == u)
                flags |= KT_ERR_DUPLICATE;
            else if (nx < u)
                flags |= KT_FLAG_UNSORTED;
        }
    }
    flags_flush(flags, &bflags, err);
    uint64_t stot;
    block_excl_sum(s0, lw, &stot);
    if (tid == 0) sym_part[blockIdx.x] = stot;
}

// Copies the kept heads in slot order: output position = tile base + rank of
// the slot among the tile's kept slots.  Every wave loads its bitmask words
// and all its heads first (coalesced whole rows of 64 slots; the dropped
// heads share their cache lines with the kept ones), so the loads are in
// flight together; one barrier publishes the per-(round, wave) counts, from
// which each wave derives all its output positions.
__global__ void __launch_bounds__(OR_THREADS)
    k_orient_emit(uint64_t nnz, const uint32_t* __restrict__ heads, const uint64_t* __restrict__ keep_bits,
                  const uint64_t* __restrict__ tile_off, uint32_t* __restrict__ oheads) {
    constexpr int NW = OR_THREADS / KT_WAVE;
    __shared__ uint32_t wcnt[OR_ITEMS][NW];
    const uint64_t tile0 = (uint64_t)blockIdx.x * OR_TILE;
    const unsigned tid = threadIdx.x, wave = tid / KT_WAVE, lane = lane_id();
    uint64_t bits[OR_ITEMS];
    uint32_t hv[OR_ITEMS];
#pragma unroll
    for (int k = 0; k < OR_ITEMS; ++k) {
        const uint64_t w0 = tile0 + k * OR_THREADS + wave * KT_WAVE;
        bits[k] = (w0 < nnz) ? keep_bits[w0 / 64] : 0;
        hv[k] = (w0 + lane < nnz) ? heads[w0 + lane] : 0u;
    }
    uint64_t base = tile_off[blockIdx.x];
#pragma unroll
    for (int k = 0; k < OR_ITEMS; ++k)
        if (lane == 0) wcnt[k][wave] = (uint32_t)__popcll(bits[k]);
    __syncthreads();
#pragma unroll
    for (int k = 0; k < OR_ITEMS; ++k) {
        uint32_t before = 0, round = 0;
#pragma unroll
        for (unsigned w = 0; w < NW; ++w) {
            const uint32_t c = wcnt[k][w];
            before += (w < wave) ? c : 0;
            round += c;
        }
        if ((bits[k] >> lane) & 1) oheads[base + before + __popcll(bits[k] & lanemask_lt())] = hv[k];
        base += round;
    }
}

// Rebuilds (row<<32 | head) keys of the oriented CSR (unsorted input path).
__global__ void k_make_keys(uint32_t n, uint64_t m, const uint64_t* __restrict__ ooff,
                            const uint32_t* __restrict__ oheads, uint64_t* __restrict__ keys) {
    uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= m) return;
    uint32_t r = row_of_slot(ooff, 0, n - 1, p);
    keys[p] = ((uint64_t)r << 32) | oheads[p];
}

// After sorting (row<<32|head) keys: extract heads, flag duplicates.
__global__ void k_keys_to_heads(uint64_t m, const uint64_t* __restrict__ keys, uint32_t* __restrict__ oheads,
                                uint32_t* __restrict__ err) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    uint64_t k = keys[i];
    oheads[i] = (uint32_t)k;
    if (i + 1 < m && keys[i + 1] == k) atomicOr(err, KT_ERR_DUPLICATE);
}

// ---------------------------------------------------------------------------
// segmented sort of the oriented lists (unsorted input)
//
// Every list N+(v) = oheads[ooff[v] .. ooff[v+1]) is sorted in place; lists
// never mix, so no keys with row ids and no second buffer are needed.  By
// length: up to 16 ids a thread sorts its own list in registers (a bitonic
// network over 16 slots padded with the largest id); 17..1024 ids a wave
// sorts the list in registers, K = 1..16 ids per lane (slot i = r * 64 +
// lane: merge steps of distance >= 64 stay inside a lane, shorter ones are
// lane shuffles); 1025..SEG_MAX ids a workgroup sorts it in LDS (bitonic over
// the next power of two).  Longer lists go to the global radix sort.  Equal
// neighbours after the sort are duplicate edges.
// ---------------------------------------------------------------------------
constexpr uint32_t SEG_THREAD_MAX = 16;
constexpr uint32_t SEG_WAVE_MAX = 1024;
constexpr uint32_t SEG_MAX = 4096;
constexpr int SEG_BLOCK = 256;

__device__ __forceinline__ void cas(uint32_t& a, uint32_t& b) {
    const uint32_t lo = min(a, b), hi = max(a, b);
    a = lo;
    b = hi;
}

// Sorts p[0 .. len) (len <= 64 K) with the whole wave; returns the duplicate
// flag of this lane's slots.
template <int K>
__device__ __forceinline__ uint32_t wave_sort_list(uint32_t* __restrict__ p, uint32_t len, uint32_t lane) {
    constexpr uint32_t P = 64 * K;
    uint32_t x[K];
#pragma unroll
    for (int r = 0; r < K; ++r) {
        const uint32_t i = r * 64 + lane;
        x[r] = i < len ? p[i] : 0xffffffffu;
    }
#pragma unroll
    for (uint32_t k = 2; k <= P; k <<= 1)
#pragma unroll
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            if (j >= 64) {
                // partner slot i ^ j is register r ^ (j / 64) of this lane;
                // the direction bit (i & k) depends on r only
#pragma unroll
                for (int r = 0; r < K; ++r) {
                    const int q = r ^ (int)(j / 64);
                    if (q > r) {
                        if ((r & (int)(k / 64)) == 0)
                            cas(x[r], x[q]);
                        else
                            cas(x[q], x[r]);
                    }
                }
            } else {
                const bool lower = (lane & j) == 0;
#pragma unroll
                for (int r = 0; r < K; ++r) {
                    const uint32_t y = (uint32_t)__shfl_xor((int)x[r], (int)j, KT_WAVE);
                    const bool up = k < 64 ? (lane & k) == 0 : (r & (int)(k / 64)) == 0;
                    x[r] = (lower == up) ? min(x[r], y) : max(x[r], y);
                }
            }
        }
    uint32_t f = 0;
#pragma unroll
    for (int r = 0; r < K; ++r) {
        const uint32_t i = r * 64 + lane;
        if (i < len) p[i] = x[r];
        // the next slot: the next lane's register r, or lane 0's register r + 1
        uint32_t nx = (uint32_t)__shfl_down((int)x[r], 1, KT_WAVE);
        if (r + 1 < K) {
            const uint32_t w = (uint32_t)__builtin_amdgcn_readfirstlane((int)x[r + 1 < K ? r + 1 : r]);
            if (lane == 63) nx = w;
        }
        if (i + 1 < len && nx == x[r]) f = KT_ERR_DUPLICATE;
    }
    return f;
}

// Rows of 64-row chunks, one chunk per wave: lists of <= 16 ids sorted by
// their own lane, then the chunk's lists of 17..1024 ids one after another by
// the whole wave.
__global__ void __launch_bounds__(SEG_BLOCK) k_segsort_small(uint32_t n, const uint64_t* __restrict__ ooff,
                                                             uint32_t* __restrict__ oheads,
                                                             uint32_t* __restrict__ err) {
    const uint32_t lane = lane_id();
    const uint64_t nw = (uint64_t)gridDim.x * (SEG_BLOCK / KT_WAVE);
    uint32_t flags = 0;
    for (uint64_t c = (uint64_t)blockIdx.x * (SEG_BLOCK / KT_WAVE) + threadIdx.x / KT_WAVE; c * KT_WAVE < n; c += nw) {
        const uint64_t v = c * KT_WAVE + lane;
        uint64_t b = 0, e = 0;
        if (v < n) {
            b = ooff[v];
            e = ooff[v + 1];
        }
        const uint32_t len = (uint32_t)(e - b);
        if (len >= 2 && len <= SEG_THREAD_MAX) {
            uint32_t a[SEG_THREAD_MAX];
#pragma unroll
            for (uint32_t j = 0; j < SEG_THREAD_MAX; ++j) a[j] = j < len ? oheads[b + j] : 0xffffffffu;
            // bitonic network on 16 slots (static indices: stays in registers)
#pragma unroll
            for (uint32_t k = 2; k <= SEG_THREAD_MAX; k <<= 1)
#pragma unroll
                for (uint32_t j = k >> 1; j > 0; j >>= 1)
#pragma unroll
                    for (uint32_t i = 0; i < SEG_THREAD_MAX; ++i) {
                        const uint32_t l = i ^ j;
                        if (l > i) {
                            if ((i & k) == 0)
                                cas(a[i], a[l]);
                            else
                                cas(a[l], a[i]);
                        }
                    }
#pragma unroll
            for (uint32_t j = 0; j < SEG_THREAD_MAX; ++j) {
                if (j < len) oheads[b + j] = a[j];
                if (j + 1 < len && a[j] == a[j + 1]) flags |= KT_ERR_DUPLICATE;
            }
        }
        // the chunk's lists of 17..1024 ids, one at a time, by the whole wave
        uint64_t mid = __ballot(len > SEG_THREAD_MAX && len <= SEG_WAVE_MAX);
        while (mid) {
            const int r = __ffsll((unsigned long long)mid) - 1;
            mid &= mid - 1;
            const uint64_t rb = (uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)b, r) |
                                ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(b >> 32), r) << 32);
            const uint32_t rl = (uint32_t)__builtin_amdgcn_readlane((int)len, r);
            uint32_t* p = oheads + rb;
            if (rl <= 64)
                flags |= wave_sort_list<1>(p, rl, lane);
            else if (rl <= 128)
                flags |= wave_sort_list<2>(p, rl, lane);
            else if (rl <= 256)
                flags |= wave_sort_list<4>(p, rl, lane);
            else if (rl <= 512)
                flags |= wave_sort_list<8>(p, rl, lane);
            else
                flags |= wave_sort_list<16>(p, rl, lane);
        }
    }
    or_flags(err, flags);
}

// Lists of 1025..SEG_MAX ids: 256-row chunks per block, each such list sorted
// by the whole block in LDS (bitonic over the next power of two).
__global__ void __launch_bounds__(SEG_BLOCK) k_segsort_large(uint32_t n, const uint64_t* __restrict__ ooff,
                                                             uint32_t* __restrict__ oheads,
                                                             uint32_t* __restrict__ err) {
    __shared__ uint32_t s[SEG_MAX];
    __shared__ uint32_t rows[SEG_BLOCK];
    __shared__ uint32_t nrows;
    __shared__ uint32_t bflags;
    const uint32_t tid = threadIdx.x;
    if (tid == 0) bflags = 0;
    for (uint64_t c0 = (uint64_t)blockIdx.x * SEG_BLOCK; c0 < n; c0 += (uint64_t)gridDim.x * SEG_BLOCK) {
        if (tid == 0) nrows = 0;
        __syncthreads();
        const uint64_t v = c0 + tid;
        if (v < n) {
            const uint64_t len = ooff[v + 1] - ooff[v];
            if (len > SEG_WAVE_MAX && len <= SEG_MAX) rows[atomicAdd(&nrows, 1u)] = (uint32_t)v;
        }
        __syncthreads();
        const uint32_t cnt = nrows;
        for (uint32_t r = 0; r < cnt; ++r) {
            const uint32_t row = rows[r];
            const uint64_t b = ooff[row];
            const uint32_t len = (uint32_t)(ooff[row + 1] - b);
            uint32_t P = 2 * SEG_WAVE_MAX;
            while (P < len) P <<= 1;
            for (uint32_t i = tid; i < P; i += SEG_BLOCK) s[i] = i < len ? oheads[b + i] : 0xffffffffu;
            __syncthreads();
            for (uint32_t k = 2; k <= P; k <<= 1)
                for (uint32_t j = k >> 1; j > 0; j >>= 1) {
                    for (uint32_t i = tid; i < P; i += SEG_BLOCK) {
                        const uint32_t l = i ^ j;
                        if (l > i) {
                            const uint32_t a = s[i], c = s[l];
                            const bool up = (i & k) == 0;
                            if ((a > c) == up) {
                                s[i] = c;
                                s[l] = a;
                            }
                        }
                    }
                    __syncthreads();
                }
            uint32_t f = 0;
            for (uint32_t i = tid; i < len; i += SEG_BLOCK) {
                oheads[b + i] = s[i];
                if (i + 1 < len && s[i] == s[i + 1]) f = KT_ERR_DUPLICATE;
            }
            if (f) atomicOr(&bflags, f);
            __syncthreads();
        }
        __syncthreads();  // every thread has read nrows before it is reset
    }
    if (tid == 0 && bflags) or_flags(err, bflags);
}

// Largest oriented list (partials per block, then one block).
__global__ void __launch_bounds__(256) k_max_len(uint32_t n, const uint64_t* __restrict__ ooff,
                                                 uint32_t* __restrict__ part) {
    uint32_t mx = 0;
    for (uint64_t v = (uint64_t)blockIdx.x * 256 + threadIdx.x; v < n; v += (uint64_t)gridDim.x * 256)
    {
        const uint64_t d = ooff[v + 1] - ooff[v];
        mx = max(mx, d > 0xffffffffull ? 0xffffffffu : (uint32_t)d);
    }
    mx = wave_incl_max_u32(mx);
    __shared__ uint32_t w[4];
    if (lane_id() == KT_WAVE - 1) w[threadIdx.x / KT_WAVE] = mx;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = max(max(w[0], w[1]), max(w[2], w[3]));
}

__global__ void __launch_bounds__(256) k_max_final(uint32_t parts, const uint32_t* __restrict__ part,
                                                   uint64_t* __restrict__ out) {
    uint32_t mx = 0;
    for (uint32_t i = threadIdx.x; i < parts; i += 256) mx = max(mx, part[i]);
    mx = wave_incl_max_u32(mx);
    __shared__ uint32_t w[4];
    if (lane_id() == KT_WAVE - 1) w[threadIdx.x / KT_WAVE] = mx;
    __syncthreads();
    if (threadIdx.x == 0) *out = max(max(w[0], w[1]), max(w[2], w[3]));
}

constexpr uint32_t MAXLEN_PARTS = 1024;

uint64_t segsort_tmp_words() { return MAXLEN_PARTS / 2 + 1; }

void max_len_launch(uint32_t n, const uint64_t* ooff, uint64_t* tmp, uint64_t* out, hipStream_t st) {
    uint32_t* part = reinterpret_cast<uint32_t*>(tmp);
    const uint32_t g = (uint32_t)std::min<uint64_t>(MAXLEN_PARTS, ((uint64_t)n + 255) / 256 + 1);
    k_max_len<<<g, 256, 0, st>>>(n, ooff, part);
    k_max_final<<<1, 256, 0, st>>>(g, part, out);
}

uint32_t segsort_max_len() { return SEG_MAX; }

void segsort_launch(uint32_t n, const uint64_t* ooff, uint32_t* oheads, uint32_t* err, uint64_t max_len, int num_cus,
                    hipStream_t st) {
    if (n == 0 || max_len < 2) return;
    const uint64_t chunks = ((uint64_t)n + KT_WAVE - 1) / KT_WAVE;
    const uint32_t g1 = (uint32_t)std::min<uint64_t>((chunks + 3) / 4, (uint64_t)num_cus * 32);
    k_segsort_small<<<g1, SEG_BLOCK, 0, st>>>(n, ooff, oheads, err);
    if (max_len <= SEG_WAVE_MAX) return;
    const uint32_t g2 = (uint32_t)std::min<uint64_t>(((uint64_t)n + SEG_BLOCK - 1) / SEG_BLOCK, (uint64_t)num_cus * 8);
    k_segsort_large<<<g2, SEG_BLOCK, 0, st>>>(n, ooff, oheads, err);
}

// ---------------------------------------------------------------------------
// LSD radix sort of u64 keys, 8-bit digits, stable
// ---------------------------------------------------------------------------
constexpr int RS_THREADS = 256;
constexpr int RS_ROUNDS = 16;
constexpr int RS_TILE = RS_THREADS * RS_ROUNDS;

__global__ void __launch_bounds__(RS_THREADS) k_radix_hist(const uint64_t* __restrict__ keys, uint64_t len,
                                                            int shift, uint64_t* __restrict__ hist,
                                                            uint32_t ntiles) {
    __shared__ uint32_t h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    uint64_t base = (uint64_t)blockIdx.x * RS_TILE;
    for (int i = 0; i < RS_ROUNDS; ++i) {
        uint64_t idx = base + (uint64_t)i * RS_THREADS + threadIdx.x;
        if (idx < len) atomicAdd(&h[(keys[idx] >> shift) & 255], 1u);
    }
    __syncthreads();
    hist[(uint64_t)threadIdx.x * ntiles + blockIdx.x] = h[threadIdx.x];
}

// Stable scatter of one 4096-key tile by the digit at `shift`: the tile is
// first ranked into digit order in LDS (rounds of 256 keys in index order;
// within a round, waves in order and lanes in order, so equal digits keep
// their input order), then written out from LDS in that order, so that
// consecutive threads store consecutive positions of one digit's run.
__global__ void __launch_bounds__(RS_THREADS) k_radix_scatter(const uint64_t* __restrict__ in, uint64_t len,
                                                               int shift, const uint64_t* __restrict__ hist,
                                                               const uint64_t* __restrict__ hist_off,
                                                               uint32_t ntiles, uint64_t* __restrict__ out) {
    __shared__ uint64_t sk[RS_TILE];
    __shared__ uint64_t gbase[256];                   // global start of this tile's run of digit d
    __shared__ uint32_t dstart[256];                  // local start of digit d in the ranked tile
    __shared__ uint32_t run[256];                     // keys of digit d placed so far
    __shared__ uint32_t wcnt[RS_THREADS / KT_WAVE][256];
    __shared__ uint64_t lw[RS_THREADS / KT_WAVE];
    const unsigned t = threadIdx.x, w = t / KT_WAVE;
    const uint64_t base = (uint64_t)blockIdx.x * RS_TILE;
    const uint32_t nt = (uint32_t)min<uint64_t>(RS_TILE, len - base);
    {
        const uint64_t c = hist[(uint64_t)t * ntiles + blockIdx.x];
        uint64_t tot;
        dstart[t] = (uint32_t)block_excl_sum(c, lw, &tot);
        gbase[t] = hist_off[(uint64_t)t * ntiles + blockIdx.x];
        run[t] = 0;
    }
    for (int i = 0; i < RS_ROUNDS; ++i) {
        for (int q = 0; q < RS_THREADS / KT_WAVE; ++q) wcnt[q][t] = 0;
        __syncthreads();
        const uint32_t li = (uint32_t)i * RS_THREADS + t;
        const bool valid = li < nt;
        const uint64_t k = valid ? in[base + li] : 0;
        const uint32_t d = (uint32_t)((k >> shift) & 255);
        // lanes of this wave holding the same digit
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const uint64_t bb = __ballot((d >> b) & 1);
            peers &= ((d >> b) & 1) ? bb : ~bb;
        }
        const uint32_t wrank = __popcll(peers & lanemask_lt());
        if (valid && wrank == 0) wcnt[w][d] = __popcll(peers);
        __syncthreads();
        // per digit: waves' offsets within this round, and the digit's new total
        {
            uint32_t r = run[t];
#pragma unroll
            for (int q = 0; q < RS_THREADS / KT_WAVE; ++q) {
                const uint32_t c = wcnt[q][t];
                wcnt[q][t] = r;
                r += c;
            }
            run[t] = r;
        }
        __syncthreads();
        if (valid) sk[dstart[d] + wcnt[w][d] + wrank] = k;
        __syncthreads();  // wcnt is cleared for the next round
    }
    for (uint32_t p = t; p < nt; p += RS_THREADS) {
        const uint64_t k = sk[p];
        const uint32_t d = (uint32_t)((k >> shift) & 255);
        out[gbase[d] + (p - dstart[d])] = k;
    }
}

uint64_t radix_tmp_words(uint64_t len) {
    uint64_t ntiles = (len + RS_TILE - 1) / RS_TILE;
    uint64_t h = 256 * ntiles;
    return 2 * (h + 1) + scan_tmp_words(h);
}

// Sorts keys[0..len) by bits [0, bits); result ends in `keys` (alt is scratch
// of len words, tmp of radix_tmp_words(len) words).
void radix_sort_u64(uint64_t* keys, uint64_t* alt, uint64_t len, int bits, uint64_t* tmp, hipStream_t st) {
    if (len <= 1) return;
    uint32_t ntiles = (uint32_t)((len + RS_TILE - 1) / RS_TILE);
    uint64_t h = 256ull * ntiles;
    uint64_t* hist = tmp;
    uint64_t* hoff = tmp + h + 1;
    uint64_t* stmp = tmp + 2 * (h + 1);
    int passes = (bits + 7) / 8;
    uint64_t* src = keys;
    uint64_t* dst = alt;
    for (int p = 0; p < passes; ++p) {
        int shift = 8 * p;
        k_radix_hist<<<ntiles, RS_THREADS, 0, st>>>(src, len, shift, hist, ntiles);
        scan_u64(hist, hoff, h, stmp, st);
        k_radix_scatter<<<ntiles, RS_THREADS, 0, st>>>(src, len, shift, hist, hoff, ntiles, dst);
        uint64_t* t = src;
        src = dst;
        dst = t;
    }
    if (src != keys) (void)hipMemcpyAsync(keys, src, len * sizeof(uint64_t), hipMemcpyDeviceToDevice, st);
}

// ---------------------------------------------------------------------------
// host driver for the plain orientation
// ---------------------------------------------------------------------------
struct OrientWs {
    uint64_t tiles;
    uint64_t* sym;       // tiles: per-tile pair checksums
    uint64_t* tile_cnt;  // tiles + 1
    uint64_t* tile_off;  // tiles + 1 (tile_off[tiles] = kept total)
    uint64_t* bits;      // ceil(nnz / 64) + 1
    uint64_t* dplus;     // n + 1
    uint32_t* tbnd;      // 2 * tiles: rows of each tile's first and last slot
    uint64_t* stmp;
    uint8_t* ncls;       // (n + 1) / 2: four-bit degree classes
    uint32_t* tgraph;    // 2 * tiles: batch, graphs of each tile's first and last row
};

static OrientWs orient_ws(uint32_t n, uint64_t nnz, uint64_t* ws) {
    OrientWs w;
    w.tiles = (nnz + OR_TILE - 1) / OR_TILE;
    w.sym = ws;
    w.tile_cnt = ws + w.tiles + 1;
    w.tile_off = w.tile_cnt + w.tiles + 1;
    w.bits = w.tile_off + w.tiles + 1;
    w.dplus = w.bits + (nnz + 63) / 64 + 1;
    w.tbnd = reinterpret_cast<uint32_t*>(w.dplus + (uint64_t)n + 1);
    w.stmp = w.dplus + (uint64_t)n + 1 + w.tiles;
    uint64_t* after = w.stmp + std::max(scan_tmp_words(w.tiles + 1), scan_tmp_words((uint64_t)n + 1));
    w.ncls = reinterpret_cast<uint8_t*>(after);
    w.tgraph = reinterpret_cast<uint32_t*>(after + ((uint64_t)n + 15) / 16);
    return w;
}

const uint64_t* orient_keep_bits(uint32_t n, uint64_t nnz, const uint64_t* ws) {
    return orient_ws(n, nnz, const_cast<uint64_t*>(ws)).bits;
}

uint64_t orient_ws_words(uint32_t n, uint64_t nnz) {
    uint64_t tiles = (nnz + OR_TILE - 1) / OR_TILE;
    return (tiles + 1) + 2 * (tiles + 1) + (nnz + 63) / 64 + 1 + (uint64_t)n + 1 + tiles +
           std::max(scan_tmp_words(tiles + 1), scan_tmp_words((uint64_t)n + 1)) + ((uint64_t)n + 15) / 16 + tiles;
}

// Rows of the first and last slot of every `tile`-slot tile of a CSR whose
// offsets start at 0 (bnd: 2 entries per tile), for tile_rows.
void tile_bounds_launch(uint32_t n, uint64_t nnz, uint64_t tile, const uint64_t* off, uint32_t* bnd, hipStream_t st) {
    const uint64_t tiles = (nnz + tile - 1) / tile;
    if (!tiles || !n) return;
    fill_u32(bnd, 2 * tiles, 0, st);
    k_tile_bounds<<<(unsigned)(((uint64_t)n + 255) / 256), 256, 0, st>>>(n, nnz, tile, off, bnd);
}

// Degrees, validation flags, keep bitmask, d+ per row and its scan into ooff
// (ooff[n] = kept total), per-tile kept counts and their scan.
void orient_count_launch(uint32_t n, uint64_t nnz, const uint64_t* off, const uint32_t* heads, uint32_t* deg,
                         uint64_t* ooff, uint64_t* ws, uint32_t* err, const OrientCheck& chk, hipStream_t st) {
    OrientWs w = orient_ws(n, nnz, ws);
    k_degree<<<(n + 255) / 256, 256, 0, st>>>(n, nnz, off, deg, w.ncls, err);
    fill_u64(w.dplus, (uint64_t)n + 1, 0, st);
    if (w.tiles > 0) {
        const CheckArgs ck{chk.key[0], chk.key[1], chk.G ? chk.vbase : nullptr, chk.G, w.tgraph};
        fill_u32(w.tbnd, 2 * w.tiles, 0, st);
        k_tile_bounds<<<(n + 255) / 256, 256, 0, st>>>(n, nnz, OR_TILE, off, w.tbnd);
        if (ck.vbase) {
            k_tile_graphs<<<(unsigned)((w.tiles + 255) / 256), 256, 0, st>>>(n, w.tiles, w.tbnd, ck.vbase, ck.G,
                                                                            w.tgraph);
            k_orient_keep<true><<<(unsigned)w.tiles, OR_THREADS, 0, st>>>(n, nnz, off, heads, deg, w.ncls, w.tbnd,
                                                                         w.bits, w.dplus, w.tile_cnt, err, ck, w.sym);
        } else {
            k_orient_keep<false><<<(unsigned)w.tiles, OR_THREADS, 0, st>>>(n, nnz, off, heads, deg, w.ncls, w.tbnd,
                                                                          w.bits, w.dplus, w.tile_cnt, err, ck, w.sym);
        }
        scan_u64(w.tile_cnt, w.tile_off, w.tiles, w.stmp, st);
    } else {
        fill_u64(w.tile_off, 1, 0, st);
    }
    scan_u64(w.dplus, ooff, n, w.stmp, st);
    k_sym_final<<<1, SYM_FINAL_THREADS, 0, st>>>(w.sym, w.tiles, ooff + n, nnz, err);
}

uint64_t sym_ws_words(uint32_t n, uint64_t nnz) {
    (void)n;
    const uint64_t tiles = (nnz + OR_TILE - 1) / OR_TILE;
    return 2 * (tiles + 1);  // per-tile checksums, then the tiles' row bounds
}

void sym_check_launch(uint32_t n, uint64_t nnz, const uint64_t* off, const uint32_t* heads, uint64_t* ws,
                      uint32_t* err, const OrientCheck& chk, hipStream_t st) {
    const uint64_t tiles = (nnz + OR_TILE - 1) / OR_TILE;
    uint64_t* sym = ws;
    uint32_t* tbnd = reinterpret_cast<uint32_t*>(ws + tiles + 1);
    if (tiles > 0 && n > 0) {
        const CheckArgs ck{chk.key[0], chk.key[1], nullptr, 0, nullptr};
        fill_u32(tbnd, 2 * tiles, 0, st);
        k_tile_bounds<<<(n + 255) / 256, 256, 0, st>>>(n, nnz, OR_TILE, off, tbnd);
        k_sym_tiles<<<(unsigned)tiles, OR_THREADS, 0, st>>>(n, nnz, off, heads, tbnd, err, ck, sym);
    }
    k_sym_final<<<1, SYM_FINAL_THREADS, 0, st>>>(sym, (tiles > 0 && n > 0) ? tiles : 0, nullptr, nnz, err);
}

void random_keys(uint64_t key[2]) {
    std::random_device rd;
    for (int i = 0; i < 2; ++i) {
        uint64_t k = 0;
        for (int j = 0; j < 4; ++j) k = (k << 16) ^ (uint64_t)(rd() & 0xffffu);
        key[i] = k ^ (0x9e3779b97f4a7c15ull * (uint64_t)(i + 1));
    }
}

// Writes the oriented heads; only valid after orient_count_launch on the same
// input showed a symmetric, well-formed CSR (kept == nnz / 2).
void orient_write_launch(uint32_t n, uint64_t nnz, const uint32_t* heads, uint32_t* oheads, uint64_t* ws,
                         hipStream_t st) {
    OrientWs w = orient_ws(n, nnz, ws);
    if (w.tiles > 0) k_orient_emit<<<(unsigned)w.tiles, OR_THREADS, 0, st>>>(nnz, heads, w.bits, w.tile_off, oheads);
}

void make_keys(uint32_t n, uint64_t m, const uint64_t* ooff, const uint32_t* oheads, uint64_t* keys,
               hipStream_t st) {
    if (m == 0) return;
    k_make_keys<<<(unsigned)((m + 255) / 256), 256, 0, st>>>(n, m, ooff, oheads, keys);
}

void keys_to_heads(uint64_t m, const uint64_t* keys, uint32_t* oheads, uint32_t* err, hipStream_t st) {
    if (m == 0) return;
    k_keys_to_heads<<<(unsigned)((m + 255) / 256), 256, 0, st>>>(m, keys, oheads, err);
}

}  // namespace kt
