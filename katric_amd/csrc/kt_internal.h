// Internal declarations shared by the katric_amd translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

// Device-side validation flags (bitmask accumulated with atomicOr).
#define KT_ERR_OFFSETS 0x1u      // offsets not monotone / off[0] != 0 / off[n] != nnz
#define KT_ERR_HEAD_RANGE 0x2u   // head id >= n
#define KT_ERR_SELF_LOOP 0x4u    // head == tail
#define KT_ERR_DUPLICATE 0x8u    // repeated head in one list
#define KT_FLAG_UNSORTED 0x100u  // some list not ascending (not an error)
#define KT_ERR_GHOST 0x10u       // shard: remote head missing from the ghost table
#define KT_ERR_MESSAGE 0x20u     // shard: malformed message list
#define KT_ERR_SEGMENT 0x40u     // batch: a graph's per-vertex sum is not a multiple of 3
#define KT_ERR_ASYMMETRIC 0x80u  // some slot v -> u has no slot u -> v (pair checksum or kept count)
#define KT_ERR_MASK 0xffu

namespace kt {

// scans / sorts / orientation (kt_prep.hip)
uint64_t scan_tmp_words(uint64_t len);
void scan_u64(const uint64_t* in, uint64_t* out, uint64_t len, uint64_t* tmp, hipStream_t st);
uint64_t radix_tmp_words(uint64_t len);
void radix_sort_u64(uint64_t* keys, uint64_t* alt, uint64_t len, int bits, uint64_t* tmp, hipStream_t st);
void tile_bounds_launch(uint32_t n, uint64_t nnz, uint64_t tile, const uint64_t* off, uint32_t* bnd, hipStream_t st);
uint64_t orient_ws_words(uint32_t n, uint64_t nnz);
// Validation of a symmetric CSR beside the orientation count pass.  Every
// slot v -> u adds +h(min, max) (v < u) or -h(min, max) (v > u) of the
// unordered pair to a 64-bit sum, h being a keyed 64-bit mix; without
// duplicates in a row, a symmetric CSR sums to zero and an asymmetric one
// does not, except when the keyed sum collides (probability about 2^-64 for
// a key drawn at random).  With vbase (G + 1 entries, a disjoint union of G
// graphs), a head outside its row's graph is an error.
struct OrientCheck {
    uint64_t key[2];               // hash keys (key[0] is used)
    const uint64_t* vbase = nullptr;
    uint32_t G = 0;
};
void orient_count_launch(uint32_t n, uint64_t nnz, const uint64_t* off, const uint32_t* heads, uint32_t* deg,
                         uint64_t* ooff, uint64_t* ws, uint32_t* err, const OrientCheck& chk, hipStream_t st);
// Pair checksum, self-loop, range and (sorted-row) duplicate checks of a whole
// symmetric CSR without orienting it; ws holds sym_ws_words(n, nnz) words.
uint64_t sym_ws_words(uint32_t n, uint64_t nnz);
void sym_check_launch(uint32_t n, uint64_t nnz, const uint64_t* off, const uint32_t* heads, uint64_t* ws,
                      uint32_t* err, const OrientCheck& chk, hipStream_t st);
// A random key pair for OrientCheck.
void random_keys(uint64_t key[2]);
void orient_write_launch(uint32_t n, uint64_t nnz, const uint32_t* heads, uint32_t* oheads, uint64_t* ws,
                         hipStream_t st);
// The keep bitmask of the input slots (bit i: slot i is an out-edge) inside
// the orientation workspace; valid after orient_count_launch.
const uint64_t* orient_keep_bits(uint32_t n, uint64_t nnz, const uint64_t* ws);
void make_keys(uint32_t n, uint64_t m, const uint64_t* ooff, const uint32_t* oheads, uint64_t* keys,
               hipStream_t st);
void keys_to_heads(uint64_t m, const uint64_t* keys, uint32_t* oheads, uint32_t* err, hipStream_t st);
// Segmented in-place sort of the oriented lists oheads[ooff[v] .. ooff[v+1])
// of up to segsort_max_len() ids each, max_len being the longest (duplicates
// set KT_ERR_DUPLICATE); max_len_launch writes the longest list's length to
// *out, using tmp (segsort_tmp_words() words).
uint64_t segsort_tmp_words();
void max_len_launch(uint32_t n, const uint64_t* ooff, uint64_t* tmp, uint64_t* out, hipStream_t st);
uint32_t segsort_max_len();
void segsort_launch(uint32_t n, const uint64_t* ooff, uint32_t* oheads, uint32_t* err, uint64_t max_len, int num_cus,
                    hipStream_t st);

// fill kernels (kt_util.hip); used instead of hipMemsetAsync so that every
// pipeline stays capturable in a HIP graph.
void fill_u64(uint64_t* p, uint64_t count, uint64_t value, hipStream_t st);
void fill_u32(uint32_t* p, uint64_t count, uint32_t value, hipStream_t st);

// Intersection tuning knobs (bin thresholds, table sizes); see kt_tc.hip.
struct TcParams {
    uint32_t wave_max_deg;    // d+(v) <= this: one wave per tail vertex (LDS hash)
    uint32_t group_max_deg;   // d+(v) <= this: one workgroup (8 waves) per tail vertex (LDS hash)
    uint32_t team_chunk;      // neighbours per work grab inside a workgroup team (1..64)
    uint32_t grid_per_cu;     // persistent workgroups per CU
    uint32_t strategy;        // TC_STREAMED or TC_GROUPED* (hash-probe tails)
    uint32_t order;           // TC_BY_DEGREE or TC_BY_VERTEX (work-queue order)
    uint32_t table_shift = 3; // hash tables hold >= 2^table_shift slots per key where they fit (8: smaller
                              // tables to clear and flush per tail; 16 gave the same RMAT-22 time, batches 1 % slower)
    uint32_t wave_grab = 32;  // wave-bin tails a wave takes per queue access (1..64), vertex order
    uint32_t bsearch_min_deg = 8192;  // |S(x)| above this: binary-search probes in global memory (<= 8192)
    uint32_t pseudo2core = 0; // skip tails with |S(x)| < 2 and count them (Config::pseudo2core)
    uint32_t min_side = 1;    // plain counting: intersect every edge from the side with the shorter T list
};

// How a team walks the T lists of a chunk of neighbours: the whole wave on
// one list at a time, 256 ids per step (TC_STREAMED, long lists), or groups
// of G lanes on 64 / G lists at a time (TC_GROUPED<G>, short lists).
enum : uint32_t {
    TC_STREAMED = 0,
    TC_GROUPED16 = 1,
    TC_GROUPED8 = 2,
    TC_GROUPED32 = 3,
    TC_GROUPED4 = 4,
    TC_GROUPED2 = 5
};
// Order of the work queues: tails by degree bucket, largest first, in 13
// buckets (TC_BY_DEGREE), or by vertex id in two classes -- workgroup tails,
// then wave tails (TC_BY_VERTEX: tails of one graph of a batch, and their
// data, stay together in time and cache).
enum : uint32_t { TC_BY_DEGREE = 0, TC_BY_VERTEX = 1 };

// Generic "tail list" view for the intersection kernel:
//   tails x in [0, X): hash set S(x) = sheads[s_off[x] .. s_off[x+1]),
//   iterated neighbours R(x) = rheads[r_beg[x] .. r_end[x]) (a subrange that
//   may alias S), and for each u in R(x) its list T(u) =
//   theads[t_beg[u] .. t_end[u]).  Every hit w in S(x) cap T(u) emits the
//   triangle {pv_tail(x), u, w}.
struct TcView {
    uint32_t X;                 // number of tails
    const uint64_t* s_off;      // X+1 (used when s_beg == nullptr)
    const uint64_t* s_beg;      // X (nullable): S(x) = sheads[s_beg[x] .. s_end[x])
    const uint64_t* s_end;      // X
    const uint32_t* sheads;
    const uint64_t* r_beg;      // X (nullptr: R = S)
    const uint64_t* r_end;      // X
    const uint32_t* rheads;     // (nullptr: = sheads)
    const uint64_t* t_beg;      // per u (indexed by u)
    const uint64_t* t_end;      // per u (nullptr: t_end[u] = t_beg[u+1])
    const uint32_t* theads;
    const uint32_t* tail_id;    // X (nullptr: identity) -> per-vertex index of the tail
    const uint32_t* tail_graph; // X (nullptr: all graph 0) -> index into per-graph totals
    // work counters of a message phase: stat_off (X+1, nullable) gives the
    // length each tail's message adds to intersection_size per receiver (the
    // whole received message, not only its ghost part S); with
    // stat_skip_empty, receivers with an empty T(u) add nothing (they are
    // not in the contracted graph)
    const uint64_t* stat_off = nullptr;
    uint32_t stat_skip_empty = 0;
    // min-side walks (plain counting views; nullable): the symmetric input
    // CSR the orientation was built from and its keep bitmask (bit i: slot i
    // is an out-edge).  in_hash_max is set by the launch.
    const uint64_t* in_off = nullptr;
    const uint32_t* in_heads = nullptr;
    const uint64_t* in_keep = nullptr;
    uint64_t in_hash_max = 0;
};

struct TcOut {
    uint64_t* per_vertex;       // nullable
    uint64_t* totals;           // per graph (>= 1 entry)
    // triangle listing (plain views only): a counting launch with tail_count
    // set records every tail's number of emitted triangles; a listing launch
    // (tri set) writes tail x's triangles as (x, u, w) id triples from
    // position tail_off[x] on
    uint64_t* tail_count = nullptr;
    const uint64_t* tail_off = nullptr;
    uint32_t* tri = nullptr;
};

// Batch pipeline over a disjoint-union CSR (kt_batch.hip), shared with the
// serving loop.
struct BatchBufs {
    uint64_t* small;
    const uint64_t* off;
    const uint32_t* heads;
    uint64_t* vbase;
    uint32_t* deg;
    uint64_t* ooff;
    uint32_t* oheads;
    uint64_t* ws_orient;
    uint64_t* ws_tc;
    uint64_t* pv;
    uint64_t* totals;
    uint64_t* keys;
    uint64_t* alt;
    uint64_t* ws_radix;
};
// How a batch pipeline sorts the oriented lists: not at all (sorted input),
// in place per list (every list at most segsort_max_len() ids), or by the
// global radix sort (its scratch is in the layout).
enum BatchSort : uint32_t { SORT_NONE = 0, SORT_SEGMENTED = 1, SORT_RADIX = 2 };
uint64_t batch_layout(void* base, uint32_t G, uint64_t N, uint64_t nnz, bool host_in, bool own_pv, bool unsorted,
                      BatchBufs& B);
hipError_t batch_pipeline(hipStream_t st, int num_cus, const TcParams& prm, const BatchBufs& B, uint32_t G, uint64_t N,
                          uint64_t nnz, uint64_t* pv, uint64_t* h_totals, uint64_t* h_small, uint32_t sort,
                          const uint64_t sym_key[2]);
// After a pipeline found unsorted lists: the sort a rerun needs, from the
// longest oriented list of that run (B as it ran; synchronises st).
hipError_t batch_sort_mode(hipStream_t st, const BatchBufs& B, uint64_t N, uint64_t* h_small, uint32_t& mode);

uint64_t tc_ws_words(uint32_t X);
// Enqueues binning and the intersection kernels; returns the launch status.
hipError_t tc_launch(const TcView& view, const TcOut& out, const TcParams& prm, uint64_t* ws, int num_cus,
                     hipStream_t st, hipEvent_t ev_begin, hipEvent_t ev_end);

}  // namespace kt
