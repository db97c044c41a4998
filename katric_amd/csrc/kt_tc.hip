// Per-edge set intersection on the oriented CSR -- the hot loop of the local
// phase.  Replaces CetricEdgeIterator::run_local_* (counters/
// cetric_edge_iterator.h:229-388, 888-1066) and cetric::intersection /
// merge_intersection (counters/intersection.h:40-96) with a degree-binned
// hash-probe design for CDNA (wave64, LDS):
//
//   * a "team" -- one wave, or all eight waves of a 512-thread workgroup for
//     heavier tails -- takes a tail x and builds a linear-probing hash of
//     S(x) = N+(x) in LDS (load factor <= 1/16 where the table allows, <= 1/2
//     at the bins' largest sets): few buckets are full, so a probe rarely
//     has to continue into the next bucket;
//   * the team walks R(x) = N+(x) one neighbour u per wave at a time (u, and
//     the bounds of T(u) = N+(u), are wave-uniform scalar loads); the wave
//     streams T(u) with coalesced 16-byte loads, four ids per lane, the next
//     block prefetched while the current one is probed;
//   * every hit w is one emitted Triangle{x, u, w}; per-vertex counts are
//     privatised -- per-slot LDS counters for w, a wave sum for u, a team sum
//     for x -- and flushed with one global atomic per non-zero counter;
//   * tails are bucketed by ceil(log2 |S(x)|) and taken largest first from
//     per-bucket work queues (dynamic, LPT-like balance across the grid);
//   * tails above the workgroup bin go to 16-wave teams whose table fills a
//     workgroup's LDS (16384 slots, |S| <= 8192); only larger S(x) are probed
//     by binary search in the sorted S(x) in global memory.
//
// Counts are exact integers; the hash only changes where a set member is
// stored, never which members match.  T lists are read in aligned 16-byte
// blocks and may be touched up to 3 ids past their end: every heads array
// handed to this kernel carries 4 ids of padding.
#include <algorithm>

#include "kt_device.h"
#include "kt_internal.h"

namespace kt {

constexpr int TC_THREADS = 512;
constexpr int TC_WAVES = TC_THREADS / KT_WAVE;           // 8
constexpr uint32_t GROUP_SLOTS = 4096;                   // shared table of the 8-wave team
constexpr uint32_t WAVE_SLOTS = GROUP_SLOTS / TC_WAVES;  // private table of a 1-wave team (512)
constexpr uint32_t MIN_LOGS = 6;
// Work-queue buckets.  By degree: bucket b < 12 holds tails with |S| in
// (2^(b-1), 2^b].  By vertex: buckets 0..7 hold the workgroup tails and 8..15
// the wave tails of the eight contiguous eighths of the tail range, one per
// XCD, each in vertex order.  Bucket 17 holds the tails above the workgroup
// bin whose table still fits a workgroup's whole LDS (k_tc_big), bucket 16
// those beyond (binary search, k_tc_huge).
constexpr int NBUCKET = 18;
constexpr int BUCKET_HUGE = 16;
constexpr int BUCKET_BIG = 17;
// k_tc_big: one 1024-thread workgroup per CU with a 16384-slot table (128 KB
// of LDS for keys and counters), for |S| <= 8192
constexpr int BIG_THREADS = 1024;
constexpr int BIG_WAVES = BIG_THREADS / KT_WAVE;
constexpr uint32_t BIG_SLOTS = 16384;
constexpr int NXCD = 8;  // MI355X: workgroup i runs on XCD i % 8
constexpr uint32_t WAVE_GRAB = 8;  // wave-phase tails taken per queue access

// work-queue header (u32 words at the start of the workspace); every counter
// that waves update concurrently sits on its own 128-byte line so that the
// queues of different buckets land on different L2 channels:
//   [0, 18)               tails per bucket (written by binning, then read-only)
//   Q_PULL + 32 * b       pull cursor of bucket b
//   Q_FILL + 32 * b       fill cursor of bucket b (binning pass 2)
enum : int { Q_CNT = 0, Q_PULL = 32, Q_FILL = 32 * (NBUCKET + 1), Q_STRIDE = 32, Q_WORDS = 32 * (2 * NBUCKET + 2) };
// stats words after the queue header (the reference's work counters, and
// the T-list ids the probes stream)
enum : int { ST_WEDGE = 0, ST_ISZ = 1, ST_SKIP = 2, ST_STREAM = 3, ST_WORDS = 4 };

template <bool PLAIN>
struct View {
    TcView v;
    __device__ __forceinline__ void s_range(uint32_t x, uint64_t& b, uint64_t& e) const {
        if (PLAIN || v.s_beg == nullptr) {
            b = v.s_off[x];
            e = v.s_off[x + 1];
        } else {
            b = v.s_beg[x];
            e = v.s_end[x];
        }
    }
    // Min-side walk (plain views with in_off set): tail x walks its whole
    // input row N(x) instead of N+(x).  The entry of an out-neighbour u
    // streams T(u) here when |T(u)| <= |S(x)| (or when T(u)'s owner has no
    // LDS table), the entry of an in-neighbour y streams T(y) here when
    // |T(y)| < |S(x)| and S(x) is in an LDS table: every oriented edge is
    // intersected exactly once, from the side whose streamed list is shorter.
    __device__ __forceinline__ bool min_side() const { return PLAIN && v.in_off != nullptr; }
    __device__ __forceinline__ void r_range(uint32_t x, uint64_t& b, uint64_t& e) const {
        if (min_side()) {
            b = v.in_off[x];
            e = v.in_off[x + 1];
        } else if (PLAIN || v.r_beg == nullptr) {
            s_range(x, b, e);
        } else {
            b = v.r_beg[x];
            e = v.r_end[x];
        }
    }
    __device__ __forceinline__ uint32_t r_head(uint64_t i) const {
        return (PLAIN || v.rheads == nullptr) ? v.sheads[i] : v.rheads[i];
    }
    __device__ __forceinline__ void t_range(uint32_t u, uint64_t& b, uint64_t& e) const {
        if (PLAIN) {
            b = v.s_off[u];
            e = v.s_off[u + 1];
        } else {
            b = v.t_beg[u];
            e = v.t_end ? v.t_end[u] : v.t_beg[u + 1];
        }
    }
    __device__ __forceinline__ const uint32_t* t_heads() const { return PLAIN ? v.sheads : v.theads; }
    // Entry i of the walk, in two dependent steps so that the first can be
    // issued a chunk ahead: entry_head loads the neighbour u (and, for a
    // min-side walk, the keep-bitmask word of the slot); entry_bounds then
    // gives the bounds of T(u) (empty when T(u) is not streamed by this tail,
    // whose |S(x)| is ns) and what the entry adds to the reference's work
    // counters (st_len to wedge_checks; st_cnt: one receiver for
    // intersection_size).
    __device__ __forceinline__ void entry_head(uint64_t i, uint32_t& u, uint32_t& kw) const {
        if (min_side()) {
            u = v.in_heads[i];
            kw = reinterpret_cast<const uint32_t*>(v.in_keep)[i >> 5];
        } else {
            u = r_head(i);
            kw = 0;
        }
    }
    __device__ __forceinline__ void entry_bounds(uint64_t i, uint32_t u, uint32_t kw, uint64_t ns, uint64_t& b,
                                                 uint64_t& e, uint64_t& st_len, bool& st_cnt) const {
        t_range(u, b, e);
        if (min_side()) {
            const bool fwd = (kw >> (i & 31)) & 1u;
            const uint64_t dr = e - b;
            st_len = fwd ? dr : 0;
            st_cnt = fwd;
            const bool here = fwd ? (dr <= ns || dr > v.in_hash_max) : (dr < ns && ns <= v.in_hash_max);
            if (!here) e = b;
            return;
        }
        st_len = e - b;
        st_cnt = !(stat_skip_empty() && e == b);
    }
    __device__ __forceinline__ void walk_entry(uint64_t i, uint64_t ns, uint32_t& u, uint64_t& b, uint64_t& e,
                                               uint64_t& st_len, bool& st_cnt) const {
        uint32_t kw;
        entry_head(i, u, kw);
        entry_bounds(i, u, kw, ns, b, e, st_len, st_cnt);
    }
    __device__ __forceinline__ uint32_t tail_pv(uint32_t x) const {
        return (PLAIN || v.tail_id == nullptr) ? x : v.tail_id[x];
    }
    __device__ __forceinline__ uint32_t tail_graph(uint32_t x) const {
        return v.tail_graph ? v.tail_graph[x] : 0u;
    }
    // length tail x charges to intersection_size per receiver (ns = |S(x)|)
    __device__ __forceinline__ uint64_t stat_len(uint32_t x, uint64_t ns) const {
        return (!PLAIN && v.stat_off) ? v.stat_off[x + 1] - v.stat_off[x] : ns;
    }
    __device__ __forceinline__ bool stat_skip_empty() const { return !PLAIN && v.stat_skip_empty; }
};

__device__ __forceinline__ uint32_t ceil_log2(uint64_t x) { return x <= 1 ? 0 : 64 - __clzll(x - 1); }

__device__ __forceinline__ uint32_t uniform(uint32_t x) { return __builtin_amdgcn_readfirstlane(x); }
__device__ __forceinline__ uint64_t uniform64(uint64_t x) {
    uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)x);
    uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(x >> 32));
    return ((uint64_t)hi << 32) | lo;
}

// ---------------------------------------------------------------------------
// binning: bucket tails by ceil(log2 |S(x)|), two passes (count, fill)
// ---------------------------------------------------------------------------
template <bool PLAIN>
__device__ __forceinline__ int tail_bucket(const View<PLAIN>& V, const TcParams& prm, uint32_t x, uint64_t& s,
                                           uint64_t& r) {
    uint64_t sb, se, rb, re;
    V.s_range(x, sb, se);
    V.r_range(x, rb, re);
    s = se - sb;
    r = re - rb;
    if (s == 0 || r == 0 || ((PLAIN || prm.pseudo2core) && s < 2)) return -1;
    if (s > prm.bsearch_min_deg) return BUCKET_HUGE;
    if (s > prm.group_max_deg) return BUCKET_BIG;
    int b = (int)ceil_log2(s);  // 0..11
    // a min-side walk covers the whole input row: a long row goes to a
    // bigger bucket (a workgroup team beyond 512 entries), whatever |S(x)|
    if (V.min_side()) b = max(b, min((int)ceil_log2(r) - 2, (int)ceil_log2(prm.group_max_deg)));
    if (prm.order == TC_BY_VERTEX) {  // two classes per XCD, each in vertex order
        const int xcd = (int)(((uint64_t)x * NXCD) / V.v.X);
        return b <= (int)ceil_log2(prm.wave_max_deg) ? NXCD + xcd : xcd;
    }
    return b;
}

// Grid-stride over tails; per-block bucket counts in LDS, one global atomic
// per bucket (and per stat) per block.
template <bool PLAIN>
__global__ void __launch_bounds__(256) k_tc_bin_count(View<PLAIN> V, TcParams prm, uint32_t* __restrict__ q,
                                                      uint64_t* __restrict__ stats) {
    __shared__ uint32_t cnt[NBUCKET];
    __shared__ uint64_t red[3][4];
    if (threadIdx.x < NBUCKET) cnt[threadIdx.x] = 0;
    __syncthreads();
    uint64_t wedge = 0, isz = 0, skipped = 0;
    for (uint32_t x = blockIdx.x * blockDim.x + threadIdx.x; x < V.v.X; x += gridDim.x * blockDim.x) {
        uint64_t s, r;
        int b = tail_bucket(V, prm, x, s, r);
        if (b >= 0) atomicAdd(&cnt[b], 1u);
        // pseudo2core: tails with fewer than two out-neighbours are skipped
        // and counted (cetric_edge_iterator.h:239-242)
        if (prm.pseudo2core && s < 2) ++skipped;
        // a plain tail with |N+(x)| == 1 emits nothing but still counts one
        // intersection in the reference's counters (cetric_edge_iterator.h:279-281)
        if (PLAIN && s == 1 && !prm.pseudo2core) {
            uint64_t sb, se, tb, te;
            V.s_range(x, sb, se);
            V.t_range(V.v.sheads[sb], tb, te);
            wedge += te - tb;
            isz += 1 + (te - tb);
        }
        // a message with receivers but no ghost entry finds no triangle and
        // is not binned; its receivers still count (cetric_edge_iterator.h:1850-1856)
        if (!PLAIN && s == 0 && r != 0) {
            uint64_t rb, re;
            V.r_range(x, rb, re);
            const uint64_t len = V.stat_len(x, 0);
            for (uint64_t i = rb; i < re; ++i) {
                uint64_t tb, te, sl;
                uint32_t u;
                bool sc;
                V.walk_entry(i, 0, u, tb, te, sl, sc);
                if (!sc) continue;
                wedge += sl;
                isz += len + sl;
            }
        }
    }
    wedge = wave_sum(wedge);
    isz = wave_sum(isz);
    skipped = wave_sum(skipped);
    if (lane_id() == 0) {
        red[0][threadIdx.x / KT_WAVE] = wedge;
        red[1][threadIdx.x / KT_WAVE] = isz;
        red[2][threadIdx.x / KT_WAVE] = skipped;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint64_t a = red[0][0] + red[0][1] + red[0][2] + red[0][3];
        uint64_t c = red[1][0] + red[1][1] + red[1][2] + red[1][3];
        uint64_t d = red[2][0] + red[2][1] + red[2][2] + red[2][3];
        if (a) atomic_add_u64(&stats[ST_WEDGE], a);
        if (c) atomic_add_u64(&stats[ST_ISZ], c);
        if (d) atomic_add_u64(&stats[ST_SKIP], d);
    }
    if (threadIdx.x < NBUCKET && cnt[threadIdx.x]) atomicAdd(&q[Q_CNT + threadIdx.x], cnt[threadIdx.x]);
}

// Bucket regions in one list: bucket b starts at sum of counts of b' < b.
__device__ __forceinline__ uint32_t bucket_base(const uint32_t* q, int b) {
    uint32_t s = 0;
    for (int i = 0; i < b; ++i) s += q[Q_CNT + i];
    return s;
}

// Each block owns a contiguous range of tails: it counts its tails per
// bucket, reserves one range per bucket with a single global atomic, then
// writes its tails into the reserved ranges.
template <bool PLAIN>
__global__ void __launch_bounds__(256) k_tc_bin_fill(View<PLAIN> V, TcParams prm, uint32_t* __restrict__ q,
                                                     uint32_t* __restrict__ list) {
    __shared__ uint32_t cnt[NBUCKET];
    __shared__ uint32_t pos[NBUCKET];
    const uint32_t X = V.v.X;
    const uint32_t per = (X + gridDim.x - 1) / gridDim.x;
    const uint32_t x0 = min(X, blockIdx.x * per), x1 = min(X, x0 + per);
    if (threadIdx.x < NBUCKET) cnt[threadIdx.x] = 0;
    __syncthreads();
    for (uint32_t x = x0 + threadIdx.x; x < x1; x += blockDim.x) {
        uint64_t s, r;
        int b = tail_bucket(V, prm, x, s, r);
        if (b >= 0) atomicAdd(&cnt[b], 1u);
    }
    __syncthreads();
    if (threadIdx.x < NBUCKET) {
        const uint32_t c = cnt[threadIdx.x];
        pos[threadIdx.x] = c ? bucket_base(q, threadIdx.x) + atomicAdd(&q[Q_FILL + Q_STRIDE * threadIdx.x], c) : 0;
    }
    __syncthreads();
    for (uint32_t x = x0 + threadIdx.x; x < x1; x += blockDim.x) {
        uint64_t s, r;
        int b = tail_bucket(V, prm, x, s, r);
        if (b >= 0) list[atomicAdd(&pos[b], 1u)] = x;
    }
}

// ---------------------------------------------------------------------------
// the team worker
// ---------------------------------------------------------------------------
// Per-wave scratch of the grouped scan: the lanes of the chunk's non-empty
// lists in order, and the hit counters of the chunk's (up to 64) neighbours.
struct PackLds {
    uint32_t marks[KT_WAVE / 4];
    uint32_t ucnt[KT_WAVE];
};

struct Lds {
    uint32_t keys[GROUP_SLOTS];
    uint32_t cnt[GROUP_SLOTS];
    uint32_t ovf[GROUP_SLOTS / 128];  // a bit per bucket (the team's table, or 4 words per wave table)
    PackLds pk[TC_WAVES];
    uint64_t red[TC_WAVES];
    unsigned long long list_ctr[TC_WAVES];  // listing: triangles of the current tail written so far
    uint32_t next_u;
    uint32_t bcast;
    uint32_t full;  // the team table has a full bucket
};

// Team control words of k_tc_big (its table is dynamic LDS).
struct LdsBig {
    uint32_t ovf[BIG_SLOTS / 128];  // a bit per bucket of the dynamic-LDS table
    PackLds pk[BIG_WAVES];
    uint64_t red[BIG_WAVES];
    unsigned long long list_ctr[BIG_WAVES];
    uint32_t next_u;
    uint32_t bcast;
    uint32_t full;
};

template <int NW>
__device__ __forceinline__ void team_sync() {
    if (NW > 1)
        __syncthreads();
    else
        wave_sync();
}

// Bucketed hash set in LDS: 4-slot buckets (16 bytes, read with one
// ds_read_b128), filled left to right; a key lives in its home bucket
// hash(w) or, when that was full at insertion, in a later bucket (linear
// probing over buckets).  A lookup therefore ends at the first bucket that
// holds w or has an empty last slot.
// Home bucket of id w as a byte offset into the table: w xor w rotated right
// by 12, masked with boff_mask = (buckets - 1) << 4 -- two full-rate
// instructions (one v_bitop3 after the rotation) that give the LDS address
// offset directly (a multiplicative hash costs a quarter-rate multiply).  The
// rotation folds the high bits onto the low ones, so ids that share their low
// bits still spread; on vertex ids it fills buckets like a multiplicative hash.
__device__ __forceinline__ uint32_t bucket_off(uint32_t w, uint32_t boff_mask) {
    return (w ^ __builtin_amdgcn_alignbit(w, w, 12)) & boff_mask;
}

__device__ __forceinline__ uint32_t* at_byte(uint32_t* base, uint32_t off) {
    return reinterpret_cast<uint32_t*>(reinterpret_cast<char*>(base) + off);
}
__device__ __forceinline__ const uint4* bucket_at(const uint32_t* keys, uint32_t off) {
    return reinterpret_cast<const uint4*>(reinterpret_cast<const char*>(keys) + off);
}

// Inserts w; returns true when w went past its home bucket (which then gets
// its bit in `ovf`, one bit per bucket).  A lookup continues past its home
// bucket only when that bucket is full and has the bit; in a table where no
// insertion went past its home bucket, every lookup ends there.
__device__ __forceinline__ bool table_insert(uint32_t* keys, uint32_t* ovf, uint32_t boff_mask, uint32_t w) {
    const uint32_t home = bucket_off(w, boff_mask);
    uint32_t b = home;
    bool moved = false;
    while (true) {
        uint32_t* bk = at_byte(keys, b);
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            uint32_t old = atomicCAS(&bk[s], KT_EMPTY, w);
            if (old == KT_EMPTY || old == w) return moved;
        }
        if (!moved) atomicOr(&ovf[home >> 9], 1u << ((home >> 4) & 31));
        b = (b + 16) & boff_mask;
        moved = true;
    }
}

// Probes the first N (2 or 4) ids of this lane (id e valid iff valid[e]) and
// returns this lane's number of hits.  Each hit bumps the w-counter of its slot and calls
// on_hit(e).  The four slot compares of a probe stay lane masks; hits are
// counted per lane (one add), so the scalar unit only sees the branch masks.
// Lookups continue past a full home bucket only in a table where some key
// went past its home bucket (moved, wave-uniform), and only from a bucket
// with its `ovf` bit.
template <int N, class OnHit>
__device__ __forceinline__ uint32_t probe_ids(const uint32_t* keys, uint32_t* cnt, const uint32_t* ovf,
                                           uint32_t boff_mask, bool moved, const uint32_t (&w)[4],
                                           const bool (&valid)[4], OnHit on_hit) {
    uint32_t bo[4];
    uint4 kk[4];
#pragma unroll
    for (int e = 0; e < N; ++e) {
        bo[e] = bucket_off(w[e], boff_mask);
        kk[e] = *bucket_at(keys, bo[e]);
    }
    uint32_t nh = 0;
    bool hit[4];
#pragma unroll
    for (int e = 0; e < N; ++e) {
        const bool m0 = kk[e].x == w[e], m1 = kk[e].y == w[e], m2 = kk[e].z == w[e], m3 = kk[e].w == w[e];
        hit[e] = valid[e] && (m0 || m1 || m2 || m3);
        nh += hit[e] ? 1u : 0u;
        if (hit[e]) {
            atomicAdd(at_byte(cnt, bo[e]) + (m0 ? 0u : m1 ? 1u : m2 ? 2u : 3u), 1u);
            on_hit(e);
        }
    }
    if (!moved) return nh;
    // a full home bucket without w: w may sit in a later bucket
    bool need[4];
#pragma unroll
    for (int e = 0; e < N; ++e) need[e] = valid[e] && !hit[e] && kk[e].w != KT_EMPTY;
    // one branch for the N ids, skipped as a whole when no lane needs it
    bool any = false;
#pragma unroll
    for (int e = 0; e < N; ++e) any = any || need[e];
    if (any) {
#pragma unroll
        for (int e = 0; e < N; ++e) {
            if (!need[e]) continue;
            // nothing went past w's home bucket: w is not in the table
            if (!((ovf[bo[e] >> 9] >> ((bo[e] >> 4) & 31)) & 1u)) continue;
            uint32_t b = bo[e];
            while (true) {
                b = (b + 16) & boff_mask;
                const uint4 k = *bucket_at(keys, b);
                const bool n0 = k.x == w[e], n1 = k.y == w[e], n2 = k.z == w[e], n3 = k.w == w[e];
                if (n0 || n1 || n2 || n3) {
                    atomicAdd(at_byte(cnt, b) + (n0 ? 0u : n1 ? 1u : n2 ? 2u : 3u), 1u);
                    on_hit(e);
                    ++nh;
                    break;
                }
                if (k.w == KT_EMPTY) break;
            }
        }
    }
    return nh;
}

struct NoHitAction {
    __device__ __forceinline__ void operator()(int) const {}
};

// 16 bytes at 4-byte alignment (T lists start anywhere); one global_load_dwordx4.
struct __attribute__((packed, aligned(4))) Ids4 {
    uint32_t x, y, z, w;
};
__device__ __forceinline__ Ids4 load_ids4(const uint32_t* p) { return *reinterpret_cast<const Ids4*>(p); }

// A T list of n ids is walked 64 << sh ids per step, 1 << sh per lane: two
// ids per lane up to 128 ids (one step, half the probes of four per lane),
// four beyond.  One id per lane for lists up to 64 was slower: a step's four
// (or two) independent bucket reads per lane are what hide its latency.
__device__ __forceinline__ uint32_t lane_shift(uint32_t n) { return n <= 2 * KT_WAVE ? 1u : 2u; }

// Binary-search probe of the sorted S(x) in global memory (huge tails).
__device__ __forceinline__ uint32_t probe_bsearch(const uint32_t* __restrict__ s, uint64_t ns, uint64_t* pv,
                                                  uint32_t w) {
    uint64_t lo = 0, hi = ns;
    while (lo < hi) {
        uint64_t mid = (lo + hi) >> 1;
        if (s[mid] < w)
            lo = mid + 1;
        else
            hi = mid;
    }
    if (lo < ns && s[lo] == w) {
        if (pv) atomic_add_u64(&pv[w], 1);
        return 1;
    }
    return 0;
}

__device__ __forceinline__ uint32_t readlane32(uint32_t x, int j) { return __builtin_amdgcn_readlane(x, j); }
__device__ __forceinline__ uint64_t readlane64(uint64_t x, int j) {
    uint32_t lo = __builtin_amdgcn_readlane((uint32_t)x, j);
    uint32_t hi = __builtin_amdgcn_readlane((uint32_t)(x >> 32), j);
    return ((uint64_t)hi << 32) | lo;
}

// Triangle listing: the calling lanes (exactly those that hit) write their
// triangles (x, u, w) of the current tail x from the tail's first output
// position on, after a team-wide running count kept in LDS.
// Writes stay inside the tail's own range [base, end) of the output even if
// this pass found more triangles than the counting pass recorded.
struct Emitter {
    uint32_t* tri;
    uint64_t base, end;
    unsigned long long* ctr;
    uint32_t x;
    __device__ __forceinline__ void operator()(uint32_t u, uint32_t w) const {
        const uint64_t m = __builtin_amdgcn_ballot_w64(true);
        const uint32_t leader = (uint32_t)__ffsll((unsigned long long)m) - 1;
        unsigned long long first = 0;
        if (lane_id() == leader) first = atomicAdd(ctr, (unsigned long long)__popcll(m));
        first = __shfl(first, (int)leader, 64);
        const uint64_t pos = base + first + (uint64_t)__popcll(m & lanemask_lt());
        if (pos < end) {
            uint32_t* t = tri + 3 * pos;
            t[0] = x;
            t[1] = u;
            t[2] = w;
        }
    }
};

// Per-wave work counters: the reference's wedge_checks and
// intersection_size, and the T-list ids this design streams through probes.
struct WorkStats {
    uint64_t wedge = 0, isz = 0, stream = 0;
    uint64_t stream_lane = 0;  // per lane, summed into stream once per wave at the end (grouped walks)
    // the wave's total of stream and stream_lane (every lane calls this)
    __device__ __forceinline__ uint64_t streamed() const { return stream + uniform64(wave_sum(stream_lane)); }
};

struct ProbeCtx {
    Emitter em;  // listing launches only
    uint32_t* keys;
    uint32_t* cnt;
    uint32_t* ovf;   // one bit per bucket: some key went past it
    uint32_t boff;   // (buckets - 1) << 4: bucket byte-offset mask
    bool full;       // some key went past its home bucket (wave-uniform)
    const uint32_t* s;  // sorted S(x) (binary-search probes)
    uint64_t ns;
    uint64_t* pv;
};

// Streams the T lists of up to 64 neighbours R[k0, k0+cnt) through the probe.
// Lane j first loads u_j and the bounds of T(u_j) (one coalesced load plus two
// gathers for the whole chunk); the wave then walks the non-empty lists in
// order, four ids per lane per step (two for lists of at most 128 ids; one
// 16-byte load per lane either way, which needs only 4-byte alignment: a load
// whose width followed the list would have to complete before the branches
// join), prefetching the next block -- of the same list or the next one --
// before probing the current block.  Positions
// and list bounds are wave-uniform 32-bit values; hits are counted per lane
// and summed across the wave once per list.  Returns the chunk's hits.
template <bool HASH, bool PLAIN, bool LIST>
__device__ __forceinline__ uint64_t scan_chunk(const View<PLAIN>& V, uint64_t k0, uint32_t cnt, uint32_t uj,
                                               uint32_t kw, const ProbeCtx& P, uint64_t nstat, WorkStats& wk) {
    const uint32_t lane = lane_id();
    uint64_t tbj = 0, tej = 0;
    uint64_t stl = 0;
    bool stc = false;
    if (lane < cnt) {
        V.entry_bounds(k0 + lane, uj, kw, P.ns, tbj, tej, stl, stc);
    }
    uint64_t pending = __ballot(lane < cnt && tej > tbj);
    {
        const uint64_t cw = uniform64(wave_sum(stl));
        const uint32_t nrecv = (uint32_t)__popcll(__ballot(stc));
        wk.wedge += cw;
        wk.isz += cw + (uint64_t)nrecv * nstat;
    }
    if (pending == 0) return 0;
    const uint32_t* __restrict__ T = V.t_heads();

    int j = __ffsll((unsigned long long)pending) - 1;
    pending &= pending - 1;
    uint32_t u = readlane32(uj, j);
    const uint32_t* base = T + readlane64(tbj, j);  // list start (uniform)
    // |T(u)| < 2^32: an out-list holds distinct 32-bit ids
    uint32_t len = (uint32_t)(readlane64(tej, j) - readlane64(tbj, j));
    wk.stream += len;
    uint32_t pos = 0;               // block start within the list (uniform)
    uint32_t sh = lane_shift(len);  // 1 << sh ids per lane (uniform)
    // Every lane loads (lanes past the list end re-read its start; their ids
    // are masked by `valid`): an unconditional load keeps the number of
    // loads in flight static, so the prefetch below is not waited for early.
    // Addresses are a wave-uniform base plus a 32-bit lane offset; at most
    // three ids past a list's end are read.
    Ids4 q = load_ids4(base + ((lane << sh) < len ? (lane << sh) : 0u));
    uint32_t nh = 0;    // this lane's hits on the current list
    uint32_t uacc = 0;  // hits on this lane's own list (T(uj)), added at the list's end
    uint64_t total = 0;
    while (true) {
        // next block: same list, or the first block of the next non-empty list
        const uint32_t* nbase = base;
        uint32_t npos = pos + (KT_WAVE << sh), nlen = len, nu = u, nsh = sh;
        int nj = j;
        bool more = true, switch_u = false;
        if (npos >= len) {
            switch_u = true;
            if (pending) {
                nj = __ffsll((unsigned long long)pending) - 1;
                pending &= pending - 1;
                nu = readlane32(uj, nj);
                const uint64_t ntb = readlane64(tbj, nj);
                nbase = T + ntb;
                nlen = (uint32_t)(readlane64(tej, nj) - ntb);
                wk.stream += nlen;
                nsh = lane_shift(nlen);
                npos = 0;
            } else {
                more = false;
            }
        }
        const uint32_t noff = npos + (lane << nsh);
        const Ids4 qn = load_ids4(nbase + (noff < nlen ? noff : 0u));
        // probe the current block: id e < (1 << sh) of this lane is list
        // entry pos + (lane << sh) + e
        const int lim = (int)(len - pos) - (int)(lane << sh);
        const bool valid[4] = {0 < lim, 1 < lim, 2 < lim, 3 < lim};  // ids 2, 3 only when sh == 2
        const uint32_t w[4] = {q.x, q.y, q.z, q.w};
        if (HASH) {
            if constexpr (LIST) {
                auto hit = [&](int e) { P.em(u, w[e]); };
                if (sh == 2)
                    nh += probe_ids<4>(P.keys, P.cnt, P.ovf, P.boff, P.full, w, valid, hit);
                else
                    nh += probe_ids<2>(P.keys, P.cnt, P.ovf, P.boff, P.full, w, valid, hit);
            } else {
                if (sh == 2)
                    nh += probe_ids<4>(P.keys, P.cnt, P.ovf, P.boff, P.full, w, valid, NoHitAction{});
                else
                    nh += probe_ids<2>(P.keys, P.cnt, P.ovf, P.boff, P.full, w, valid, NoHitAction{});
            }
        } else {
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                if (e >= (1 << sh) || !valid[e]) continue;
                const uint32_t h = probe_bsearch(P.s, P.ns, P.pv, w[e]);
                nh += h;
                if (LIST && h) P.em(u, w[e]);
            }
        }
        if (switch_u) {
            const uint32_t h = wave_total_dpp(nh);
            total += h;
            uacc += (lane == (uint32_t)j) ? h : 0u;
            nh = 0;
        }
        if (!more) break;
        u = nu;
        j = nj;
        base = nbase;
        len = nlen;
        pos = npos;
        sh = nsh;
        q = qn;
    }
    // one atomic per neighbour with hits, issued once for the chunk
    if (P.pv && uacc) atomic_add_u64(&P.pv[uj], uacc);
    return total;
}

// Grouped variant of scan_chunk for the hash probe: the wave splits into
// 64 / G groups of G lanes, and group q streams the chunk's non-empty T lists
// of rank q, q + 64/G, q + 2*64/G, ... -- four ids per lane per step, G * 4
// contiguous ids per group (one coalesced run per group), the next block of
// the same list or the first block of the group's next list prefetched.  A
// step therefore covers 64/G lists at once: short lists no longer leave most
// of the wave idle.  Group state (list start, length, position) lives in
// vector registers, uniform within the group; the next list's start and
// length are read from the lane that loaded them (ds_bpermute, every lane
// active).  Hits are added to the neighbour's LDS counter as they occur and
// flushed with one global atomic per neighbour.
template <int G, bool PLAIN>
__device__ __forceinline__ uint64_t scan_chunk_grouped(const View<PLAIN>& V, uint64_t k0, uint32_t cnt, uint32_t uj,
                                                       uint32_t kw, const ProbeCtx& P, uint64_t nstat,
                                                       WorkStats& wk, PackLds& pk) {
    constexpr uint32_t NQ = KT_WAVE / G;
    const uint32_t lane = lane_id();
    uint64_t tbj = 0, tej = 0, stl = 0;
    bool stc = false;
    if (lane < cnt) {
        V.entry_bounds(k0 + lane, uj, kw, P.ns, tbj, tej, stl, stc);
    }
    const bool nonempty = lane < cnt && tej > tbj;
    const uint64_t nem = __builtin_amdgcn_ballot_w64(nonempty);
    const uint32_t nne = (uint32_t)__popcll(nem);
    {
        const uint64_t cw = uniform64(wave_sum(stl));
        wk.wedge += cw;
        wk.isz += cw + (uint64_t)__popcll(__builtin_amdgcn_ballot_w64(stc)) * nstat;
        wk.stream_lane += nonempty ? tej - tbj : 0ull;
    }
    if (nne == 0) return 0;
    // order[r] = lane of the r-th non-empty list
    uint8_t* order = reinterpret_cast<uint8_t*>(pk.marks);
    if (nonempty) order[__builtin_amdgcn_mbcnt_hi((uint32_t)(nem >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)nem, 0))] =
        (uint8_t)lane;
    pk.ucnt[lane] = 0;
    wave_sync();
    const uint32_t tb_lo = (uint32_t)tbj, tb_hi = (uint32_t)(tbj >> 32);
    const uint32_t len32 = (uint32_t)(tej - tbj);  // distinct 32-bit ids: < 2^32
    const uint32_t* __restrict__ T = V.t_heads();
    const uint32_t q = lane / G, s4 = 4 * (lane % G);
    // state of list rank r: source lane, start, length (every lane calls this)
    auto fetch = [&](uint32_t r, uint32_t& src, const uint32_t*& base, uint32_t& len) {
        const bool ok = r < nne;
        src = order[ok ? r : 0];
        const int a = (int)(src * 4);
        const uint32_t lo = (uint32_t)__builtin_amdgcn_ds_bpermute(a, (int)tb_lo);
        const uint32_t hi = (uint32_t)__builtin_amdgcn_ds_bpermute(a, (int)tb_hi);
        const uint32_t l = (uint32_t)__builtin_amdgcn_ds_bpermute(a, (int)len32);
        base = T + (((uint64_t)hi << 32) | lo);
        len = ok ? l : 0u;
    };
    uint32_t k = q;  // rank of the group's current list
    uint32_t csrc, nsrc, clen, nlen;
    const uint32_t *cbase, *nbase;
    fetch(k, csrc, cbase, clen);
    fetch(k + NQ, nsrc, nbase, nlen);
    uint32_t pos = 0;
    // every lane loads (lanes past their list's end re-read its start, masked
    // by `valid`): a static number of loads in flight keeps the prefetch from
    // being waited for early
    Ids4 cur = load_ids4(s4 < clen ? cbase + s4 : cbase);
    uint32_t nh = 0;
    while (__builtin_amdgcn_ballot_w64(clen != 0)) {
        // the next block of this group: same list, or its next list
        const bool same = pos + 4 * G < clen;
        const uint32_t npos = same ? pos + 4 * G : 0u;
        const uint32_t* bb = same ? cbase : nbase;
        const uint32_t bl = same ? clen : nlen;
        const Ids4 nxt = load_ids4(npos + s4 < bl ? bb + npos + s4 : (bl ? bb : cbase));
        const int lim = (int)clen - (int)(pos + s4);
        const bool valid[4] = {0 < lim, 1 < lim, 2 < lim, 3 < lim};
        const uint32_t w[4] = {cur.x, cur.y, cur.z, cur.w};
        nh += probe_ids<4>(P.keys, P.cnt, P.ovf, P.boff, P.full, w, valid,
                           [&](int) { atomicAdd(&pk.ucnt[csrc], 1u); });
        // advance; a group that switched lists fetches its list after next
        uint32_t fsrc, flen;
        const uint32_t* fbase;
        fetch(same ? k + NQ : k + 2 * NQ, fsrc, fbase, flen);
        if (!same) {
            k += NQ;
            csrc = nsrc;
            cbase = nbase;
            clen = nlen;
            nsrc = fsrc;
            nbase = fbase;
            nlen = flen;
        }
        pos = npos;
        cur = nxt;
    }
    wave_sync();
    const uint32_t h = lane < cnt ? pk.ucnt[lane] : 0u;
    if (h && P.pv) atomic_add_u64(&P.pv[uj], h);
    return wave_total_dpp(nh);
}

// Bounds of a tail's S and R ranges and its counter length.
struct TailBounds {
    uint64_t sb, se, rb, re, nstat;
};

template <bool PLAIN>
__device__ __forceinline__ TailBounds tail_bounds(const View<PLAIN>& V, uint32_t x) {
    TailBounds t;
    V.s_range(x, t.sb, t.se);
    V.r_range(x, t.rb, t.re);
    t.nstat = V.stat_len(x, t.se - t.sb);
    return t;
}

// Process tail x, whose bounds are tb, with a team of NW waves (NW == 1:
// this wave alone).
template <int NW, bool HASH, bool PLAIN, int GL, bool LIST, class LdsT>
__device__ __forceinline__ void team_process(const View<PLAIN>& V, uint32_t x, const TailBounds& tb, LdsT& L,
                                             uint32_t* keys, uint32_t* cnt, uint32_t* ovf, uint32_t cap_slots,
                                             uint32_t shift, uint32_t chunk, const TcOut& out,
                                             uint64_t& my_total, WorkStats& wk) {
    const unsigned lane = lane_id();
    const unsigned wave = threadIdx.x / KT_WAVE;
    const unsigned twave = NW > 1 ? wave : 0;
    const unsigned tid = twave * KT_WAVE + lane;

    const uint64_t sb = uniform64(tb.sb), se = uniform64(tb.se);
    const uint64_t rb = uniform64(tb.rb), re = uniform64(tb.re);
    const uint64_t ns = se - sb, nr = re - rb;
    const uint64_t nstat = uniform64(tb.nstat);
    // The first chunk's neighbours are loaded before the tail's table is
    // built (clearing it and inserting S(x) hide the round trip); every later
    // chunk's neighbours are loaded while the chunk before it is scanned (the
    // bounds of their T lists depend on them).
    uint64_t k = (uint64_t)twave * chunk;
    uint32_t u_nx = 0, kw_nx = 0;
    if (k < nr && lane < min<uint64_t>(chunk, nr - k)) V.entry_head(rb + k + lane, u_nx, kw_nx);

    ProbeCtx P;
    P.keys = keys;
    P.cnt = cnt;
    P.ovf = ovf;
    P.boff = 0;
    P.full = false;
    P.s = V.v.sheads + sb;
    P.ns = ns;
    P.pv = out.per_vertex;
    uint32_t slots = 0;
    if (HASH) {
        uint32_t logS = ceil_log2(ns << shift);  // load <= 2^-shift unless capped
        if (logS < MIN_LOGS) logS = MIN_LOGS;
        slots = 1u << logS;
        if (slots > cap_slots) {  // bins guarantee cap_slots >= 2 * ns
            slots = cap_slots;
            logS = ceil_log2(cap_slots);
        }
        P.boff = ((slots >> 2) - 1) << 4;
        for (uint32_t i = tid; i < slots; i += NW * KT_WAVE) {
            keys[i] = KT_EMPTY;
            cnt[i] = 0;
        }
        for (uint32_t i = tid; i < (slots + 127) / 128; i += NW * KT_WAVE) ovf[i] = 0;  // one bit per 4-slot bucket
        team_sync<NW>();
        bool full = false;
        for (uint64_t i = tid; i < ns; i += NW * KT_WAVE)
            full |= table_insert(keys, ovf, P.boff, V.v.sheads[sb + i]);
        if (NW > 1) {
            if (tid == 0) L.full = 0;
            team_sync<NW>();
            if (__builtin_amdgcn_ballot_w64(full) && lane == 0) atomicOr(&L.full, 1u);
            team_sync<NW>();
            P.full = uniform(L.full) != 0;
        } else {
            P.full = __builtin_amdgcn_ballot_w64(full) != 0;
        }
    }
    if (NW > 1 && tid == 0) L.next_u = NW * chunk;
    if (LIST) {
        P.em = Emitter{out.tri, out.tail_off[x], out.tail_off[x + 1], &L.list_ctr[NW > 1 ? 0 : wave], x};
        if (tid == 0) *P.em.ctr = 0;
    }
    team_sync<NW>();

    uint64_t tail_hits = 0;
    while (k < nr) {
        const uint32_t c = (uint32_t)min<uint64_t>(chunk, nr - k);
        const uint32_t u_cur = u_nx, kw_cur = kw_nx;
        uint64_t nk;
        if (NW > 1) {
            uint32_t g = 0;
            if (lane == 0) g = atomicAdd(&L.next_u, chunk);
            nk = uniform(g);
        } else {
            nk = k + chunk;
        }
        if (nk < nr && lane < min<uint64_t>(chunk, nr - nk)) V.entry_head(rb + nk + lane, u_nx, kw_nx);
        if (GL < KT_WAVE && HASH)
            tail_hits += scan_chunk_grouped<GL, PLAIN>(V, rb + k, c, u_cur, kw_cur, P, nstat, wk, L.pk[wave]);
        else
            tail_hits += scan_chunk<HASH, PLAIN, LIST>(V, rb + k, c, u_cur, kw_cur, P, nstat, wk);
        k = nk;
    }

    // team reduction of the tail's hits
    uint64_t hsum = tail_hits;  // wave-uniform
    if (NW > 1) {
        if (lane == 0) L.red[wave] = hsum;
        __syncthreads();
        hsum = 0;
#pragma unroll
        for (int i = 0; i < NW; ++i) hsum += L.red[i];
    }
    if (HASH) {
        team_sync<NW>();
        if (out.per_vertex) {
            for (uint32_t i = tid; i < slots; i += NW * KT_WAVE) {
                uint32_t c = cnt[i];
                if (c) atomic_add_u64(&out.per_vertex[keys[i]], c);
            }
        }
    }
    if (tid == 0 && out.tail_count) out.tail_count[x] = hsum;
    if (tid == 0 && hsum) {
        if (out.per_vertex) atomic_add_u64(&out.per_vertex[V.tail_pv(x)], hsum);
        if (V.v.tail_graph)
            atomic_add_u64(&out.totals[V.tail_graph(x)], hsum);
        else
            my_total += hsum;
    }
    team_sync<NW>();  // table reuse by the next tail
}

// Streamed walks fit 64 VGPRs (8 waves per SIMD, four workgroups per CU);
// grouped walks keep more list state per lane and get 80 VGPRs (6 waves per
// SIMD), which keeps their step loop free of register spills.
template <bool PLAIN, int GL, bool LIST>
__global__ void __launch_bounds__(TC_THREADS) __attribute__((amdgpu_waves_per_eu(GL < KT_WAVE ? 6 : 8, 8))) k_tc_main(View<PLAIN> V, TcOut out, TcParams prm,
                                                         uint32_t* __restrict__ q,
                                                         const uint32_t* __restrict__ list,
                                                         uint64_t* __restrict__ stats) {
    __shared__ Lds L;
    const unsigned wave = threadIdx.x / KT_WAVE;
    const unsigned lane = lane_id();
    uint64_t my_total = 0;
    WorkStats wk;
    const int wave_bucket_max = (int)ceil_log2(prm.wave_max_deg);

    const bool by_vertex = prm.order == TC_BY_VERTEX;
    const int xcd = (int)(blockIdx.x % NXCD);
    // team phase: by degree, the workgroup-bin buckets from the largest down;
    // by vertex, this XCD's workgroup tails first, then the other XCDs'
    const int team_hi = by_vertex ? NXCD - 1 : BUCKET_HUGE - 1;
    const int team_lo = by_vertex ? 0 : wave_bucket_max + 1;
    for (int k = team_hi; k >= team_lo; --k) {
        const int b = by_vertex ? (xcd + NXCD - 1 - k) % NXCD : k;
        const uint32_t cnt = q[Q_CNT + b];
        if (cnt == 0) continue;
        const uint32_t base = bucket_base(q, b);
        uint32_t* pull = &q[Q_PULL + Q_STRIDE * b];
        while (true) {
            if (threadIdx.x == 0)
                L.bcast = (__hip_atomic_load(pull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= cnt)
                              ? cnt
                              : atomicAdd(pull, 1u);
            __syncthreads();
            const uint32_t i = L.bcast;
            __syncthreads();
            if (i >= cnt) break;
            const uint32_t x = list[base + i];
            team_process<TC_WAVES, true, PLAIN, GL, LIST>(V, x, tail_bounds(V, x), L, L.keys, L.cnt, L.ovf, GROUP_SLOTS,
                                                          prm.table_shift, prm.team_chunk, out, my_total, wk);
        }
    }
    // wave phase: independent waves with private tables; by degree, largest
    // buckets first; by vertex, this XCD's wave tails first
    {
        uint32_t* keys = L.keys + wave * WAVE_SLOTS;
        uint32_t* tcnt = L.cnt + wave * WAVE_SLOTS;
        const int wave_hi = by_vertex ? NXCD - 1 : wave_bucket_max;
        for (int k = wave_hi; k >= 0; --k) {
            const int b = by_vertex ? NXCD + (xcd + NXCD - 1 - k) % NXCD : k;
            const uint32_t cnt = q[Q_CNT + b];
            if (cnt == 0) continue;
            const uint32_t base = bucket_base(q, b);
            uint32_t* pull = &q[Q_PULL + Q_STRIDE * b];
            // more tails per grab for the small buckets (less queue traffic);
            // in vertex order, the grab set per launch
            const uint32_t grab = by_vertex ? min(max(prm.wave_grab, 1u), (uint32_t)KT_WAVE)
                                            : (b <= 3 ? 4 * WAVE_GRAB : (b <= 5 ? 2 * WAVE_GRAB : WAVE_GRAB));
            while (true) {
                uint32_t i = 0;
                if (lane == 0)
                    i = (__hip_atomic_load(pull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= cnt)
                            ? cnt
                            : atomicAdd(pull, grab);
                i = uniform(i);
                if (i >= cnt) break;
                // the grabbed tails (grab <= 64) and their bounds, loaded by
                // one lane each before the first is processed: one round
                // trip for the batch instead of two per tail
                const uint32_t ng = min(cnt, i + grab) - i;
                uint32_t xl = 0;
                TailBounds bl{0, 0, 0, 0, 0};
                if (lane < ng) {
                    xl = list[base + i + lane];
                    bl = tail_bounds(V, xl);
                }
                for (uint32_t j = 0; j < ng; ++j) {
                    const TailBounds tb{readlane64(bl.sb, j), readlane64(bl.se, j), readlane64(bl.rb, j),
                                        readlane64(bl.re, j), readlane64(bl.nstat, j)};
                    team_process<1, true, PLAIN, GL, LIST>(V, readlane32(xl, j), tb, L, keys, tcnt,
                                                           L.ovf + wave * (WAVE_SLOTS / 128), WAVE_SLOTS,
                                                           prm.table_shift, KT_WAVE, out, my_total, wk);
                }
            }
        }
    }
    // flush the wave-uniform accumulators: one atomic per workgroup and counter
    const uint64_t streamed = wk.streamed();
    __syncthreads();
    if (lane == 0) {
        L.red[wave] = my_total;
        L.keys[wave] = (uint32_t)wk.wedge;
        L.keys[TC_WAVES + wave] = (uint32_t)(wk.wedge >> 32);
        L.cnt[wave] = (uint32_t)wk.isz;
        L.cnt[TC_WAVES + wave] = (uint32_t)(wk.isz >> 32);
        L.keys[2 * TC_WAVES + wave] = (uint32_t)streamed;
        L.keys[3 * TC_WAVES + wave] = (uint32_t)(streamed >> 32);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint64_t t = 0, w = 0, z = 0, y = 0;
        for (int i = 0; i < TC_WAVES; ++i) {
            t += L.red[i];
            w += ((uint64_t)L.keys[TC_WAVES + i] << 32) | L.keys[i];
            z += ((uint64_t)L.cnt[TC_WAVES + i] << 32) | L.cnt[i];
            y += ((uint64_t)L.keys[3 * TC_WAVES + i] << 32) | L.keys[2 * TC_WAVES + i];
        }
        if (t) atomic_add_u64(&out.totals[0], t);
        if (w) atomic_add_u64(&stats[ST_WEDGE], w);
        if (z) atomic_add_u64(&stats[ST_ISZ], z);
        if (y) atomic_add_u64(&stats[ST_STREAM], y);
    }
}

// Tails whose S(x) exceeds the LDS table: 8-wave teams, binary-search probes
// of the sorted S(x) in global memory.
template <bool PLAIN, bool LIST>
__global__ void __launch_bounds__(TC_THREADS) k_tc_huge(View<PLAIN> V, TcOut out, TcParams prm,
                                                         uint32_t* __restrict__ q,
                                                         const uint32_t* __restrict__ list,
                                                         uint64_t* __restrict__ stats) {
    __shared__ Lds L;
    uint64_t my_total = 0;
    WorkStats wk;
    const uint32_t cnt = q[Q_CNT + BUCKET_HUGE];
    const uint32_t base = bucket_base(q, BUCKET_HUGE);
    uint32_t* pull = &q[Q_PULL + Q_STRIDE * BUCKET_HUGE];
    while (true) {
        if (threadIdx.x == 0)
            L.bcast = (__hip_atomic_load(pull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= cnt)
                          ? cnt
                          : atomicAdd(pull, 1u);
        __syncthreads();
        const uint32_t i = L.bcast;
        __syncthreads();
        if (i >= cnt) break;
        const uint32_t x = list[base + i];
        team_process<TC_WAVES, false, PLAIN, KT_WAVE, LIST>(V, x, tail_bounds(V, x), L, nullptr, nullptr, nullptr, 0, 0,
                                                            prm.team_chunk, out, my_total, wk);
    }
    const uint64_t streamed = wk.streamed();
    if (lane_id() == 0) {
        if (my_total) atomic_add_u64(&out.totals[0], my_total);
        if (wk.wedge) atomic_add_u64(&stats[ST_WEDGE], wk.wedge);
        if (wk.isz) atomic_add_u64(&stats[ST_ISZ], wk.isz);
        if (streamed) atomic_add_u64(&stats[ST_STREAM], streamed);
    }
}

// Tails above the workgroup bin whose hash table fits one workgroup's LDS
// (|S| <= 8192): a 16-wave team per tail with a 16384-slot table in dynamic
// LDS, one workgroup per CU.
template <bool PLAIN, bool LIST>
__global__ void __launch_bounds__(BIG_THREADS) k_tc_big(View<PLAIN> V, TcOut out, TcParams prm,
                                                        uint32_t* __restrict__ q, const uint32_t* __restrict__ list,
                                                        uint64_t* __restrict__ stats) {
    __shared__ LdsBig L;
    extern __shared__ uint32_t big_table[];  // keys[BIG_SLOTS], then cnt[BIG_SLOTS]
    uint64_t my_total = 0;
    WorkStats wk;
    const uint32_t cnt = q[Q_CNT + BUCKET_BIG];
    const uint32_t base = bucket_base(q, BUCKET_BIG);
    uint32_t* pull = &q[Q_PULL + Q_STRIDE * BUCKET_BIG];
    while (true) {
        if (threadIdx.x == 0)
            L.bcast = (__hip_atomic_load(pull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= cnt)
                          ? cnt
                          : atomicAdd(pull, 1u);
        __syncthreads();
        const uint32_t i = L.bcast;
        __syncthreads();
        if (i >= cnt) break;
        const uint32_t x = list[base + i];
        team_process<BIG_WAVES, true, PLAIN, KT_WAVE, LIST>(V, x, tail_bounds(V, x), L, big_table,
                                                            big_table + BIG_SLOTS, L.ovf, BIG_SLOTS, prm.table_shift,
                                                            prm.team_chunk, out, my_total, wk);
    }
    const uint64_t streamed = wk.streamed();
    if (lane_id() == 0) {
        if (my_total) atomic_add_u64(&out.totals[0], my_total);
        if (wk.wedge) atomic_add_u64(&stats[ST_WEDGE], wk.wedge);
        if (wk.isz) atomic_add_u64(&stats[ST_ISZ], wk.isz);
        if (streamed) atomic_add_u64(&stats[ST_STREAM], streamed);
    }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
uint64_t tc_ws_words(uint32_t X) {
    // queue header (u32) + stats (u64) + one u32 list of X
    return Q_WORDS / 2 + ST_WORDS + (uint64_t(X) + 1) / 2 + 1;
}

uint64_t* tc_stats_ptr(uint64_t* ws) { return ws + Q_WORDS / 2; }

template <bool PLAIN>
static hipError_t launch_impl(const TcView& view, const TcOut& out, const TcParams& prm, uint64_t* ws, int num_cus,
                              hipStream_t st, hipEvent_t ev_begin, hipEvent_t ev_end) {
    (void)hipGetLastError();  // a status left by an earlier call is not this launch's
    View<PLAIN> V;
    V.v = view;
    // min-side walks move triangles between tails: listing launches (and
    // their per-tail counting pass) keep every triangle at its first vertex
    if (out.tri || out.tail_count) V.v.in_off = nullptr;
    uint32_t* q = reinterpret_cast<uint32_t*>(ws);
    uint64_t* stats = ws + Q_WORDS / 2;
    uint32_t* list = reinterpret_cast<uint32_t*>(stats + ST_WORDS);
    fill_u64(ws, Q_WORDS / 2 + ST_WORDS, 0, st);
    TcParams p = prm;
    if (p.wave_max_deg > WAVE_SLOTS / 2) p.wave_max_deg = WAVE_SLOTS / 2;
    if (p.group_max_deg > GROUP_SLOTS / 2) p.group_max_deg = GROUP_SLOTS / 2;
    if (p.bsearch_min_deg > BIG_SLOTS / 2) p.bsearch_min_deg = BIG_SLOTS / 2;
    // bucket edges are powers of two: round the wave threshold up to one
    uint32_t wm = 1;
    while (wm < p.wave_max_deg) wm <<= 1;
    p.wave_max_deg = wm;
    // |S| up to this is probed in an LDS table (k_tc_main or k_tc_big)
    V.v.in_hash_max = p.bsearch_min_deg;
    // a min-side walk covers in- and out-neighbours and streams from only
    // some of them: four times the entries per team grab
    if (V.v.in_off) p.team_chunk = std::min<uint32_t>(4 * std::max<uint32_t>(p.team_chunk, 1), KT_WAVE);
    if (view.X > 0) {
        const unsigned g = std::min<unsigned>((view.X + 255) / 256, (unsigned)num_cus * 8);
        k_tc_bin_count<PLAIN><<<g, 256, 0, st>>>(V, p, q, stats);
        k_tc_bin_fill<PLAIN><<<g, 256, 0, st>>>(V, p, q, list);
    }
    unsigned grid = (unsigned)num_cus * (p.grid_per_cu ? p.grid_per_cu : 4);
    constexpr uint32_t big_lds = 2 * BIG_SLOTS * sizeof(uint32_t);
    if constexpr (PLAIN) {
        if (out.tri) {  // listing: streamed walks only
            k_tc_huge<PLAIN, true><<<(unsigned)num_cus, TC_THREADS, 0, st>>>(V, out, p, q, list, stats);
            static const hipError_t big_list_ok =
                hipFuncSetAttribute(reinterpret_cast<const void*>(&k_tc_big<PLAIN, true>),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)big_lds);
            if (big_list_ok != hipSuccess) return big_list_ok;
            k_tc_big<PLAIN, true><<<(unsigned)num_cus, BIG_THREADS, big_lds, st>>>(V, out, p, q, list, stats);
            if (ev_begin) (void)hipEventRecord(ev_begin, st);
            k_tc_main<PLAIN, KT_WAVE, true><<<grid, TC_THREADS, 0, st>>>(V, out, p, q, list, stats);
            if (ev_end) (void)hipEventRecord(ev_end, st);
            return hipGetLastError();
        }
    }
    // huge and big tails are rare: one workgroup per CU; they exit at once
    // when there are none
    k_tc_huge<PLAIN, false><<<(unsigned)num_cus, TC_THREADS, 0, st>>>(V, out, p, q, list, stats);
    static const hipError_t big_ok =
        hipFuncSetAttribute(reinterpret_cast<const void*>(&k_tc_big<PLAIN, false>),
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)big_lds);
    if (big_ok != hipSuccess) return big_ok;
    k_tc_big<PLAIN, false><<<(unsigned)num_cus, BIG_THREADS, big_lds, st>>>(V, out, p, q, list, stats);
    if (ev_begin) (void)hipEventRecord(ev_begin, st);
    // lanes per list: the whole wave (streamed) or groups of 32, 16 or 8
    switch (p.strategy) {
        case TC_GROUPED32: k_tc_main<PLAIN, 32, false><<<grid, TC_THREADS, 0, st>>>(V, out, p, q, list, stats); break;
        case TC_GROUPED16: k_tc_main<PLAIN, 16, false><<<grid, TC_THREADS, 0, st>>>(V, out, p, q, list, stats); break;
        case TC_GROUPED8: k_tc_main<PLAIN, 8, false><<<grid, TC_THREADS, 0, st>>>(V, out, p, q, list, stats); break;
        case TC_GROUPED4: k_tc_main<PLAIN, 4, false><<<grid, TC_THREADS, 0, st>>>(V, out, p, q, list, stats); break;
        case TC_GROUPED2: k_tc_main<PLAIN, 2, false><<<grid, TC_THREADS, 0, st>>>(V, out, p, q, list, stats); break;
        default: k_tc_main<PLAIN, KT_WAVE, false><<<grid, TC_THREADS, 0, st>>>(V, out, p, q, list, stats); break;
    }
    if (ev_end) (void)hipEventRecord(ev_end, st);
    return hipGetLastError();  // a launch that did not start (e.g. LDS beyond the device's limit)
}

hipError_t tc_launch(const TcView& view, const TcOut& out, const TcParams& prm, uint64_t* ws, int num_cus,
                     hipStream_t st, hipEvent_t ev_begin, hipEvent_t ev_end) {
    const bool plain = view.s_beg == nullptr && view.r_beg == nullptr && view.rheads == nullptr &&
                       view.t_beg == nullptr && view.theads == nullptr && view.tail_id == nullptr;
    return plain ? launch_impl<true>(view, out, prm, ws, num_cus, st, ev_begin, ev_end)
                 : launch_impl<false>(view, out, prm, ws, num_cus, st, ev_begin, ev_end);
}

}  // namespace kt
