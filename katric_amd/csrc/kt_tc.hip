// Per-edge set intersection on the oriented CSR -- the hot loop of the local
// phase.  Replaces CetricEdgeIterator::run_local_* (counters/
// cetric_edge_iterator.h:229-388, 888-1066) and cetric::intersection /
// merge_intersection (counters/intersection.h:40-96) with a degree-binned
// hash-probe design for CDNA (wave64, LDS):
//
//   * a "team" (one wave, or the four waves of a workgroup for heavier tails)
//     takes a tail x and builds a linear-probing hash of S(x) = N+(x) in LDS;
//   * the team walks R(x) = N+(x) in chunks of 64 neighbours u; the lists
//     T(u) = N+(u) of a chunk are concatenated and probed 64 elements per
//     step, one per lane (load-balanced by a marker/max-scan owner search),
//     so short and long T(u) lists both keep all lanes busy;
//   * every hit w is one emitted Triangle{x, u, w}; its per-vertex counts are
//     privatised -- per-slot LDS counters for w, per-lane counters for u, a
//     register for x -- and flushed with one global atomic per non-zero
//     counter.
//   * tails whose S(x) exceeds the LDS table are probed by binary search in
//     the sorted S(x) in global memory instead.
//
// Counts are exact integers; the hash only changes where a set member is
// stored, never which members match.
#include "kt_device.h"
#include "kt_internal.h"

namespace kt {

constexpr int TC_THREADS = 256;
constexpr int TC_WAVES = TC_THREADS / KT_WAVE;
constexpr uint32_t GROUP_SLOTS = 4096;               // shared table of a 4-wave team
constexpr uint32_t WAVE_SLOTS = GROUP_SLOTS / TC_WAVES;  // private table of a 1-wave team
constexpr uint32_t MIN_LOGS = 6;

// work-queue header layout (u32 words at the start of the workspace)
enum : int {
    Q_CNT_HUGE = 0,
    Q_CNT_GROUP = 1,
    Q_CNT_WAVE = 2,
    Q_PULL_HUGE = 3,
    Q_PULL_GROUP = 4,
    Q_PULL_WAVE = 5,
    Q_WORDS = 16
};
// stats accumulators (u64) follow the queue header
enum : int { ST_WEDGE = 0, ST_ISZ = 1, ST_WORDS = 4 };

template <bool PLAIN>
struct View {
    TcView v;
    __device__ __forceinline__ void s_range(uint32_t x, uint64_t& b, uint64_t& e) const {
        b = v.s_off[x];
        e = v.s_off[x + 1];
    }
    __device__ __forceinline__ void r_range(uint32_t x, uint64_t& b, uint64_t& e) const {
        if (PLAIN || v.r_beg == nullptr) {
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
    __device__ __forceinline__ uint32_t t_head(uint64_t i) const { return PLAIN ? v.sheads[i] : v.theads[i]; }
    __device__ __forceinline__ uint32_t tail_pv(uint32_t x) const {
        return (PLAIN || v.tail_id == nullptr) ? x : v.tail_id[x];
    }
    __device__ __forceinline__ uint32_t tail_graph(uint32_t x) const {
        return v.tail_graph ? v.tail_graph[x] : 0u;
    }
};

__device__ __forceinline__ uint32_t ceil_log2(uint64_t x) { return x <= 1 ? 0 : 64 - __clzll(x - 1); }

// ---------------------------------------------------------------------------
// binning: split tails into huge / group / wave lists
// ---------------------------------------------------------------------------
__device__ __forceinline__ void wave_append(uint32_t* list, uint32_t* cnt, bool pred, uint32_t item) {
    uint64_t m = __ballot(pred);
    if (m == 0) return;
    uint32_t leader = __ffsll((unsigned long long)m) - 1;
    uint32_t base = 0;
    if (lane_id() == leader) base = atomicAdd(cnt, (uint32_t)__popcll(m));
    base = __shfl(base, leader, 64);
    if (pred) list[base + __popcll(m & lanemask_lt())] = item;
}

template <bool PLAIN>
__global__ void __launch_bounds__(256) k_tc_bin(View<PLAIN> V, TcParams prm, uint32_t* __restrict__ q,
                                                uint32_t* __restrict__ l_huge, uint32_t* __restrict__ l_group,
                                                uint32_t* __restrict__ l_wave, uint64_t* __restrict__ stats) {
    uint32_t x = blockIdx.x * blockDim.x + threadIdx.x;
    bool in = x < V.v.X;
    uint64_t sb = 0, se = 0, rb = 0, re = 0;
    if (in) {
        V.s_range(x, sb, se);
        V.r_range(x, rb, re);
    }
    uint64_t s = se - sb, r = re - rb;
    bool active = in && s > 0 && r > 0 && !(PLAIN && s < 2);
    // a plain tail with |N+(x)| == 1 emits nothing but still counts one
    // intersection in the reference's counters (cetric_edge_iterator.h:279-281)
    uint64_t wedge = 0, isz = 0;
    if (PLAIN && in && s == 1) {
        uint64_t tb, te;
        V.t_range(V.r_head(rb), tb, te);
        wedge = te - tb;
        isz = 1 + (te - tb);
    }
    wedge = wave_sum(wedge);
    isz = wave_sum(isz);
    if (lane_id() == 0 && (wedge | isz)) {
        atomic_add_u64(&stats[ST_WEDGE], wedge);
        atomic_add_u64(&stats[ST_ISZ], isz);
    }
    bool huge = active && s > prm.group_max_deg;
    bool group = active && !huge && s > prm.wave_max_deg;
    bool wave = active && !huge && !group;
    wave_append(l_huge, &q[Q_CNT_HUGE], huge, x);
    wave_append(l_group, &q[Q_CNT_GROUP], group, x);
    wave_append(l_wave, &q[Q_CNT_WAVE], wave, x);
}

// ---------------------------------------------------------------------------
// the team worker
// ---------------------------------------------------------------------------
struct Lds {
    uint32_t keys[GROUP_SLOTS];
    uint32_t cnt[GROUP_SLOTS];
    int marker[TC_WAVES][KT_WAVE];
    uint32_t ucnt[TC_WAVES][KT_WAVE];
    uint64_t red[TC_WAVES];
    uint32_t bcast;
};

template <int NW>
__device__ __forceinline__ void team_sync() {
    if (NW > 1)
        __syncthreads();
    else
        wave_sync();
}

// Probe policy: LDS hash table (HASH) or binary search of the sorted S(x) in
// global memory (huge tails).
template <bool HASH>
struct Probe {
    uint32_t* keys;
    uint32_t* cnt;
    uint32_t logS;
    uint32_t mask;
    const uint32_t* s;  // sorted S(x) (global) for !HASH
    uint64_t ns;
    uint64_t* per_vertex;

    // Returns true on hit and records the w-side count.
    __device__ __forceinline__ bool probe(uint32_t w) const {
        if (HASH) {
            uint32_t h = hash_slot(w, logS);
            while (true) {
                uint32_t k = keys[h];
                if (k == w) {
                    atomicAdd(&cnt[h], 1u);
                    return true;
                }
                if (k == KT_EMPTY) return false;
                h = (h + 1) & mask;
            }
        } else {
            uint64_t lo = 0, hi = ns;
            while (lo < hi) {
                uint64_t mid = (lo + hi) >> 1;
                if (s[mid] < w)
                    lo = mid + 1;
                else
                    hi = mid;
            }
            if (lo < ns && s[lo] == w) {
                if (per_vertex) atomic_add_u64(&per_vertex[w], 1);
                return true;
            }
            return false;
        }
    }
};

// Process tail x with a team of NW waves.  Returns (in lane 0 of wave 0 of
// the team) nothing; accumulates per-lane hits/stats into the references.
template <int NW, bool HASH, bool PLAIN>
__device__ __forceinline__ void team_process(const View<PLAIN>& V, uint32_t x, Lds& L, uint32_t* keys,
                                             uint32_t* cnt, uint32_t cap_slots, const TcOut& out,
                                             uint64_t& my_total, uint64_t& st_wedge, uint64_t& st_isz) {
    const unsigned lane = lane_id();
    const unsigned wave = threadIdx.x / KT_WAVE;
    const unsigned twave = NW > 1 ? wave : 0;  // wave index within the team
    const unsigned tid = twave * KT_WAVE + lane;
    int* marker = L.marker[wave];
    uint32_t* ucnt = L.ucnt[wave];

    uint64_t sb, se, rb, re;
    V.s_range(x, sb, se);
    V.r_range(x, rb, re);
    const uint64_t ns = se - sb, nr = re - rb;

    Probe<HASH> P;
    P.per_vertex = out.per_vertex;
    P.s = V.v.sheads + sb;
    P.ns = ns;
    P.keys = keys;
    P.cnt = cnt;
    P.logS = 0;
    P.mask = 0;
    uint32_t slots = 0;
    if (HASH) {
        uint32_t logS = ceil_log2(2 * ns);
        if (logS < MIN_LOGS) logS = MIN_LOGS;
        slots = 1u << logS;
        if (slots > cap_slots) slots = cap_slots;  // never happens: bins respect the cap
        P.logS = logS;
        P.mask = slots - 1;
        for (uint32_t i = tid; i < slots; i += NW * KT_WAVE) {
            keys[i] = KT_EMPTY;
            cnt[i] = 0;
        }
        team_sync<NW>();
        for (uint64_t i = tid; i < ns; i += NW * KT_WAVE) {
            uint32_t w = V.v.sheads[sb + i];
            uint32_t h = hash_slot(w, logS);
            while (true) {
                uint32_t old = atomicCAS(&keys[h], KT_EMPTY, w);
                if (old == KT_EMPTY || old == w) break;
                h = (h + 1) & P.mask;
            }
        }
        team_sync<NW>();
    }

    uint64_t hits = 0;
    for (uint64_t c = (uint64_t)twave * KT_WAVE; c < nr; c += (uint64_t)NW * KT_WAVE) {
        const uint64_t idx = rb + c + lane;
        const bool valid = c + lane < nr;
        uint32_t u = 0;
        uint64_t tb = 0, te = 0;
        if (valid) {
            u = V.r_head(idx);
            V.t_range(u, tb, te);
        }
        const uint64_t du = te - tb;
        st_wedge += du;
        st_isz += valid ? ns + du : 0;
        const uint64_t incl = wave_incl_sum(du);
        const uint64_t start = incl - du;
        const uint64_t total = __shfl(incl, KT_WAVE - 1, 64);
        ucnt[lane] = 0;
        for (uint64_t t = 0; t < total; t += KT_WAVE) {
            marker[lane] = -1;
            wave_sync();
            if (du > 0 && start >= t && start < t + KT_WAVE) marker[start - t] = (int)lane;
            wave_sync();
            const int mk = marker[lane];
            const uint64_t ball = __ballot(incl > t);
            const int owner_t = __ffsll((unsigned long long)ball) - 1;
            const int own = max(owner_t, wave_incl_max(mk));
            const uint64_t p = t + lane;
            const uint64_t own_tb = __shfl(tb, own, 64);
            const uint64_t own_start = __shfl(start, own, 64);
            if (p < total) {
                const uint32_t w = V.t_head(own_tb + (p - own_start));
                if (P.probe(w)) atomicAdd(&ucnt[own], 1u);
            }
            wave_sync();
        }
        const uint32_t cu = ucnt[lane];
        if (valid && cu) {
            hits += cu;
            if (out.per_vertex) atomic_add_u64(&out.per_vertex[u], cu);
        }
        wave_sync();
    }

    // team reduction of hits for the tail's own count and the graph total
    uint64_t h = wave_sum(hits);
    if (NW > 1) {
        if (lane == 0) L.red[wave] = h;
        __syncthreads();
        h = 0;
#pragma unroll
        for (int i = 0; i < NW; ++i) h += L.red[i];
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
    if (tid == 0 && h) {
        if (out.per_vertex) atomic_add_u64(&out.per_vertex[V.tail_pv(x)], h);
        if (V.v.tail_graph)
            atomic_add_u64(&out.totals[V.tail_graph(x)], h);
        else
            my_total += h;
    }
    team_sync<NW>();  // table / marker reuse by the next tail
}

template <bool PLAIN>
__global__ void __launch_bounds__(TC_THREADS) k_tc_main(View<PLAIN> V, TcOut out, uint32_t* __restrict__ q,
                                                         const uint32_t* __restrict__ l_huge,
                                                         const uint32_t* __restrict__ l_group,
                                                         const uint32_t* __restrict__ l_wave,
                                                         uint64_t* __restrict__ stats) {
    __shared__ Lds L;
    const unsigned wave = threadIdx.x / KT_WAVE;
    const unsigned lane = lane_id();
    uint64_t my_total = 0, st_wedge = 0, st_isz = 0;

    // phase 0: huge tails, 4-wave team, binary search probes
    {
        const uint32_t cnt = q[Q_CNT_HUGE];
        while (true) {
            if (threadIdx.x == 0) L.bcast = atomicAdd(&q[Q_PULL_HUGE], 1u);
            __syncthreads();
            const uint32_t i = L.bcast;
            __syncthreads();
            if (i >= cnt) break;
            team_process<TC_WAVES, false, PLAIN>(V, l_huge[i], L, nullptr, nullptr, 0, out, my_total, st_wedge,
                                                 st_isz);
        }
    }
    // phase 1: group tails, 4-wave team, shared LDS table
    {
        const uint32_t cnt = q[Q_CNT_GROUP];
        while (true) {
            if (threadIdx.x == 0) L.bcast = atomicAdd(&q[Q_PULL_GROUP], 1u);
            __syncthreads();
            const uint32_t i = L.bcast;
            __syncthreads();
            if (i >= cnt) break;
            team_process<TC_WAVES, true, PLAIN>(V, l_group[i], L, L.keys, L.cnt, GROUP_SLOTS, out, my_total,
                                                st_wedge, st_isz);
        }
    }
    // phase 2: wave tails, independent waves with private LDS tables
    {
        const uint32_t cnt = q[Q_CNT_WAVE];
        uint32_t* keys = L.keys + wave * WAVE_SLOTS;
        uint32_t* tcnt = L.cnt + wave * WAVE_SLOTS;
        while (true) {
            uint32_t i = 0;
            if (lane == 0) i = atomicAdd(&q[Q_PULL_WAVE], 1u);
            i = __shfl(i, 0, 64);
            if (i >= cnt) break;
            team_process<1, true, PLAIN>(V, l_wave[i], L, keys, tcnt, WAVE_SLOTS, out, my_total, st_wedge,
                                         st_isz);
        }
    }
    // flush per-lane accumulators
    my_total = wave_sum(my_total);
    st_wedge = wave_sum(st_wedge);
    st_isz = wave_sum(st_isz);
    if (lane == 0) {
        if (my_total) atomic_add_u64(&out.totals[0], my_total);
        if (st_wedge) atomic_add_u64(&stats[ST_WEDGE], st_wedge);
        if (st_isz) atomic_add_u64(&stats[ST_ISZ], st_isz);
    }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
uint64_t tc_ws_words(uint32_t X) {
    // queue header (u32) + stats (u64) + three u32 lists of X
    return Q_WORDS / 2 + ST_WORDS + (3ull * X + 1) / 2 + 1;
}

uint64_t* tc_stats_ptr(uint64_t* ws) { return ws + Q_WORDS / 2; }

template <bool PLAIN>
static void launch_impl(const TcView& view, const TcOut& out, const TcParams& prm, uint64_t* ws, int num_cus,
                        hipStream_t st, hipEvent_t ev_begin, hipEvent_t ev_end) {
    View<PLAIN> V;
    V.v = view;
    uint32_t* q = reinterpret_cast<uint32_t*>(ws);
    uint64_t* stats = ws + Q_WORDS / 2;
    uint32_t* lists = reinterpret_cast<uint32_t*>(stats + ST_WORDS);
    uint32_t* l_huge = lists;
    uint32_t* l_group = lists + view.X;
    uint32_t* l_wave = lists + 2ull * view.X;
    fill_u64(ws, Q_WORDS / 2 + ST_WORDS, 0, st);
    TcParams p = prm;
    if (p.wave_max_deg > WAVE_SLOTS / 2) p.wave_max_deg = WAVE_SLOTS / 2;
    if (p.group_max_deg > GROUP_SLOTS / 2) p.group_max_deg = GROUP_SLOTS / 2;
    if (view.X > 0) k_tc_bin<PLAIN><<<(view.X + 255) / 256, 256, 0, st>>>(V, p, q, l_huge, l_group, l_wave, stats);
    unsigned grid = (unsigned)num_cus * (p.grid_per_cu ? p.grid_per_cu : 4);
    if (ev_begin) hipEventRecord(ev_begin, st);
    k_tc_main<PLAIN><<<grid, TC_THREADS, 0, st>>>(V, out, q, l_huge, l_group, l_wave, stats);
    if (ev_end) hipEventRecord(ev_end, st);
}

void tc_launch(const TcView& view, const TcOut& out, const TcParams& prm, uint64_t* ws, int num_cus,
               hipStream_t st, hipEvent_t ev_begin, hipEvent_t ev_end) {
    const bool plain = view.r_beg == nullptr && view.rheads == nullptr && view.t_beg == nullptr &&
                       view.theads == nullptr && view.tail_id == nullptr;
    if (plain)
        launch_impl<true>(view, out, prm, ws, num_cus, st, ev_begin, ev_end);
    else
        launch_impl<false>(view, out, prm, ws, num_cus, st, ev_begin, ev_end);
}

}  // namespace kt
