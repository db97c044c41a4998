/*
 * CPU oracle for the per-rank triangle-counting core -- TEST INFRASTRUCTURE.
 *
 * Restates the reference (KaTric) algorithm for the local counting path in
 * plain C.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg load this library, as the checker or as the timed CPU baseline; the
 * product (katric_amd) never does.
 *
 * What pins each function (all checks live in tests/test_oracle.py):
 *   or_*_intersection : the reference's hand vectors and its random-vs-
 *                       std::set_intersection test (tests/intersection_test.cpp
 *                       :41-52, :101-173), restated against Python's set
 *                       intersection; binary compared unordered as there.
 *   or_log2           : closed form floor(log2) over all powers of two +-1.
 *   or_orient         : the orientation invariants of
 *                       tests/orientation_test.cpp:81-133 (out u in = N(v),
 *                       every undirected edge kept once, lists sorted), the
 *                       reference fixture examples/metis-sample.first_out/.head.
 *   or_run_shmem      : known totals examples/examples.yaml (triangle=1,
 *                       metis-sample=4, hiv=45); brute force (or_brute_force)
 *                       and networkx.triangles golden vectors for per-vertex
 *                       counts; sum_x t[x] = 3 * total; K_n closed form
 *                       C(n,3), per-vertex C(n-1,2); relabelling invariance.
 *   or_brute_force    : K_n / empty / star / wheel closed forms; networkx.
 *   or_rank_sim       : or_rank_closed_form per rank for P in {1,2,3,4,16};
 *                       sum over ranks = total; per-vertex sums over ranks
 *                       equal the single-rank per-vertex counts.
 *   or_rank_closed_form: brute-force triangle enumeration by owner pattern
 *                       in tests; P = 1 gives local = total.
 */
#include "oracle.h"

#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------ */
/* small helpers                                                       */
/* ------------------------------------------------------------------ */

uint64_t or_log2(uint64_t x) {
    /* cetric::log2, include/cetric/util.h:136-150 (8*sizeof - clzl(x) - 1). */
    if (x == 0) return 0;
    return 63u - (uint64_t)__builtin_clzll(x);
}

static int cmp_u32(const void* a, const void* b) {
    uint32_t x = *(const uint32_t*)a, y = *(const uint32_t*)b;
    return (x > y) - (x < y);
}
static int cmp_u64(const void* a, const void* b) {
    uint64_t x = *(const uint64_t*)a, y = *(const uint64_t*)b;
    return (x > y) - (x < y);
}

static inline void add_pv(uint64_t* pv, uint32_t x, int atomic) {
    if (!pv) return;
    if (atomic)
        __atomic_fetch_add(&pv[x], 1, __ATOMIC_RELAXED);
    else
        pv[x] += 1;
}

/* ------------------------------------------------------------------ */
/* intersection kernels, include/cetric/counters/intersection.h        */
/* Generated for uint32 (public) and uint64 keys (rank-encoded sort    */
/* orders in the rank simulation).                                     */
/* ------------------------------------------------------------------ */

#define OR_DEFINE_ISECT(T, SFX)                                                                              \
    /* merge_intersection, intersection.h:81-96 */                                                            \
    static size_t merge_##SFX(const T* a, size_t na, const T* b, size_t nb, T* out) {                         \
        size_t i = 0, j = 0, k = 0;                                                                           \
        while (i != na && j != nb) {                                                                          \
            if (a[i] < b[j])                                                                                  \
                ++i;                                                                                          \
            else if (b[j] < a[i])                                                                             \
                ++j;                                                                                          \
            else {                                                                                            \
                if (out) out[k] = a[i];                                                                       \
                ++k;                                                                                          \
                ++i;                                                                                          \
                ++j;                                                                                          \
            }                                                                                                 \
        }                                                                                                     \
        return k;                                                                                             \
    }                                                                                                         \
    static size_t lower_bound_##SFX(const T* a, size_t lo, size_t hi, T key) {                                \
        while (lo < hi) {                                                                                     \
            size_t mid = lo + (hi - lo) / 2;                                                                  \
            if (a[mid] < key)                                                                                 \
                lo = mid + 1;                                                                                 \
            else                                                                                              \
                hi = mid;                                                                                     \
        }                                                                                                     \
        return lo;                                                                                            \
    }                                                                                                         \
    /* binary_search_intersection, intersection.h:98-117: the shorter list probes the longer one */          \
    static size_t bsi_##SFX(const T* a, size_t na, const T* b, size_t nb, T* out) {                           \
        size_t k = 0;                                                                                         \
        const T *s = a, *l = b;                                                                               \
        size_t ns = na, nl = nb;                                                                              \
        if (!(na <= nb)) {                                                                                    \
            s = b; l = a; ns = nb; nl = na;                                                                   \
        }                                                                                                     \
        for (size_t i = 0; i < ns; ++i) {                                                                     \
            size_t p = lower_bound_##SFX(l, 0, nl, s[i]);                                                     \
            if (p != nl && !(s[i] < l[p])) {                                                                  \
                if (out) out[k] = s[i];                                                                       \
                ++k;                                                                                          \
            }                                                                                                 \
        }                                                                                                     \
        return k;                                                                                             \
    }                                                                                                         \
    /* binary_intersection, iterative stack form, intersection.h:129-205 */                                  \
    static size_t bin_##SFX(const T* A, size_t na, const T* B, size_t nb, T* out, size_t cutoff) {            \
        typedef struct { size_t f1, l1, f2, l2; } task_t;                                                     \
        size_t cap = 64, top = 0, k = 0;                                                                      \
        task_t* st = (task_t*)malloc(cap * sizeof(task_t));                                                   \
        st[top++] = (task_t){0, na, 0, nb};                                                                   \
        while (top) {                                                                                         \
            task_t t = st[--top];                                                                             \
            size_t d1 = t.l1 - t.f1, d2 = t.l2 - t.f2;                                                        \
            if (d1 <= cutoff && d2 <= cutoff) {                                                               \
                k += merge_##SFX(A + t.f1, d1, B + t.f2, d2, out ? out + k : NULL);                           \
                continue;                                                                                     \
            }                                                                                                 \
            if (d1 == 0 || d2 == 0) continue;                                                                 \
            if (top + 2 > cap) {                                                                              \
                cap *= 2;                                                                                     \
                st = (task_t*)realloc(st, cap * sizeof(task_t));                                              \
            }                                                                                                 \
            if (d1 <= d2) {                                                                                   \
                size_t p1 = t.f1 + d1 / 2;                                                                    \
                size_t p2 = lower_bound_##SFX(B, t.f2, t.l2, A[p1]);                                          \
                if (!(p1 == t.l1 && p2 == t.l2)) st[top++] = (task_t){t.f1, p1, t.f2, p2};                    \
                if (p2 != t.l2 && !(A[p1] < B[p2])) {                                                         \
                    if (out) out[k] = A[p1];                                                                  \
                    ++k;                                                                                      \
                    ++p1;                                                                                     \
                    ++p2;                                                                                     \
                }                                                                                             \
                if (!(p1 == t.f1 && p2 == t.f2)) st[top++] = (task_t){p1, t.l1, p2, t.l2};                    \
            } else {                                                                                          \
                size_t p2 = t.f2 + d2 / 2;                                                                    \
                size_t p1 = lower_bound_##SFX(A, t.f1, t.l1, B[p2]);                                          \
                if (!(p2 == t.l2 && p1 == t.l1)) st[top++] = (task_t){t.f1, p1, t.f2, p2};                    \
                if (p1 != t.l1 && !(B[p2] < A[p1])) {                                                         \
                    if (out) out[k] = B[p2];                                                                  \
                    ++k;                                                                                      \
                    ++p1;                                                                                     \
                    ++p2;                                                                                     \
                }                                                                                             \
                if (!(p2 == t.f2 && p1 == t.f1)) st[top++] = (task_t){p1, t.l1, p2, t.l2};                    \
            }                                                                                                 \
        }                                                                                                     \
        free(st);                                                                                             \
        return k;                                                                                             \
    }                                                                                                         \
    /* dispatcher + hybrid rule, intersection.h:40-79 */                                                      \
    static inline size_t isect_##SFX(const T* a, size_t na, const T* b, size_t nb, T* out, int method, size_t cutoff, \
                              double scale) {                                                                 \
        switch (method) {                                                                                     \
            case OR_BINARY_SEARCH: return bsi_##SFX(a, na, b, nb, out);                                       \
            case OR_BINARY: return bin_##SFX(a, na, b, nb, out, cutoff);                                      \
            case OR_HYBRID: {                                                                                 \
                uint64_t d1 = na, d2 = nb;                                                                    \
                if (d2 < d1) {                                                                                \
                    uint64_t t = d1; d1 = d2; d2 = t;                                                         \
                }                                                                                             \
                if ((double)(d1 * or_log2(d2)) < scale * (double)(d1 + d2))                                   \
                    return bsi_##SFX(a, na, b, nb, out);                                                      \
                return merge_##SFX(a, na, b, nb, out);                                                        \
            }                                                                                                 \
            default: return merge_##SFX(a, na, b, nb, out);                                                   \
        }                                                                                                     \
    }

OR_DEFINE_ISECT(uint32_t, u32)
OR_DEFINE_ISECT(uint64_t, u64)

size_t or_merge_intersection(const uint32_t* a, size_t na, const uint32_t* b, size_t nb, uint32_t* out) {
    return merge_u32(a, na, b, nb, out);
}
size_t or_binary_search_intersection(const uint32_t* a, size_t na, const uint32_t* b, size_t nb, uint32_t* out) {
    return bsi_u32(a, na, b, nb, out);
}
size_t or_binary_intersection(const uint32_t* a, size_t na, const uint32_t* b, size_t nb, uint32_t* out,
                              size_t cutoff) {
    return bin_u32(a, na, b, nb, out, cutoff);
}
size_t or_intersection(const uint32_t* a, size_t na, const uint32_t* b, size_t nb, uint32_t* out, int method,
                       size_t binary_cutoff, double hybrid_cutoff_scale) {
    return isect_u32(a, na, b, nb, out, method, binary_cutoff, hybrid_cutoff_scale);
}

/* ------------------------------------------------------------------ */
/* orientation, distributed_graph.h:537-566 / sort :577-593            */
/* ------------------------------------------------------------------ */

/* node_ordering::degree_outward on one rank (distributed_graph.h:147-163):
 * (degree, id) lexicographic; every vertex is local. */
static inline int deg_less(const uint64_t* off, uint32_t a, uint32_t b) {
    uint64_t da = off[a + 1] - off[a], db = off[b + 1] - off[b];
    return da < db || (da == db && a < b);
}

/* Two-pointer partition of one adjacency list: in-neighbours first, out-
 * neighbours (cmp(v,u) true) last; returns first_out_offset (in-degree). */
static uint64_t orient_one(const uint64_t* off, uint32_t v, uint32_t* h, uint64_t deg) {
    int64_t left = 0, right = (int64_t)deg - 1;
    while (left <= right) {
        while (left <= right && !deg_less(off, v, h[left])) left++;
        while (right >= 0 && deg_less(off, v, h[right])) right--;
        if (left <= right) {
            uint32_t t = h[left];
            h[left] = h[right];
            h[right] = t;
        }
    }
    return (uint64_t)left;
}

/* Oriented copy: per vertex, orient then sort both parts by id.
 * fo[v] = first_out_offset (in-degree) within the full list. */
static uint32_t* orient_full(uint32_t n, const uint64_t* off, const uint32_t* heads, uint64_t* fo, int threads) {
    uint64_t nnz = off[n];
    uint32_t* h = (uint32_t*)malloc((nnz ? nnz : 1) * sizeof(uint32_t));
    memcpy(h, heads, nnz * sizeof(uint32_t));
    (void)threads;
#pragma omp parallel for schedule(dynamic, 1024) num_threads(threads > 1 ? threads : 1)
    for (int64_t v = 0; v < (int64_t)n; ++v) {
        uint64_t b = off[v], d = off[v + 1] - off[v];
        uint64_t k = orient_one(off, (uint32_t)v, h + b, d);
        fo[v] = k;
        qsort(h + b, k, sizeof(uint32_t), cmp_u32);
        qsort(h + b + k, d - k, sizeof(uint32_t), cmp_u32);
    }
    return h;
}

int or_orient(uint32_t n, const uint64_t* offsets, const uint32_t* heads, uint64_t* out_off, uint32_t* out_heads) {
    uint64_t* fo = (uint64_t*)malloc(((size_t)n + 1) * sizeof(uint64_t));
    uint32_t* h = orient_full(n, offsets, heads, fo, 1);
    uint64_t pos = 0;
    for (uint32_t v = 0; v < n; ++v) {
        out_off[v] = pos;
        uint64_t b = offsets[v] + fo[v], e = offsets[v + 1];
        memcpy(out_heads + pos, h + b, (e - b) * sizeof(uint32_t));
        pos += e - b;
    }
    out_off[n] = pos;
    free(h);
    free(fo);
    return (2 * pos == offsets[n]) ? 0 : -1;
}

/* ------------------------------------------------------------------ */
/* shmetric, shared_memory_cetric.h:71-170                             */
/* ------------------------------------------------------------------ */

int or_run_shmem(uint32_t n, const uint64_t* offsets, const uint32_t* heads, int method, int threads,
                 uint64_t* total, uint64_t* per_vertex, or_stats* stats) {
    return or_run_shmem_ex(n, offsets, heads, method, threads, 0, total, per_vertex, stats);
}

int or_run_shmem_ex(uint32_t n, const uint64_t* offsets, const uint32_t* heads, int method, int threads,
                    int pseudo2core, uint64_t* total, uint64_t* per_vertex, or_stats* stats) {
    uint64_t* fo = (uint64_t*)malloc(((size_t)n + 1) * sizeof(uint64_t));
    uint32_t* h = orient_full(n, offsets, heads, fo, threads);
    if (per_vertex) memset(per_vertex, 0, (size_t)n * sizeof(uint64_t));
    uint64_t tri = 0, wedge = 0, isz = 0, skipped = 0;
    const int atomic = threads > 1;
    /* run_local_sequential (cetric_edge_iterator.h:229-298) / node_loop_body
     * of run_local_parallel (:993-1023): for every v, for every u in N+(v),
     * intersect N+(v) with N+(u) and emit {v,u,w} per common w.  offset = 0:
     * NodeOrdering (id_outward) differs from EdgeOrientationOrdering. */
#pragma omp parallel num_threads(threads > 1 ? threads : 1) reduction(+ : tri, wedge, isz, skipped)
    {
        uint32_t* buf = NULL;
        size_t cap = 0;
#pragma omp for schedule(dynamic, 256)
        for (int64_t vv = 0; vv < (int64_t)n; ++vv) {
            uint32_t v = (uint32_t)vv;
            const uint32_t* nv = h + offsets[v] + fo[v];
            uint64_t dv = offsets[v + 1] - offsets[v] - fo[v];
            /* pseudo2core: cetric_edge_iterator.h:239-242 */
            if (pseudo2core && dv < 2) {
                skipped++;
                continue;
            }
            for (uint64_t i = 0; i < dv; ++i) {
                uint32_t u = nv[i];
                const uint32_t* nu = h + offsets[u] + fo[u];
                uint64_t du = offsets[u + 1] - offsets[u] - fo[u];
                wedge += du;
                isz += dv + du;
                size_t need = dv < du ? dv : du;
                if (need > cap) {
                    cap = need * 2;
                    buf = (uint32_t*)realloc(buf, cap * sizeof(uint32_t));
                }
                size_t k = isect_u32(nv, dv, nu, du, buf, method, 1000, 1.0);
                tri += k;
                if (per_vertex && k) {
                    for (size_t j = 0; j < k; ++j) add_pv(per_vertex, buf[j], atomic);
                    if (atomic) {
                        __atomic_fetch_add(&per_vertex[v], k, __ATOMIC_RELAXED);
                        __atomic_fetch_add(&per_vertex[u], k, __ATOMIC_RELAXED);
                    } else {
                        per_vertex[v] += k;
                        per_vertex[u] += k;
                    }
                }
            }
        }
        free(buf);
    }
    *total = tri;
    if (stats) {
        memset(stats, 0, sizeof(*stats));
        stats->wedge_checks = wedge;
        stats->intersection_size_local = isz;
        stats->local_triangles = tri;
        stats->skipped_nodes = skipped;
    }
    free(h);
    free(fo);
    return 0;
}

/* One-thread shmetric over a sample of tails, for timing: preprocessing
 * (orient + sort) of the whole graph, then run_local_sequential
 * (cetric_edge_iterator.h:229-298) over the tails v with v % stride == 0.
 * Reports both phases' seconds and the oriented edges of the sampled tails. */
int or_run_shmem_sampled(uint32_t n, const uint64_t* offsets, const uint32_t* heads, uint32_t stride,
                         uint64_t* sample_triangles, uint64_t* sample_edges, double* prep_s, double* local_s) {
    if (stride == 0) return -1;
    double t0 = omp_get_wtime();
    uint64_t* fo = (uint64_t*)malloc(((size_t)n + 1) * sizeof(uint64_t));
    uint32_t* h = orient_full(n, offsets, heads, fo, 1);
    double t1 = omp_get_wtime();
    uint64_t tri = 0, edges = 0;
    uint32_t* buf = NULL;
    size_t cap = 0;
    for (uint64_t v = 0; v < n; v += stride) {
        const uint32_t* nv = h + offsets[v] + fo[v];
        uint64_t dv = offsets[v + 1] - offsets[v] - fo[v];
        edges += dv;
        for (uint64_t i = 0; i < dv; ++i) {
            uint32_t u = nv[i];
            const uint32_t* nu = h + offsets[u] + fo[u];
            uint64_t du = offsets[u + 1] - offsets[u] - fo[u];
            size_t need = dv < du ? dv : du;
            if (need > cap) {
                cap = need * 2;
                buf = (uint32_t*)realloc(buf, cap * sizeof(uint32_t));
            }
            tri += isect_u32(nv, dv, nu, du, buf, OR_MERGE, 1000, 1.0);
        }
    }
    double t2 = omp_get_wtime();
    free(buf);
    free(h);
    free(fo);
    *sample_triangles = tri;
    *sample_edges = edges;
    *prep_s = t1 - t0;
    *local_s = t2 - t1;
    return 0;
}

/* ------------------------------------------------------------------ */
/* brute force                                                         */
/* ------------------------------------------------------------------ */

int or_brute_force(uint32_t n, const uint64_t* offsets, const uint32_t* heads, uint64_t* total, uint64_t* per_vertex) {
    if (n > 4096) return -1;
    size_t words = ((size_t)n + 63) / 64;
    uint64_t* bits = (uint64_t*)calloc((size_t)n * words + 1, sizeof(uint64_t));
    for (uint32_t v = 0; v < n; ++v)
        for (uint64_t e = offsets[v]; e < offsets[v + 1]; ++e) {
            uint32_t u = heads[e];
            if (u == v) continue;
            bits[(size_t)v * words + u / 64] |= 1ull << (u % 64);
            bits[(size_t)u * words + v / 64] |= 1ull << (v % 64);
        }
    if (per_vertex) memset(per_vertex, 0, (size_t)n * sizeof(uint64_t));
    uint64_t t = 0;
    for (uint32_t a = 0; a < n; ++a)
        for (uint32_t b = a + 1; b < n; ++b) {
            if (!(bits[(size_t)a * words + b / 64] >> (b % 64) & 1)) continue;
            for (uint32_t c = b + 1; c < n; ++c) {
                if ((bits[(size_t)a * words + c / 64] >> (c % 64) & 1) &&
                    (bits[(size_t)b * words + c / 64] >> (c % 64) & 1)) {
                    ++t;
                    if (per_vertex) {
                        per_vertex[a]++;
                        per_vertex[b]++;
                        per_vertex[c]++;
                    }
                }
            }
        }
    *total = t;
    free(bits);
    return 0;
}

/* ------------------------------------------------------------------ */
/* P-rank simulation                                                   */
/* ------------------------------------------------------------------ */

int or_owner(uint64_t x, const uint64_t* bounds, int P) {
    /* upper_bound over range ends bounds[1..P] (helpers.h:83-101). */
    int lo = 0, hi = P;
    while (lo < hi) {
        int mid = (lo + hi) / 2;
        if (x < bounds[mid + 1])
            hi = mid;
        else
            lo = mid + 1;
    }
    return lo;
}

/* id_outward(rank) key (distributed_graph.h:102-112): (rank(x) != r, id). */
static inline uint64_t key_outward(uint32_t x, int owner_x, int r) {
    return ((uint64_t)(owner_x != r) << 32) | x;
}

typedef struct {
    uint64_t* off; /* n_items + 1 */
    uint64_t* key; /* sorted keys */
} or_csr64;

static void csr64_free(or_csr64* c) {
    free(c->off);
    free(c->key);
}

/* Oriented out-lists of arbitrary vertices (global (degree,id) order, which
 * equals node_ordering::degree with exchanged ghost degrees,
 * distributed_graph.h:113-146), sorted by the key function. */
static void out_list(const uint64_t* off, const uint32_t* heads, uint32_t v, const uint64_t* bounds, int P, int r,
                     int key_kind, uint64_t** dst, uint64_t* len) {
    uint64_t d = off[v + 1] - off[v];
    uint64_t* k = (uint64_t*)malloc((d ? d : 1) * sizeof(uint64_t));
    uint64_t c = 0;
    for (uint64_t e = off[v]; e < off[v + 1]; ++e) {
        uint32_t u = heads[e];
        if (deg_less(off, v, u)) {
            k[c++] = key_kind ? key_outward(u, or_owner(u, bounds, P), r) : (uint64_t)u;
        }
    }
    qsort(k, c, sizeof(uint64_t), cmp_u64);
    *dst = k;
    *len = c;
}

typedef struct {
    uint64_t tri_local, tri_global, wedge, isz_local, isz_global;
} sim_acc;

static inline void emit3(uint64_t* pv, uint32_t a, uint32_t b, uint32_t c, int atomic) {
    add_pv(pv, a, atomic);
    add_pv(pv, b, atomic);
    add_pv(pv, c, atomic);
}

int or_rank_sim(uint32_t n, const uint64_t* offsets, const uint32_t* heads, int P, const uint64_t* bounds, int rank,
                int algorithm, int threads, uint64_t* local_count, uint64_t* global_count, uint64_t* per_vertex,
                or_stats* stats) {
    if (P < 1 || rank < 0 || rank >= P || bounds[0] != 0 || bounds[P] != n) return -1;
    const uint32_t from = (uint32_t)bounds[rank], to = (uint32_t)bounds[rank + 1];
    const uint32_t nl = to - from;
    const int nthr = threads > 1 ? threads : 1;
    const int atomic = nthr > 1;
    const int key_kind = (algorithm == 0) ? 1 : 0; /* cetric-x sorts by id_outward(rank), patric by id */
    if (per_vertex) memset(per_vertex, 0, (size_t)n * sizeof(uint64_t));

    /* Local out-lists N+(v) after preprocessing (cetric.h:61-158). */
    or_csr64 L;
    L.off = (uint64_t*)malloc(((size_t)nl + 1) * sizeof(uint64_t));
    uint64_t** tmp = (uint64_t**)malloc((nl ? nl : 1) * sizeof(uint64_t*));
    uint64_t* tlen = (uint64_t*)malloc((nl ? nl : 1) * sizeof(uint64_t));
#pragma omp parallel for schedule(dynamic, 256) num_threads(nthr)
    for (int64_t i = 0; i < (int64_t)nl; ++i)
        out_list(offsets, heads, from + (uint32_t)i, bounds, P, rank, key_kind, &tmp[i], &tlen[i]);
    L.off[0] = 0;
    for (uint32_t i = 0; i < nl; ++i) L.off[i + 1] = L.off[i] + tlen[i];
    L.key = (uint64_t*)malloc((L.off[nl] ? L.off[nl] : 1) * sizeof(uint64_t));
    for (uint32_t i = 0; i < nl; ++i) {
        memcpy(L.key + L.off[i], tmp[i], tlen[i] * sizeof(uint64_t));
        free(tmp[i]);
    }
    free(tmp);
    free(tlen);
#define KEY_ID(k) ((uint32_t)((k) & 0xffffffffull))
#define IS_LOCAL(x) ((x) >= from && (x) < to)

    sim_acc acc = {0, 0, 0, 0, 0};

    if (algorithm == 0) {
        /* ---- CeTriC "cetric-x" ---- */
        /* remove_in_edges_and_expand_ghosts (distributed_graph.h:1057-1104):
         * ghost_adj(g) = sorted {local u : g in N-(u)}, g remote.  Ghosts are
         * kept in ascending id order (the reference iterates a hash map; the
         * order affects iteration only). */
        uint64_t ng_edges = 0;
        for (uint32_t i = 0; i < nl; ++i) {
            uint32_t u = from + i;
            for (uint64_t e = offsets[u]; e < offsets[u + 1]; ++e) {
                uint32_t g = heads[e];
                if (!IS_LOCAL(g) && deg_less(offsets, g, u)) ng_edges++;
            }
        }
        uint64_t* pairs = (uint64_t*)malloc((ng_edges ? ng_edges : 1) * sizeof(uint64_t));
        uint64_t pc = 0;
        for (uint32_t i = 0; i < nl; ++i) {
            uint32_t u = from + i;
            for (uint64_t e = offsets[u]; e < offsets[u + 1]; ++e) {
                uint32_t g = heads[e];
                if (!IS_LOCAL(g) && deg_less(offsets, g, u)) pairs[pc++] = ((uint64_t)g << 32) | u;
            }
        }
        qsort(pairs, pc, sizeof(uint64_t), cmp_u64);
        /* ghost CSR */
        uint64_t ngh = 0;
        for (uint64_t i = 0; i < pc; ++i)
            if (i == 0 || (pairs[i] >> 32) != (pairs[i - 1] >> 32)) ngh++;
        uint32_t* gid = (uint32_t*)malloc((ngh ? ngh : 1) * sizeof(uint32_t));
        uint64_t* goff = (uint64_t*)malloc((ngh + 1) * sizeof(uint64_t));
        uint64_t* gkey = (uint64_t*)malloc((pc ? pc : 1) * sizeof(uint64_t));
        {
            uint64_t gi = 0;
            for (uint64_t i = 0; i < pc; ++i) {
                if (i == 0 || (pairs[i] >> 32) != (pairs[i - 1] >> 32)) {
                    gid[gi] = (uint32_t)(pairs[i] >> 32);
                    goff[gi] = i;
                    gi++;
                }
                gkey[i] = key_outward((uint32_t)(pairs[i] & 0xffffffffu), rank, rank); /* locals: key = id */
            }
            goff[ngh] = pc;
        }
        free(pairs);
        /* ghost_adj lookup by id (binary search in gid) */
#define GHOST_IDX(x, out)                                 \
    do {                                                  \
        int64_t lo_ = 0, hi_ = (int64_t)ngh;              \
        while (lo_ < hi_) {                               \
            int64_t mid_ = (lo_ + hi_) / 2;               \
            if (gid[mid_] < (x))                          \
                lo_ = mid_ + 1;                           \
            else                                          \
                hi_ = mid_;                               \
        }                                                 \
        (out) = (lo_ < (int64_t)ngh && gid[lo_] == (x)) ? lo_ : -1; \
    } while (0)

        /* run_local_sequential_with_ghosts (cetric_edge_iterator.h:300-388) */
        uint64_t t_loc = 0, wch = 0, isz = 0;
#pragma omp parallel num_threads(nthr) reduction(+ : t_loc, wch, isz)
        {
            uint64_t* buf = NULL;
            size_t cap = 0;
#pragma omp for schedule(dynamic, 128)
            for (int64_t i = 0; i < (int64_t)nl; ++i) {
                uint32_t v = from + (uint32_t)i;
                const uint64_t* nv = L.key + L.off[i];
                uint64_t dv = L.off[i + 1] - L.off[i];
                for (uint64_t j = 0; j < dv; ++j) {
                    uint32_t u = KEY_ID(nv[j]);
                    const uint64_t* nu;
                    uint64_t du;
                    if (!IS_LOCAL(u)) {
                        int64_t gi;
                        GHOST_IDX(u, gi);
                        if (gi < 0) {
                            nu = NULL;
                            du = 0;
                        } else {
                            nu = gkey + goff[gi];
                            du = goff[gi + 1] - goff[gi];
                        }
                    } else {
                        nu = L.key + L.off[u - from];
                        du = L.off[u - from + 1] - L.off[u - from];
                    }
                    wch += du;
                    isz += dv + du;
                    size_t need = dv < du ? dv : du;
                    if (need > cap) {
                        cap = 2 * need;
                        buf = (uint64_t*)realloc(buf, cap * sizeof(uint64_t));
                    }
                    size_t k = merge_u64(nv, dv, nu, du, buf);
                    t_loc += k;
                    for (size_t q = 0; q < k; ++q) emit3(per_vertex, v, u, KEY_ID(buf[q]), atomic);
                }
            }
#pragma omp for schedule(dynamic, 128)
            for (int64_t g = 0; g < (int64_t)ngh; ++g) {
                uint32_t v = gid[g];
                const uint64_t* nv = gkey + goff[g];
                uint64_t dv = goff[g + 1] - goff[g];
                for (uint64_t j = 0; j < dv; ++j) {
                    uint32_t u = KEY_ID(nv[j]);
                    const uint64_t* nu = L.key + L.off[u - from];
                    uint64_t du = L.off[u - from + 1] - L.off[u - from];
                    wch += du;
                    isz += dv + du;
                    size_t need = dv < du ? dv : du;
                    if (need > cap) {
                        cap = 2 * need;
                        buf = (uint64_t*)realloc(buf, cap * sizeof(uint64_t));
                    }
                    size_t k = merge_u64(nv, dv, nu, du, buf);
                    t_loc += k;
                    for (size_t q = 0; q < k; ++q) emit3(per_vertex, v, u, KEY_ID(buf[q]), atomic);
                }
            }
            free(buf);
        }
        acc.tri_local = t_loc;
        acc.wedge += wch;
        acc.isz_local += isz;

        /* Contraction: remove_internal_edges (distributed_graph.h:679-764)
         * keeps N+cut(v) = remote out-neighbours of interface nodes; compact
         * (:978-1055) drops the rest.  N+cut of a local u = remote suffix of
         * its id_outward-sorted list.  Contracted phase: every remote a whose
         * N+cut(a) (computed from a's owner's view) touches this rank sends
         * [a, N+cut(a)] once (last_proc_, cetric_edge_iterator.h:1362-1364;
         * enqueue_for_sending :1632-1661).  For each local u in the message
         * that survived compaction: N+cut(u) cap message -> emit {a,u,w}
         * (process_neighborhood / intersect_from_message :1797-1958). */
        uint64_t t3 = 0, wch3 = 0, isz3 = 0;
#pragma omp parallel num_threads(nthr) reduction(+ : t3, wch3, isz3)
        {
            uint64_t* buf = NULL;
            size_t cap = 0;
#pragma omp for schedule(dynamic, 256)
            for (int64_t aa = 0; aa < (int64_t)n; ++aa) {
                uint32_t a = (uint32_t)aa;
                if (IS_LOCAL(a)) continue;
                int sa = or_owner(a, bounds, P);
                /* does a have a cut out-neighbour owned by this rank? */
                int sends = 0;
                for (uint64_t e = offsets[a]; e < offsets[a + 1]; ++e) {
                    uint32_t x = heads[e];
                    if (IS_LOCAL(x) && deg_less(offsets, a, x)) {
                        sends = 1;
                        break;
                    }
                }
                if (!sends) continue;
                /* message payload: N+cut(a) from rank sa's view, sorted by id */
                uint64_t* msg;
                uint64_t mlen;
                out_list(offsets, heads, a, bounds, P, sa, 0, &msg, &mlen);
                uint64_t w = 0;
                for (uint64_t q = 0; q < mlen; ++q)
                    if (or_owner(msg[q], bounds, P) != sa) msg[w++] = msg[q];
                mlen = w;
                for (uint64_t q = 0; q < mlen; ++q) {
                    uint32_t u = (uint32_t)msg[q];
                    if (!IS_LOCAL(u)) continue;
                    /* N+cut(u): remote suffix of u's list (keys with bit 32 set) */
                    const uint64_t* nu = L.key + L.off[u - from];
                    uint64_t du = L.off[u - from + 1] - L.off[u - from];
                    uint64_t s = 0;
                    while (s < du && (nu[s] >> 32) == 0) s++;
                    if (s == du) continue; /* not interface: removed by compaction */
                    uint64_t dcut = du - s;
                    /* cut keys carry the remote flag; compare by id */
                    if (dcut > cap) {
                        cap = 2 * dcut + mlen;
                        buf = (uint64_t*)realloc(buf, cap * sizeof(uint64_t));
                    }
                    for (uint64_t z = 0; z < dcut; ++z) buf[z] = KEY_ID(nu[s + z]);
                    wch3 += dcut;
                    isz3 += mlen + dcut;
                    /* intersect in place: count matches of buf[0..dcut) with msg */
                    uint64_t i1 = 0, i2 = 0;
                    while (i1 < dcut && i2 < mlen) {
                        if (buf[i1] < msg[i2])
                            ++i1;
                        else if (msg[i2] < buf[i1])
                            ++i2;
                        else {
                            t3++;
                            emit3(per_vertex, a, u, (uint32_t)buf[i1], atomic);
                            ++i1;
                            ++i2;
                        }
                    }
                }
                free(msg);
            }
            free(buf);
        }
        acc.tri_global = t3;
        acc.wedge += wch3;
        acc.isz_global += isz3;
        free(gid);
        free(goff);
        free(gkey);
    } else {
        /* ---- DiTriC "patric" ---- */
        /* run_local_sequential (cetric_edge_iterator.h:229-298), NodeOrdering
         * = id: ghost heads `continue`; lists are id-sorted. */
        uint64_t t_loc = 0, wch = 0, isz = 0;
#pragma omp parallel num_threads(nthr) reduction(+ : t_loc, wch, isz)
        {
            uint64_t* buf = NULL;
            size_t cap = 0;
#pragma omp for schedule(dynamic, 128)
            for (int64_t i = 0; i < (int64_t)nl; ++i) {
                uint32_t v = from + (uint32_t)i;
                const uint64_t* nv = L.key + L.off[i];
                uint64_t dv = L.off[i + 1] - L.off[i];
                for (uint64_t j = 0; j < dv; ++j) {
                    uint32_t u = KEY_ID(nv[j]);
                    if (!IS_LOCAL(u)) continue;
                    const uint64_t* nu = L.key + L.off[u - from];
                    uint64_t du = L.off[u - from + 1] - L.off[u - from];
                    wch += du;
                    isz += dv + du;
                    size_t need = dv < du ? dv : du;
                    if (need > cap) {
                        cap = 2 * need;
                        buf = (uint64_t*)realloc(buf, cap * sizeof(uint64_t));
                    }
                    size_t k = merge_u64(nv, dv, nu, du, buf);
                    t_loc += k;
                    for (size_t q = 0; q < k; ++q) emit3(per_vertex, v, u, KEY_ID(buf[q]), atomic);
                }
            }
            free(buf);
        }
        acc.tri_local = t_loc;
        acc.wedge += wch;
        acc.isz_local += isz;
        /* run_distributed over interface nodes: each remote a with an out-
         * neighbour on this rank sends [a, N+(a)] (all out-neighbours,
         * skip_local_neighborhood = false); receiver intersects N+(u) with
         * the message for each local u in it (intersect_from_message). */
        uint64_t tg = 0, wchg = 0, iszg = 0;
#pragma omp parallel num_threads(nthr) reduction(+ : tg, wchg, iszg)
        {
#pragma omp for schedule(dynamic, 256)
            for (int64_t aa = 0; aa < (int64_t)n; ++aa) {
                uint32_t a = (uint32_t)aa;
                if (IS_LOCAL(a)) continue;
                int sends = 0;
                for (uint64_t e = offsets[a]; e < offsets[a + 1]; ++e) {
                    uint32_t x = heads[e];
                    if (IS_LOCAL(x) && deg_less(offsets, a, x)) {
                        sends = 1;
                        break;
                    }
                }
                if (!sends) continue;
                uint64_t* msg;
                uint64_t mlen;
                out_list(offsets, heads, a, bounds, P, or_owner(a, bounds, P), 0, &msg, &mlen);
                for (uint64_t q = 0; q < mlen; ++q) {
                    uint32_t u = (uint32_t)msg[q];
                    if (!IS_LOCAL(u)) continue;
                    const uint64_t* nu = L.key + L.off[u - from];
                    uint64_t du = L.off[u - from + 1] - L.off[u - from];
                    wchg += du;
                    iszg += mlen + du;
                    uint64_t i1 = 0, i2 = 0;
                    while (i1 < du && i2 < mlen) {
                        if (nu[i1] < msg[i2])
                            ++i1;
                        else if (msg[i2] < nu[i1])
                            ++i2;
                        else {
                            tg++;
                            emit3(per_vertex, a, u, (uint32_t)nu[i1], atomic);
                            ++i1;
                            ++i2;
                        }
                    }
                }
                free(msg);
            }
        }
        acc.tri_global = tg;
        acc.wedge += wchg;
        acc.isz_global += iszg;
    }
    csr64_free(&L);
    *local_count = acc.tri_local;
    *global_count = acc.tri_global;
    if (stats) {
        memset(stats, 0, sizeof(*stats));
        stats->wedge_checks = acc.wedge;
        stats->intersection_size_local = acc.isz_local;
        stats->intersection_size_global = acc.isz_global;
        stats->local_triangles = acc.tri_local;
        stats->type3_triangles = acc.tri_global;
    }
#undef KEY_ID
#undef IS_LOCAL
#undef GHOST_IDX
    return 0;
}

int or_rank_closed_form(uint32_t n, const uint64_t* offsets, const uint32_t* heads, int P, const uint64_t* bounds,
                        int algorithm, uint64_t* local, uint64_t* type3) {
    uint64_t* oo = (uint64_t*)malloc(((size_t)n + 1) * sizeof(uint64_t));
    uint32_t* oh = (uint32_t*)malloc((offsets[n] / 2 + 1) * sizeof(uint32_t));
    if (or_orient(n, offsets, heads, oo, oh) != 0) {
        free(oo);
        free(oh);
        return -1;
    }
    memset(local, 0, (size_t)P * sizeof(uint64_t));
    memset(type3, 0, (size_t)P * sizeof(uint64_t));
    for (uint32_t a = 0; a < n; ++a)
        for (uint64_t i = oo[a]; i < oo[a + 1]; ++i) {
            uint32_t b = oh[i];
            uint64_t p = oo[a], q = oo[b];
            while (p < oo[a + 1] && q < oo[b + 1]) {
                if (oh[p] < oh[q])
                    ++p;
                else if (oh[q] < oh[p])
                    ++q;
                else {
                    uint32_t c = oh[p];
                    int ra = or_owner(a, bounds, P), rb = or_owner(b, bounds, P), rc = or_owner(c, bounds, P);
                    if (algorithm == 0) {
                        if (ra == rb || ra == rc)
                            local[ra]++;
                        else if (rb == rc)
                            local[rb]++;
                        else
                            type3[rb]++;
                    } else {
                        if (ra == rb)
                            local[rb]++;
                        else
                            type3[rb]++;
                    }
                    ++p;
                    ++q;
                }
            }
        }
    free(oo);
    free(oh);
    return 0;
}
