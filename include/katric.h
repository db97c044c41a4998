/*
 * katric_amd -- C ABI of the MI355X (gfx950) triangle-counting core.
 *
 * The reference (KaTric, header-only C++) has no C ABI; its drop-in surface
 * for this path is a set of C++ templates.  Each entry point below names the
 * reference interface it replaces (paths relative to the reference root).
 *
 * Conventions
 *   - vertex ids uint32, CSR offsets uint64 (n+1 entries, offsets[0] = 0),
 *     triangle counts uint64; all counting is exact integer arithmetic.
 *   - input graphs are symmetric (undirected) CSRs without self-loops or
 *     duplicate entries; adjacency lists may be in any order.
 *   - pointers are host pointers unless KT_DEVICE_INPUT / KT_DEVICE_OUTPUT
 *     say otherwise; device pointers must be allocations on the context's
 *     device whose producing work has completed before the call.
 *   - every call is synchronous: outputs are complete when it returns.
 *   - status codes are 0 (KT_OK) or one of the negative KT_E* values; no
 *     C++ exception crosses the ABI.  kt_last_error() gives a message.
 *   - one kt_ctx may be used from several host threads; calls on one
 *     context are serialised by an internal lock.
 */
#ifndef KATRIC_H
#define KATRIC_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define KATRIC_ABI_VERSION 3

/* status codes */
#define KT_OK 0
#define KT_EINVAL (-1)   /* malformed CSR, ranges or arguments */
#define KT_ENOMEM (-2)   /* HBM budget or device allocation exhausted */
#define KT_EHIP (-3)     /* HIP runtime error */
#define KT_ENODEV (-4)   /* no usable GPU / bad device index */
#define KT_ENOTREADY (-5) /* serve: request still pending */
#define KT_ECANCELED (-6) /* serve: request cancelled */
#define KT_ENOTFOUND (-7) /* serve: unknown ticket */
#define KT_ESTATE (-8)    /* call not valid in the current state */
#define KT_ERANGE (-9)    /* output buffer too small (the required size is still returned) */

/* flags */
#define KT_DEVICE_INPUT 0x1u  /* offsets / heads are device pointers */
#define KT_DEVICE_OUTPUT 0x2u /* per-vertex outputs are device pointers */

typedef struct kt_ctx kt_ctx;

/* Counters of cetric::profiling::LocalStatistics for the last call
 * (include/cetric/statistics.h:193-203).  intersection_size_local is the
 * reference's W = sum over oriented edges (v,u) of |N+(v)| + |N+(u)|
 * (counters/cetric_edge_iterator.h:279-281). */
typedef struct {
    uint64_t wedge_checks;
    uint64_t intersection_size_local;
    uint64_t intersection_size_global;
    uint64_t local_triangles;
    uint64_t type3_triangles;
    uint64_t oriented_edges; /* |E+| */
    uint64_t vertices;
    uint64_t input_slots;    /* entries of the symmetric input CSR */
    uint64_t skipped_nodes;  /* tails skipped with pseudo2core (cetric_edge_iterator.h:239-242) */
    /* T-list ids the intersection streamed through its probes.  With
     * kt_params.min_side = 0 this is wedge_checks less the lists of tails
     * with a single out-neighbour (which cannot close a triangle and are not
     * walked); min_side streams fewer.  No reference counterpart. */
    uint64_t streamed_ids;
} kt_stats;

/* Device times of the last call, from HIP events on the context stream. */
typedef struct {
    float total_ms;     /* first device op to last device op */
    float h2d_ms;       /* host -> device copies (0 for device input) */
    float prep_ms;      /* orientation + oriented-CSR construction */
    float count_ms;     /* binning + intersection */
    float intersect_ms; /* the intersection kernel alone */
    float d2h_ms;       /* device -> host copies of the outputs */
} kt_timing;

/* Intersection strategy knobs (bin thresholds and list walk); the package's
 * autotuner (katric_amd.tune) chooses them per graph class.  Counts never
 * depend on them. */
#define KT_STRATEGY_STREAMED 0u  /* the wave walks one T list at a time, 256 ids per step (128 for lists of <= 128) */
#define KT_STRATEGY_GROUPED16 1u /* groups of 16 lanes walk 4 T lists at a time (short lists) */
#define KT_STRATEGY_GROUPED8 2u  /* groups of 8 lanes walk 8 T lists at a time (very short lists) */
#define KT_STRATEGY_GROUPED32 3u /* groups of 32 lanes walk 2 T lists at a time */
#define KT_STRATEGY_GROUPED4 4u  /* groups of 4 lanes walk 16 T lists at a time */
#define KT_STRATEGY_GROUPED2 5u  /* groups of 2 lanes walk 32 T lists at a time */
#define KT_ORDER_DEGREE 0u      /* work queues by degree bucket, largest first */
#define KT_ORDER_VERTEX 1u      /* work queues by vertex id (locality across a batch's graphs) */
typedef struct {
    uint32_t wave_max_deg;  /* |N+(v)| <= this: one wave per tail, LDS hash (<= 256) */
    uint32_t group_max_deg; /* |N+(v)| <= this: one workgroup per tail, LDS hash (<= 2048) */
    uint32_t grid_per_cu;   /* persistent 512-thread workgroups per compute unit (1..4 fill the CU) */
    uint32_t strategy;      /* KT_STRATEGY_* */
    uint32_t order;         /* KT_ORDER_* */
    /* |N+(v)| above this: the tail's set is probed by binary search in global
     * memory instead of an LDS hash (1..8192; 8192 = only sets beyond the
     * LDS).  1 sends every tail to binary search: the device form of
     * Config::intersection_method = binary_search (config.h:129,172). */
    uint32_t bsearch_min_deg;
    /* Config::pseudo2core (config.h:184): tails with |N+(v)| < 2 are skipped
     * and counted in kt_stats.skipped_nodes (cetric_edge_iterator.h:239-242).
     * Counts never change; the work counters do.  Shared-memory path only
     * (kt_count, kt_list_triangles, kt_batch_count, serving): the shard
     * entry points return KT_EINVAL while it is set. */
    uint32_t pseudo2core;
    /* 1 (default): kt_count, kt_batch_count and serving intersect every
     * oriented edge (v, u) from the side whose out-list is shorter -- N+(u)
     * probed against the table of N+(v) as in the reference, or N+(v)
     * against the table of N+(u) -- so the streamed ids drop from
     * sum |N+(u)| (wedge_checks) to sum min(|N+(v)|, |N+(u)|).  Counts,
     * per-vertex counts and the reference's work counters are the same
     * either way.  0: always from v (the reference's walk,
     * cetric_edge_iterator.h:888-1066).  Listing and the shard entry points
     * always walk from v.  No reference counterpart. */
    uint32_t min_side;
} kt_params;

/* Create a context on HIP device `device`.  hbm_budget_bytes bounds the
 * device memory the context may hold (0 = no bound beyond the device).
 * No reference counterpart (the reference allocates on the host). */
int kt_create(kt_ctx** ctx, int device, uint64_t hbm_budget_bytes);
void kt_destroy(kt_ctx* ctx);
const char* kt_strerror(int status);
const char* kt_last_error(const kt_ctx* ctx);
int kt_set_params(kt_ctx* ctx, const kt_params* params);
int kt_get_params(const kt_ctx* ctx, kt_params* params);
int kt_last_stats(const kt_ctx* ctx, kt_stats* stats);
int kt_last_timing(const kt_ctx* ctx, kt_timing* timing);
/* bytes of device memory currently held by the context */
uint64_t kt_device_bytes(const kt_ctx* ctx);

/* Device buffers on the context's device, for callers without a HIP
 * allocator of their own (they count against the HBM budget).  No reference
 * counterpart: the reference keeps every graph in host memory. */
int kt_buffer_alloc(kt_ctx* ctx, uint64_t bytes, void** device_ptr);
int kt_buffer_free(kt_ctx* ctx, void* device_ptr);
int kt_buffer_upload(kt_ctx* ctx, void* device_dst, const void* host_src, uint64_t bytes);
int kt_buffer_download(kt_ctx* ctx, void* host_dst, const void* device_src, uint64_t bytes);
/* Page-locked host memory (faster, asynchronous host <-> device copies). */
int kt_host_alloc(uint64_t bytes, void** host_ptr);
int kt_host_free(void* host_ptr);
/* Waits for all work of the context's device. */
int kt_synchronize(kt_ctx* ctx);

/* Shared-memory triangle count of a symmetric CSR: total and (if
 * per_vertex_out != NULL) per-vertex counts t[x] = number of triangles
 * containing x.  Replaces cetric::run_shmem
 * (include/cetric/counters/shared_memory_cetric.h:105-170), i.e.
 * preprocessing (:71-103: DistributedGraph::orient with
 * node_ordering::degree_outward + sort_neighborhoods,
 * datastructures/distributed/distributed_graph.h:537-593) followed by
 * CetricEdgeIterator::run_local (counters/cetric_edge_iterator.h:173-298,
 * 888-1066) whose emit callback counts triangles. */
int kt_count(kt_ctx* ctx, uint32_t n, const uint64_t* offsets, const uint32_t* heads, uint32_t flags,
             uint64_t* total_out, uint64_t* per_vertex_out);

/* Triangle listing: every triangle of a symmetric CSR once, as the
 * reference's emit callback receives it -- Triangle{v, u, w} with v the tail
 * (lowest in the (degree, id) order), u in N+(v), w in N+(v) and N+(u)
 * (run_local's emit, counters/cetric_edge_iterator.h:173-298, the call at
 * :289; run_shmem's emit, shared_memory_cetric.h:130-147).  triangles
 * receives 3 ids per triangle (host, or device with KT_DEVICE_OUTPUT), in no
 * particular order (the reference emits concurrently from its workers);
 * *total_out the number of triangles.  If it exceeds capacity (in
 * triangles), nothing is written and KT_ERANGE is returned with *total_out
 * set.  Two device passes: a count with per-tail totals, then the listing
 * at the scanned per-tail offsets. */
int kt_list_triangles(kt_ctx* ctx, uint32_t n, const uint64_t* offsets, const uint32_t* heads, uint32_t flags,
                      uint32_t* triangles, uint64_t capacity, uint64_t* total_out);

/* Degree orientation + oriented-CSR construction only: out_offsets[n+1],
 * out_heads[offsets[n]/2] receive, for every v, the ascending ids u with
 * (deg v, v) < (deg u, u).  Replaces the preprocessing half of run_shmem:
 * DistributedGraph::orient (distributed_graph.h:537-566) with
 * node_ordering::degree_outward (:147-163), sort_neighborhoods (:577-593)
 * and the out-only rewrite of remove_in_edges_and_expand_ghosts
 * (:1057-1104) on one rank. */
int kt_orient(kt_ctx* ctx, uint32_t n, const uint64_t* offsets, const uint32_t* heads, uint32_t flags,
              uint64_t* out_offsets, uint32_t* out_heads);

/* Batched intersection of caller-given id lists: for pair i, the strictly
 * increasing lists A_i = a_ids[a_offsets[i] .. a_offsets[i+1]) and
 * B_i = b_ids[b_offsets[i] .. b_offsets[i+1]) (offsets: num_pairs + 1
 * entries starting at 0).  counts[i] receives |A_i cap B_i| and *total_out
 * their sum; if matches != NULL the common ids are written pair after pair,
 * each pair's in ascending order (pair i's from the sum of counts[0..i) on),
 * or KT_ERANGE is returned (counts and *total_out still set) when they exceed
 * capacity (in ids).  A list that is not strictly increasing is KT_EINVAL.
 * Inputs are host arrays unless KT_DEVICE_INPUT; counts / matches are host
 * arrays unless KT_DEVICE_OUTPUT.  Replaces cetric::intersection
 * (include/cetric/counters/intersection.h:40-79, its callback receiving every
 * common element once, ascending as merge_intersection :81-96 emits them), as
 * tests/intersection_test.cpp:35-173 calls it. */
int kt_intersect(kt_ctx* ctx, uint64_t num_pairs, const uint64_t* a_offsets, const uint32_t* a_ids,
                 const uint64_t* b_offsets, const uint32_t* b_ids, uint32_t flags, uint64_t* counts,
                 uint32_t* matches, uint64_t capacity, uint64_t* total_out);

/* kt_count without any allocation: for callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops that forbid it).  No
 * counterpart in the reference.  The device workspace is a fixed array of
 * KT_NOALLOC_SCRATCH_BYTES in the library, one per GPU, held until the
 * process ends; a graph whose workspace exceeds it (about 24 bytes per vertex
 * plus 2.2 bytes per CSR entry for sorted lists, plus 16 bytes per edge for
 * unsorted ones) is KT_ENOMEM.  offsets / heads are device pointers,
 * per_vertex_out a device pointer (nullable), *total_out is written on the
 * host.  Calls on one GPU share the array: this library serialises them
 * within a process; calls that overlap on one GPU through other means race. */
#define KT_NOALLOC_SCRATCH_BYTES (64ull << 20)
int kt_count_noalloc(kt_ctx* ctx, uint32_t n, const uint64_t* offsets, const uint32_t* heads, uint64_t* total_out,
                     uint64_t* per_vertex_out);

/* Batched count of many independent graphs given as one disjoint-union CSR:
 * graph g owns the global ids [vertex_base[g], vertex_base[g+1]) (host array
 * of num_graphs+1), offsets has vertex_base[num_graphs]+1 entries and heads
 * hold global ids.  totals (host, num_graphs) receives each graph's count,
 * per_vertex (nullable) the per-vertex counts over the union.  The device
 * pipeline is recorded into a HIP graph on the first call for a batch shape
 * and replayed on later calls with the same shape and buffers.  No reference
 * counterpart (the reference counts one graph per run, apps/shmetric.cpp:138-142);
 * per graph the result equals kt_count. */
int kt_batch_count(kt_ctx* ctx, uint32_t num_graphs, const uint64_t* vertex_base, const uint64_t* offsets,
                   const uint32_t* heads, uint32_t flags, uint64_t* totals, uint64_t* per_vertex);
/* Number of replays of the context's captured batch pipeline. */
uint64_t kt_batch_replays(const kt_ctx* ctx);

#ifdef __cplusplus
}
#endif
#endif /* KATRIC_H */
