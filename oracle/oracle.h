/*
 * CPU oracle for the per-rank triangle-counting core.
 *
 * TEST INFRASTRUCTURE ONLY.  This library restates the reference
 * algorithm (KaTric: shmetric / CeTriC / DiTriC local + contracted phase)
 * in plain C so that the HIP implementation can be checked against it.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it, and only as the checker or as the timed CPU baseline.  The
 * product library (katric_amd) never links, loads or calls it.
 *
 * Each function cites the reference file:line it restates (paths relative
 * to the reference repository root).  Parity pins are listed per function
 * in the header comment of oracle.c.
 *
 * Conventions: vertex ids are uint32, CSR offsets uint64 (n+1 entries),
 * per-vertex triangle counts uint64.  All counting is exact integer work.
 */
#ifndef KT_ORACLE_H
#define KT_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Intersection methods, cetric::IntersectionMethod (include/cetric/config.h:129). */
enum { OR_MERGE = 0, OR_BINARY_SEARCH = 1, OR_BINARY = 2, OR_HYBRID = 3 };

/* Counters of cetric::profiling::LocalStatistics that the hot path touches
 * (include/cetric/statistics.h:193-203). */
typedef struct {
    uint64_t wedge_checks;
    uint64_t intersection_size_local;
    uint64_t intersection_size_global;
    uint64_t local_triangles;
    uint64_t type3_triangles;
    uint64_t skipped_nodes;
} or_stats;

/* floor(log2(x)), cetric::log2 (include/cetric/util.h:136-150); 0 for x == 0. */
uint64_t or_log2(uint64_t x);

/* Set intersections of two strictly increasing lists; matches are written to
 * `out` (capacity >= min(na, nb)) in emission order; returns the count.
 * include/cetric/counters/intersection.h:81-96, 98-117, 129-205 (iterative). */
size_t or_merge_intersection(const uint32_t* a, size_t na, const uint32_t* b, size_t nb, uint32_t* out);
size_t or_binary_search_intersection(const uint32_t* a, size_t na, const uint32_t* b, size_t nb, uint32_t* out);
size_t or_binary_intersection(const uint32_t* a, size_t na, const uint32_t* b, size_t nb, uint32_t* out,
                              size_t cutoff);
/* Dispatcher with the hybrid rule, intersection.h:40-79. */
size_t or_intersection(const uint32_t* a, size_t na, const uint32_t* b, size_t nb, uint32_t* out, int method,
                       size_t binary_cutoff, double hybrid_cutoff_scale);

/* Degree orientation + per-part id sort of a symmetric CSR, as shmetric's
 * preprocessing (shared_memory_cetric.h:71-103) with orient
 * (distributed_graph.h:537-566) and sort_neighborhoods (:577-593).
 * Writes the oriented (out-only) CSR: out_off[n+1], out_heads[offsets[n]/2].
 * Returns 0, or -1 if the kept edge count is not offsets[n]/2. */
int or_orient(uint32_t n, const uint64_t* offsets, const uint32_t* heads, uint64_t* out_off, uint32_t* out_heads);

/* shmetric: run_shmem (shared_memory_cetric.h:105-170) = preprocessing +
 * CetricEdgeIterator::run_local with (id_outward(0), degree_outward).
 * threads <= 1 -> run_local_sequential (cetric_edge_iterator.h:229-298);
 * threads > 1  -> node-parallel run_local_parallel (:888-1066), OpenMP.
 * per_vertex (nullable, n entries): t[x] += 1 for x in {v,u,w} of every
 * emitted Triangle{v,u,w}.  stats nullable. */
int or_run_shmem(uint32_t n, const uint64_t* offsets, const uint32_t* heads, int method, int threads,
                 uint64_t* total, uint64_t* per_vertex, or_stats* stats);
/* The same with Config::pseudo2core (config.h:184): tails with |N+(v)| < 2
 * are skipped and counted in stats->skipped_nodes
 * (cetric_edge_iterator.h:239-242). */
int or_run_shmem_ex(uint32_t n, const uint64_t* offsets, const uint32_t* heads, int method, int threads,
                    int pseudo2core, uint64_t* total, uint64_t* per_vertex, or_stats* stats);

/* Timing helper: one-thread preprocessing of the whole graph, then the local
 * phase over the tails v % stride == 0 only (a sample of the workload);
 * returns the sample's triangles and oriented edges and both phases' seconds. */
int or_run_shmem_sampled(uint32_t n, const uint64_t* offsets, const uint32_t* heads, uint32_t stride,
                         uint64_t* sample_triangles, uint64_t* sample_edges, double* prep_s, double* local_s);

/* Brute force over adjacency bitsets (n <= 4096): every triple a<b<c with all
 * three edges present.  Independent of the orientation machinery. */
int or_brute_force(uint32_t n, const uint64_t* offsets, const uint32_t* heads, uint64_t* total, uint64_t* per_vertex);

/* Owner of vertex x under contiguous ranges bounds[0..P] (bounds[0]=0,
 * bounds[P]=n): upper_bound over the range ends, get_PE_from_node_ranges
 * (datastructures/distributed/helpers.h:83-101). */
int or_owner(uint64_t x, const uint64_t* bounds, int P);

/* In-process P-rank simulation of one rank's share of the distributed
 * algorithms on a symmetric CSR partitioned into contiguous ranges.
 *   algorithm 0 = CeTriC "cetric-x" preset (configs/cetric.toml):
 *     run_cetric_new (counters/cetric.h:665-1002): preprocessing with global
 *     degrees (:744-762), remove_in_edges_and_expand_ghosts
 *     (distributed_graph.h:1057-1104), run_local_sequential_with_ghosts
 *     (cetric_edge_iterator.h:300-388), contraction remove_internal_edges +
 *     compact (distributed_graph.h:679-764, 978-1055), and the contracted
 *     phase enqueue_for_sending / process_neighborhood /
 *     intersect_from_message (cetric_edge_iterator.h:1328-1379, 1632-1661,
 *     1797-1958).
 *   algorithm 1 = DiTriC "patric" preset (configs/ditric.toml): run_patric
 *     (counters/cetric.h:160-316) with run_local_sequential (`continue` on
 *     ghost heads) and run_distributed over interface nodes.
 * Outputs for rank r: local_count (local phase emits), global_count
 * (contracted / distributed phase emits), per_vertex (nullable, n entries,
 * indexed by global id): triangles emitted on rank r containing x.
 * threads > 1 parallelises the loops over tails (same emits). */
int or_rank_sim(uint32_t n, const uint64_t* offsets, const uint32_t* heads, int P, const uint64_t* bounds, int rank,
                int algorithm, int threads, uint64_t* local_count, uint64_t* global_count, uint64_t* per_vertex,
                or_stats* stats);

/* Closed-form attribution of every triangle to ranks, by enumeration over the
 * oriented graph (independent of the simulation above):
 *   cetric-x: local[r] = #{T: >= 2 vertices owned by r};
 *             type3[r] = #{T: owners pairwise distinct, owner(middle) = r};
 *   patric:   local[r] = #{T: owner(lowest) = owner(middle) = r};
 *             type3[r] = #{T: owner(middle) = r != owner(lowest)};
 * where lowest/middle/highest are taken in the (degree, id) order. */
int or_rank_closed_form(uint32_t n, const uint64_t* offsets, const uint32_t* heads, int P, const uint64_t* bounds,
                        int algorithm, uint64_t* local, uint64_t* type3);

#ifdef __cplusplus
}
#endif
#endif
