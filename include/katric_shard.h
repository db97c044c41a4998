/*
 * katric_amd -- one rank's shard of the distributed algorithms (CeTriC
 * "cetric-x" and DiTriC "patric") on one GPU.
 *
 * A rank owns the contiguous vertex range [range_bounds[rank],
 * range_bounds[rank+1]).  It holds the full adjacency of its owned vertices,
 * the global degrees of its ghosts (non-owned neighbours) and the messages it
 * receives in the distributed phase.  kt_count_shard runs the rank's local
 * phase and its contracted / distributed phase on the device and returns the
 * rank's counts, exactly as the reference's rank would emit them.
 */
#ifndef KATRIC_SHARD_H
#define KATRIC_SHARD_H

#include "katric.h"

#ifdef __cplusplus
extern "C" {
#endif

#define KT_SHARD_CETRIC 0u /* configs/cetric.toml: algorithm = "cetric-x" */
#define KT_SHARD_DITRIC 1u /* configs/ditric.toml: algorithm = "patric" */

typedef struct {
    uint32_t algorithm;          /* KT_SHARD_CETRIC or KT_SHARD_DITRIC */
    uint32_t rank;
    uint32_t num_ranks;
    uint32_t flags;              /* KT_DEVICE_INPUT: every array below except range_bounds is a device pointer */
    const uint64_t* range_bounds; /* host, num_ranks + 1 contiguous vertex-id bounds */
    /* full adjacency (global ids) of the owned vertices */
    const uint64_t* offsets;     /* n_local + 1 */
    const uint32_t* heads;       /* offsets[n_local] */
    /* ghosts: every non-owned id occurring in heads, ascending, with its global degree */
    uint32_t n_ghost;
    const uint32_t* ghost_ids;
    const uint64_t* ghost_degree;
    /* messages received in the distributed phase: message i is
     * [msg_src[i], msg_heads[msg_offsets[i] .. msg_offsets[i+1])], ids ascending;
     * cetric-x: N+cut(a), the out-neighbours of a not owned by a's owner
     * (enqueue_for_sending after contraction, counters/cetric_edge_iterator.h:1632-1661);
     * patric: N+(a), all out-neighbours of a. */
    uint32_t n_msg;
    const uint32_t* msg_src;
    const uint64_t* msg_offsets; /* n_msg + 1 */
    const uint32_t* msg_heads;
} kt_shard_in;

typedef struct {
    uint64_t local_triangles; /* emits of the local phase */
    uint64_t type3_triangles; /* emits of the contracted (cetric-x) / distributed (patric) phase */
    uint64_t* per_vertex;     /* nullable; n_local + n_ghost entries: owned vertices, then ghosts in
                                 ghost_ids order; triangles emitted on this rank containing the vertex
                                 (device pointer with KT_DEVICE_OUTPUT in flags) */
} kt_shard_out;

/* One rank's counts.  Replaces, for that rank:
 *   cetric-x: run_cetric_new (counters/cetric.h:665-1002) -- preprocessing
 *     with exchanged ghost degrees (:744-762; node_ordering::degree,
 *     datastructures/distributed/distributed_graph.h:113-146),
 *     remove_in_edges_and_expand_ghosts (:1057-1169),
 *     run_local_sequential_with_ghosts (cetric_edge_iterator.h:300-388),
 *     contraction remove_internal_edges + compact (distributed_graph.h:679-764,
 *     978-1055) and the receiving side of the contracted phase,
 *     process_neighborhood / intersect_from_message (cetric_edge_iterator.h:1797-1958);
 *   patric: run_patric (counters/cetric.h:160-316) -- run_local_sequential
 *     (cetric_edge_iterator.h:229-298, ghost heads skipped) and the receiving
 *     side of run_distributed.
 * local_triangles / type3_triangles are the rank's emit counts of the two
 * phases (their sum over ranks is the global count reduced by MPI_Reduce at
 * cetric.h:293/979). */
int kt_count_shard(kt_ctx* ctx, const kt_shard_in* in, kt_shard_out* out);

/* One rank's counts from a full symmetric CSR (n vertices, offsets[n+1],
 * heads[offsets[n]], ascending or not; device-resident with KT_DEVICE_INPUT):
 * the rank's kt_shard_in -- local slice, ghost ids and global degrees, the
 * senders of the distributed phase and their message lists -- is built on the
 * device from the resident graph (what the reference's MPI exchange delivers:
 * the 1D partitioned read, DegreeCommunicator,
 * datastructures/distributed/graph_communicator.h:48-266, and the message
 * queue of the contracted / distributed phase,
 * counters/cetric_edge_iterator.h:1328-1379, 1632-1661; ghosts as in
 * find_ghosts, distributed_graph.h:1287-1298), then counted exactly like
 * kt_count_shard.  out->per_vertex (nullable) has n entries indexed by global
 * id: the triangles emitted on this rank containing each vertex (zero for
 * vertices the rank does not see); with KT_SHARD_ACCUMULATE they are added to
 * its contents instead, so a sweep over all ranks sums to kt_count's
 * per-vertex counts.  range_bounds is a host array as in kt_shard_in.  No
 * single reference counterpart: it fuses the exchange of the inputs of one
 * rank with run_cetric_new / run_patric on that rank. */
#define KT_SHARD_ACCUMULATE 0x10u
int kt_count_shard_graph(kt_ctx* ctx, uint32_t n, const uint64_t* offsets, const uint32_t* heads, uint32_t flags,
                         uint32_t algorithm, uint32_t num_ranks, const uint64_t* range_bounds, uint32_t rank,
                         kt_shard_out* out);

/* Host-side producer of kt_shard_in arrays for one rank from a full
 * symmetric CSR (it computes what MPI delivers to that rank in the reference:
 * the local CSR slice, the ghost degrees and the received messages). */
typedef struct {
    uint32_t rank, num_ranks, algorithm;
    uint32_t n_local;
    uint64_t* offsets;
    uint32_t* heads;
    uint32_t n_ghost;
    uint32_t* ghost_ids;
    uint64_t* ghost_degree;
    uint32_t n_msg;
    uint32_t* msg_src;
    uint64_t* msg_offsets;
    uint32_t* msg_heads;
} kt_shard_host;

int kt_shard_make(uint32_t n, const uint64_t* offsets, const uint32_t* heads, uint32_t num_ranks,
                  const uint64_t* range_bounds, uint32_t rank, uint32_t algorithm, kt_shard_host* out);
void kt_shard_host_free(kt_shard_host* s);

#ifdef __cplusplus
}
#endif
#endif
