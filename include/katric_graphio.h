/*
 * katric_amd graph input layer (host only): seeded generators and file
 * formats feeding the counting entry points of katric.h.
 *
 * Reference counterpart: graph-io (graphio::read_graph / gen_local_graph,
 * called from apps/shmetric.cpp:119 and apps/katric.cpp:233-264); the library
 * is not vendored in the reference tree, so the generators are this
 * project's own seeded restatements of G(n,m) and R-MAT, and the binary format
 * is the one of examples/metis-sample.first_out / .head (uint64 LE).
 *
 * All returned graphs are canonical symmetric CSRs (uint32 ids, uint64
 * offsets, no self-loops or duplicates, ascending lists) except the readers,
 * which keep the file's adjacency order.  Buffers are malloc'ed by the library
 * and released with kt_host_graph_free.
 */
#ifndef KATRIC_GRAPHIO_H
#define KATRIC_GRAPHIO_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define KT_GIO_OK 0
#define KT_GIO_EINVAL (-1)
#define KT_GIO_ENOMEM (-2)
#define KT_GIO_EIO (-10)
#define KT_GIO_EFORMAT (-11)

typedef struct {
    uint32_t n;
    uint64_t nnz;       /* offsets[n] */
    uint64_t* offsets;  /* n+1 */
    uint32_t* heads;    /* nnz */
} kt_host_graph;

void kt_host_graph_free(kt_host_graph* g);

/* G(n, m): `samples` endpoint pairs drawn uniformly from [0,n) by a
 * SplitMix64 stream keyed by seed; self-loops and duplicates removed,
 * symmetrised. */
int kt_gen_gnm(uint32_t n, uint64_t samples, uint64_t seed, kt_host_graph* out);

/* R-MAT: `samples` pairs of `scale`-bit ids by recursive quadrant choice with
 * probabilities (a, b, c, 1-a-b-c); pairs with an id >= n_limit are rejected
 * (n_limit = 0: 2^scale); ids optionally relabelled by a seeded Fisher-Yates
 * permutation; self-loops and duplicates removed, symmetrised. */
int kt_gen_rmat(uint32_t scale, uint64_t samples, double a, double b, double c, uint64_t seed, uint32_t n_limit,
                int permute, kt_host_graph* out);

int kt_read_metis(const char* path, kt_host_graph* out);
int kt_write_metis(const char* path, uint32_t n, const uint64_t* offsets, const uint32_t* heads);
int kt_read_graphio_binary(const char* first_out_path, const char* head_path, kt_host_graph* out);
int kt_write_graphio_binary(const char* first_out_path, const char* head_path, uint32_t n, const uint64_t* offsets,
                            const uint32_t* heads);

/* A batch of independent graphs as one disjoint-union CSR: graph g owns the
 * global ids [vertex_base[g], vertex_base[g+1]); heads hold global ids. */
typedef struct {
    uint32_t num_graphs;
    uint64_t n_total;
    uint64_t nnz_total;
    uint64_t* vertex_base; /* num_graphs + 1 */
    uint64_t* offsets;     /* n_total + 1 */
    uint32_t* heads;       /* nnz_total */
    uint8_t* kind;         /* per graph: 0 = G(n,m), 1 = R-MAT */
} kt_host_batch;

void kt_host_batch_free(kt_host_batch* b);

/* Mixed batch: graph g draws from its own seeded stream n = floor(exp(U[ln
 * n_min, ln n_max])), an average degree d ~ U[deg_min, deg_max] and its kind
 * (R-MAT (a, b, c) with probability rmat_fraction, else G(n,m)); both kinds
 * take n*d/2 endpoint pairs, R-MAT at scale ceil(log2 n) rejecting ids >= n,
 * ids permuted.  Every graph is canonical (as above). */
int kt_gen_batch(uint32_t num_graphs, uint64_t seed, uint32_t n_min, uint32_t n_max, double deg_min, double deg_max,
                 double rmat_fraction, double a, double b, double c, kt_host_batch* out);

/* FNV-1a 64 over n, offsets[0..n] and heads[0..offsets[n]). */
uint64_t kt_csr_checksum(uint32_t n, const uint64_t* offsets, const uint32_t* heads);

#ifdef __cplusplus
}
#endif
#endif
