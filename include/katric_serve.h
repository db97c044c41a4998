/*
 * katric_amd -- serving loop: continuous batching of triangle-count requests.
 *
 * kt_serve_start launches a worker thread on the context's device.  Requests
 * (symmetric CSRs in host memory) are admitted FIFO into device batches while
 * a batch fits its share of the HBM budget (half of the budget per batch
 * slot; two slots alternate, so one batch uploads while the other counts).
 * Requests that arrive while batches run join the next batch as soon as a slot
 * frees up.  Every result equals the one-graph count of kt_count.
 *
 * No reference counterpart: the reference counts one graph per run
 * (apps/shmetric.cpp:132-146).
 */
#ifndef KATRIC_SERVE_H
#define KATRIC_SERVE_H

#include "katric.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    uint64_t hbm_budget_bytes;  /* device bytes for both batch slots (0: the context budget, else half the free memory) */
    uint32_t max_batch_graphs;  /* 0: 4096 */
    uint32_t large_request_edges; /* requests with more undirected edges run alone, at most one such batch at a
                                     time, and smaller requests behind them may go ahead (0: max_batch_edges / 4) */
    uint64_t max_batch_edges;   /* undirected edges per batch (0: 2^25) */
} kt_serve_config;

typedef struct {
    uint64_t ticket;
    int32_t status;        /* KT_OK, KT_EINVAL (malformed request), KT_ENOMEM (exceeds the budget), KT_ECANCELED */
    uint32_t batch_graphs; /* requests in the batch that counted this one */
    uint64_t total;        /* triangle count */
    uint64_t batch_id;
    double queue_ms;       /* submit -> batch launch */
    double latency_ms;     /* submit -> result published */
} kt_serve_result;

typedef struct {
    uint64_t submitted;
    uint64_t batches;
    uint64_t graphs;
    uint64_t edges;
    uint64_t isolation_reruns; /* batches re-run one request at a time after a validation failure */
    uint64_t pending;
} kt_serve_stats;

int kt_serve_start(kt_ctx* ctx, const kt_serve_config* config);
/* Stops after every submitted request is answered. */
int kt_serve_stop(kt_ctx* ctx);
/* offsets / heads (and per_vertex_out, n entries, nullable) must stay valid
 * until the request's result has been collected. */
int kt_serve_submit(kt_ctx* ctx, uint32_t n, const uint64_t* offsets, const uint32_t* heads,
                    uint64_t* per_vertex_out, uint64_t* ticket);
/* KT_ENOTREADY while pending; collecting a result releases the ticket. */
int kt_serve_poll(kt_ctx* ctx, uint64_t ticket, kt_serve_result* result);
/* timeout_ms < 0 waits without bound. */
int kt_serve_wait(kt_ctx* ctx, uint64_t ticket, double timeout_ms, kt_serve_result* result);
int kt_serve_cancel(kt_ctx* ctx, uint64_t ticket);
int kt_serve_get_stats(kt_ctx* ctx, kt_serve_stats* stats);

#ifdef __cplusplus
}
#endif
#endif
