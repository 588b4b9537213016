/*
 * plk_serve.h -- likelihood serving loop of libplk (continuous batching on one GPU).
 *
 * Requests are whole likelihood problems (tree, alignment patterns, model) with one of
 * the PhyloTree jobs IQ-TREE's callers issue: a tree logL (computeLikelihood,
 * tree/phylotree.cpp:1258-1299) or a Newton branch-length solve (optimizeOneBranch
 * tree/phylotree.cpp:2620-2670 / optimizeAllBranches :2724-2801).  Requests are admitted
 * in arrival order whenever one of the nworkers request slots is free and their HBM
 * footprint (plk_memory_required, after PhyloTree::getMemoryRequired
 * tree/phylotree.cpp:992-1052) fits in the remaining budget.  Continuous batching: at
 * every step the traversals + logL of all requests in flight (DNA or 20-state, ncat <= 4)
 * run as ONE persistent launch over their (request, tile) work items; Newton phases run on
 * the request's own stream on three worker threads, so requests that arrive while others
 * are on the GPU join the next shared launch and their work overlaps on the device.
 * Trees and device buffers are pooled per (states, categories, precision, scaling) and
 * reshaped per request; partial buffers stay resident for the whole request (all Newton
 * iterations).  Results do not depend on what else is in flight: every reduction has a
 * fixed order per request.
 */
#ifndef PLK_SERVE_H
#define PLK_SERVE_H

#include <stdint.h>

#include "plk.h"

#ifdef __cplusplus
extern "C" {
#endif

#define PLK_REQ_LOGL 0        /* full traversal + logL at the root branch */
#define PLK_REQ_OPT_BRANCH 1  /* logL, then Newton on branch (opt_a, opt_b), then logL */
#define PLK_REQ_OPT_ALL 2     /* optimizeAllBranches(iterations, tol) */

typedef struct plk_request {
    /* problem (host arrays must stay valid until the request completes) */
    int nstates, ncat, ntaxa, nptn, precision, scaling, state_unknown;
    const double *eval, *evec, *inv_evec, *rates, *props, *tip_lh;
    const uint8_t *states;
    const double *freq, *invar;
    const int32_t *nei;
    const double *len;
    /* job */
    int kind, opt_a, opt_b, iterations;
    double tol;
    /* outputs */
    double logl_before, logl_after, opt_len;
    int status;           /* PLK_OK or error code */
    double t_submit, t_admit, t_done;  /* seconds on the server clock */
    int64_t footprint;    /* bytes of HBM the request held */
} plk_request;

typedef struct plk_server plk_server;

/* nworkers requests in flight at most; hbm_budget bytes of device memory for requests
 * (their footprints; pooled trees keep their buffers until the server is destroyed). */
int plk_server_create(plk_server **out, int nworkers, int64_t hbm_budget);
/* Queue a request (non-blocking).  t_submit is stamped now. */
int plk_server_submit(plk_server *s, plk_request *req);
/* Block until every submitted request has completed. */
int plk_server_drain(plk_server *s);
/* Seconds on the server clock. */
double plk_server_now(plk_server *s);
/* stats[0..7]: completed, failed, p50 latency s, p99 latency s, mean latency s,
 * makespan s, peak HBM bytes in use, peak concurrent requests. */
int plk_server_stats(plk_server *s, double *stats);
/* out[0..2]: shared traversal launches, request traversals run in them, pooled trees. */
int plk_server_counters(plk_server *s, int64_t *out);
/* Waits for every submitted request, then stops the threads and frees the pool. */
int plk_server_destroy(plk_server *s);

#ifdef __cplusplus
}
#endif
#endif
