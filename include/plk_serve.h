/*
 * plk_serve.h -- likelihood serving loop of libplk (continuous batching on one GPU).
 *
 * Requests are whole likelihood problems (tree, alignment patterns, model) with one of
 * the PhyloTree jobs IQ-TREE's callers issue: a tree logL (computeLikelihood,
 * tree/phylotree.cpp:1258-1299) or a Newton branch-length solve (optimizeOneBranch
 * tree/phylotree.cpp:2620-2670 / optimizeAllBranches :2724-2801).  Requests are admitted
 * in arrival order whenever a worker slot is free and their HBM footprint
 * (plk_memory_required, after PhyloTree::getMemoryRequired tree/phylotree.cpp:992-1052)
 * fits in the remaining budget; each admitted request gets its own context and HIP
 * stream, so requests that arrive while others run overlap on the device.  Partial
 * buffers stay resident for the whole request (all Newton iterations).  Results do not
 * depend on what else is running: every reduction has a fixed order.
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

/* nworkers concurrent requests at most; hbm_budget bytes of device memory for requests. */
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
int plk_server_destroy(plk_server *s);

#ifdef __cplusplus
}
#endif
#endif
