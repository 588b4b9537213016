/*
 * plk_batch.h -- batched unlinked partitions (libplk).
 *
 * IQ-TREE evaluates a partitioned analysis by looping over per-partition PhyloTrees
 * (PhyloSuperTree::computeLikelihood, tree/phylosupertree.cpp:709-734, OpenMP over
 * partitions ordered by cost, :662-707).  A batch runs that loop as ONE P(t) launch plus
 * ONE persistent traversal launch over all partitions' 64-pattern tiles, replayed as a
 * HIP graph while the launch shape is unchanged.  Each partition is a plk_ctx whose ops
 * were queued with plk_update_partials; results per partition are bit-identical to
 * evaluating that context alone.
 */
#ifndef PLK_BATCH_H
#define PLK_BATCH_H

#include <stdint.h>

#include "plk.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct plk_batch plk_batch;

/* All contexts share one precision form; state counts 4 or 20, ncat <= 4. */
int plk_batch_create(plk_batch **out, plk_ctx *const *parts, int nparts);
int plk_batch_destroy(plk_batch *b);
/* Replace the batch's partitions (same rules as plk_batch_create), keeping its device and
 * staging buffers where they are large enough: the serving loop re-forms its batch from
 * the requests in flight at every step (PhyloSuperTree's partition loop over a changing
 * set, tree/phylosupertree.cpp:709-734).  Drops any captured graph. */
int plk_batch_rebind(plk_batch *b, plk_ctx *const *parts, int nparts);
/* grid: persistent workgroups (0 = one per work item); use_graph: replay a captured graph. */
int plk_batch_set_launch(plk_batch *b, int grid, int use_graph);
/* Per partition i: run its queued ops and the branch logL for node refs a[i], b[i] at
 * length t[i] (computeLikelihoodBranch of every partition).  logl (nullable) receives the
 * per-partition values, total their sum. */
int plk_batch_loglik(plk_batch *b, const int *a, const int *bref, const double *t, double *logl, double *total);
/* Measure the persistent grid candidates on the currently queued traversal (kept queued)
 * and keep the fastest. */
int plk_batch_autotune(plk_batch *b, const int *a, const int *bref, const double *t, int reps, int *best_grid,
                       double *best_ms);
/* Wall time (HIP events around each batched evaluation, ms), evaluations, updates. */
int plk_batch_timing(plk_batch *b, double *total_ms, int64_t *launches, int64_t *updates);
int plk_batch_items(plk_batch *b);

#ifdef __cplusplus
}
#endif
#endif
