/*
 * plk_rell.h -- RELL evaluation of IQ-TREE's ultrafast bootstrap on the device (libplk).
 *
 * UFBoot keeps, per bootstrap replicate s, a vector of pattern counts (boot_samples[s],
 * float, tree/iqtree.cpp:357-420, drawn once by createBootstrapAlignment).  Every tree the
 * search visits is scored on every replicate by RELL: rell[s] = sum_p pattern_lh[p] *
 * boot_samples[s][p] with pattern_lh the per-pattern logL rounded to float
 * (IQTree::saveCurrentTree, tree/iqtree.cpp:3852-3935; computePatternLikelihood
 * tree/phylotree.cpp:1483-1516), and the replicate's best tree is replaced under the
 * epsilon / random tie-break rule of tree/iqtree.cpp:3917-3931.
 *
 * Here the replicate counts stay resident in HBM and the per-pattern logL never leaves
 * the device: plk_rell_save_tree reads the per-pattern logL that the last
 * plk_branch_loglik (or plk_loglik_from_theta) of a context left on that context's
 * device, and applies the scoring and the update rule to every replicate in one launch.
 * Products of the float logL and the float count are exact in FP64 and are summed in
 * FP64 in a fixed order (the reference sums in float SIMD lanes; the FP64 sum is the
 * exact value those lanes approximate).  The tie-break's uniform deviate per replicate
 * is supplied by the caller instead of being drawn from the reference's random stream.
 */
#ifndef PLK_RELL_H
#define PLK_RELL_H

#include <stdint.h>

#include "plk.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct plk_rell plk_rell;

/* counts: nsamples x nptn pattern counts (boot_samples).  The replicate state starts as
 * in tree/iqtree.cpp:386-390: boot_logl = boot_orig_logl = -DBL_MAX, boot_counts = 0.
 * Device memory: nsamples x round_up(nptn, 64) floats. */
int plk_rell_create(plk_rell **out, int nsamples, int nptn, const float *counts);
int plk_rell_destroy(plk_rell *r);

/* rell[s] for the per-pattern logL currently on ctx's device (nptn must match). */
int plk_rell_eval(plk_rell *r, plk_ctx *ctx, double *rell);

/* saveCurrentTree's replicate loop (tree/iqtree.cpp:3909-3931) for a tree of score
 * cur_logl: score every replicate, then with better = rell > boot_logl + eps, or, for
 * |rell - boot_logl| within eps, uniform[s] <= 1 / (boot_counts + 1): on better,
 * boot_counts becomes count + 1 (tie) or 1, boot_logl = max(boot_logl, rell),
 * boot_orig_logl = cur_logl.  improved[s] (nullable) = 1 where the replicate's best tree
 * is now this tree (the caller stores the tree string), *nimproved (nullable) their count. */
int plk_rell_save_tree(plk_rell *r, plk_ctx *ctx, double cur_logl, double eps, const double *uniform,
                       int32_t *improved, int *nimproved);

/* plk_rell_eval without any device allocation, for callers that cannot allocate (code
 * under HIP graph capture, tight serving loops, hosts that forbid hipMalloc).  No
 * counterpart in the reference: it is plk_rell_eval on caller-owned device memory.
 * d_counts: nsamples rows of pattern counts in device memory, row s at d_counts + s *
 * count_stride (16-byte aligned, count_stride a multiple of 4 and >= nptn rounded up to 64;
 * entries nptn .. that rounded size of each row are read, weighted by 0, and must be
 * finite); d_rell: nsamples doubles of device memory.  The float copy of the per-pattern
 * logL is staged in a library-owned static device array of 2^20 patterns (4 MiB per
 * device, held until the process ends): nptn above 2^20 - 63 is rejected.  Every call on
 * one device shares that array, so calls that overlap on one device race.  Runs on ctx's
 * stream and returns when the scores are written. */
int plk_rell_eval_noalloc(plk_ctx *ctx, int nsamples, int nptn, const float *d_counts, int64_t count_stride,
                          double *d_rell);

/* Read the replicate state (any pointer nullable). */
int plk_rell_get(plk_rell *r, double *boot_logl, int32_t *boot_counts, double *boot_orig_logl);

#ifdef __cplusplus
}
#endif
#endif
