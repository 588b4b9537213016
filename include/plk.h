/*
 * plk.h -- C ABI of the MI355X Felsenstein likelihood engine (libplk.so).
 *
 * The engine replaces the four PhyloTree kernel function pointers that IQ-TREE's
 * setLikelihoodKernel installs (tree/phylotreesse.cpp:91-199) and calls only through
 * the wrappers at tree/phylotreesse.cpp:212-236:
 *
 *   computePartialLikelihoodPointer     tree/phylotree.h:902-904   -> plk_update_partials
 *   computeLikelihoodBranchPointer      tree/phylotree.h:948-951   -> plk_branch_loglik
 *   computeLikelihoodDervPointer        tree/phylotree.h:1343-1346 -> plk_branch_derivs
 *   computeLikelihoodFromBufferPointer  tree/phylotree.h:1002-1004 -> plk_loglik_from_theta
 *
 * plus the buffers PhyloTree owns for them (initializeAllPartialLh,
 * tree/phylotree.cpp:905-960 and 1085-1213): one context holds the partial slots,
 * scale counters (scale_num), theta buffer, pattern weights and model arrays of
 * one alignment on one GPU, resident in HBM.
 *
 * Conventions.
 *   - Every function returns PLK_OK (0) or a negative error code; plk_last_error()
 *     returns a thread-local message for the last failure.  No function aborts.
 *   - A "node reference" is a slot id (>= 0) holding a directional partial, or a tip,
 *     encoded as -(taxon+1).
 *   - Partials use IQ-TREE's eigen-space representation in the FP64 form
 *     (tree/phylokernelnew.h:1989-1994) and are laid out on the device exactly like
 *     the reference's SIMD layout with 64 lanes: element (ptn, c, x) at
 *     (floor(ptn/64)*block + c*n + x)*64 + ptn%64, block = ncat*n.
 *   - scale counts follow scale_num (tree/phylonode.h:17, uint16): one per pattern in
 *     NORM mode, one per (pattern, category) in SAFE mode
 *     (tree/phylokernelnew.h:1924-1936, mode rule tree/phylotreesse.cpp:95-96).
 *   - Contexts are not re-entrant (like PhyloTree); different contexts may be used
 *     from different host threads concurrently, each on its own HIP stream.
 */
#ifndef PLK_H
#define PLK_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PLK_OK 0
#define PLK_ERR_ARG (-1)      /* invalid argument */
#define PLK_ERR_HIP (-2)      /* HIP runtime failure */
#define PLK_ERR_STATE (-3)    /* call out of order (e.g. no model / no theta) */
#define PLK_ERR_NUMERIC (-4)  /* non-finite derivative in NORM mode: the reference's
                                 "use -safe" outError (tree/phylokernelnew.h:2594-2596) */
#define PLK_ERR_NOMEM (-5)    /* device memory exhausted */

/* precision form */
#define PLK_FP64 0            /* reference semantics, eigen-space partials */
#define PLK_FP32 1            /* FP32 partials, per-pattern power-of-two exponents,
                                 FP64 eigen system / exponent sums / reductions */
/* underflow scaling mode (tree/phylotree.h:61-62 SAFE_LH / NORM_LH) */
#define PLK_SCALE_NORM 0
#define PLK_SCALE_SAFE 1

typedef struct plk_ctx plk_ctx;

/* One directional partial update: out_slot <- children.  One element of IQ-TREE's
 * traversal_info (TraversalInfo, tree/phylotree.h:320-331), planned on the host by
 * computeTraversalInfo (tree/phylotree.cpp:6031-6139). */
typedef struct plk_op {
    int32_t out_slot;
    int32_t child[2];  /* slot id or -(taxon+1) */
    int32_t reserved;  /* must be 0 */
    double len[2];     /* branch length from the output node to each child */
} plk_op;

const char *plk_last_error(void);
int plk_version(void);
/* Number of visible GPUs (0 when none); never fails. */
int plk_device_count(void);
/* Compute units of the current GPU (0 when none).  The 20-state FP32 traversal runs the
 * last tiles of a launch whose tile count is just above a multiple of this number as
 * single 16-pattern-block units; no counterpart in the reference (a device query). */
int plk_device_units(void);

/* PhyloTree::initializeAllPartialLh (tree/phylotree.cpp:905-960).  nslots directional
 * partial buffers of nptn x ncat x nstates reals each. */
int plk_create(plk_ctx **out, int nstates, int ncat, int ntaxa, int nptn, int nslots,
               int precision, int scaling);
int plk_destroy(plk_ctx *ctx);
/* Reuse a context for ntaxa taxa, nptn patterns and nslots slots (same state count,
 * categories, precision, scaling): device buffers are kept where large enough and grown
 * (stream-ordered) otherwise; patterns must be set again, the model stays.  The
 * re-initialisation of a PhyloTree's buffers for a new alignment (initializeAllPartialLh,
 * tree/phylotree.cpp:905-960) without giving the memory back. */
int plk_reshape(plk_ctx *ctx, int ntaxa, int nptn, int nslots);
/* Bytes of device memory a context of this shape holds (PhyloTree::getMemoryRequired,
 * tree/phylotree.cpp:992-1052). */
int64_t plk_memory_required(int nstates, int ncat, int ntaxa, int nptn, int nslots,
                            int precision, int scaling);

/* Pattern-compressed alignment: states[taxon*nptn + ptn] (uint8, STATE_UNKNOWN
 * alignment/alignment.cpp:1116-1128), ptn_freq and ptn_invar
 * (tree/phylotreesse.cpp:537-612).  invar may be NULL (all zero). */
int plk_set_patterns(plk_ctx *ctx, const uint8_t *states, const double *ptn_freq,
                     const double *ptn_invar);

/* Model arrays (ModelSubst::getEigenvalues/getEigenvectors/getInverseEigenvectors,
 * model/modelsubst.h:363-375; RateHeterogeneity::getRate/getProp model/rategamma.h:106,120)
 * and the tip table tip_partial_lh[(state_unknown+1) x nstates]
 * (tree/phylotreesse.cpp:354-375).  All FP64, row-major. */
int plk_set_model(plk_ctx *ctx, const double *eval, const double *evec, const double *inv_evec,
                  const double *rates, const double *props, int state_unknown,
                  const double *tip_lh);

/* computePartialLikelihood over one traversal (tree/phylokernelnew.h:1293-2043): ops in
 * post-order.  The ops are queued and run fused with the next branch call, or at
 * plk_flush. */
int plk_update_partials(plk_ctx *ctx, const plk_op *ops, int nops);
int plk_flush(plk_ctx *ctx);

/* computeLikelihoodBranch (tree/phylokernelnew.h:2658-3184) for the branch joining the
 * subtrees held by node references a and b (at most one may be a tip), length t.
 * ptn_lh (nullable) receives per-pattern log-likelihoods (_pattern_lh, save_log_value). */
int plk_branch_loglik(plk_ctx *ctx, int a, int b, double t, double *logl, double *ptn_lh);

/* computeLikelihoodDerv (tree/phylokernelnew.h:2237-2650): first and second derivative of
 * logL with respect to t.  reuse_theta != 0 reuses the device-resident theta buffer of
 * the previous call on the same branch (PhyloTree::theta_computed). */
int plk_branch_derivs(plk_ctx *ctx, int a, int b, double t, int reuse_theta, double *df,
                      double *ddf);

/* computeLikelihoodDerv for nbr branches in one launch (every branch's gradient_vector /
 * hessian_diagonal entry, tree/phylokernelnew.h:2567-2568): a[i], b[i] node references,
 * t[i] lengths.  The theta buffer is not kept. */
int plk_branch_derivs_multi(plk_ctx *ctx, int nbr, const int *a, const int *b, const double *t, double *df,
                            double *ddf);

/* computeLikelihoodFromBuffer (tree/phylokernelnew.h:3284-3450): logL at length t from the
 * theta buffer of the last plk_branch_derivs call. */
int plk_loglik_from_theta(plk_ctx *ctx, double t, double *logl);

/* Optimization::minimizeNewton (utils/optimization.cpp:422-499) on the length of branch
 * (a, b), with computeFuncDerv (tree/phylotree.cpp:2555-2562, f = -dlogL/dt) as the function,
 * run device-resident: evaluation 0 is a full traversal building theta at
 * clamp(xguess, x1, x2); every later evaluation reuses theta, and the bracket/step logic
 * runs on the device after each one, so iterations are queued launches with no host
 * round trip.  The arithmetic, stopping rules and error messages are those of the
 * reference routine; *out receives the minimiser, *evals (nullable) the number of
 * derivative evaluations.  theta stays valid for plk_loglik_from_theta afterwards. */
int plk_newton_branch(plk_ctx *ctx, int a, int b, double x1, double xguess, double x2, double xacc,
                      int max_nr, double *out, int *evals);

/* nbr independent Newton solves at once (minimizeNewton per branch, as plk_newton_branch,
 * with bounds x1, x2, tolerance xacc and max_nr shared): every launch evaluates the
 * derivatives of all unfinished branches, over (branch, tile) work items, from the two
 * partials (theta is not stored), and the step logic runs on the device.  out[i] is the
 * minimiser, logl[i] (nullable) the logL at out[i] (computeLikelihoodFromBuffer's value
 * there), status[i] (nullable) PLK_OK or PLK_ERR_NUMERIC per solve (with status NULL a
 * failed solve fails the call).  max_nr = -1 solves nothing: logl[i] = the logL at
 * clamp(xguess[i], x1, x2) (optimizeOneBranch's divergence check needs it). */
int plk_newton_multi(plk_ctx *ctx, int nbr, const int *a, const int *b, double x1, const double *xguess,
                     double x2, double xacc, int max_nr, double *out, double *logl, int *status);

/* Read back one slot: partial in canonical order [ptn][c][x] as FP64 (the FP32 form
 * returns its own representation), and its scale counts (NORM: nptn uint16; SAFE:
 * nptn*ncat uint16 [ptn][c]).  Either pointer may be NULL. */
int plk_get_partial(plk_ctx *ctx, int slot, double *partial, uint16_t *scale);

/* FP32 form only: the per-pattern power-of-two exponents of one slot (true partial =
 * stored value * 2^exponent), nptn int32. */
int plk_get_exponents(plk_ctx *ctx, int slot, int32_t *exponents);

/* FP32 form, 20/61-state models: tile shape of the matrix-core kernel.  0 = one
 * wavefront per 16-pattern block over all categories (ncat <= 4; no cross-wave exchange);
 * 1, 2 or 4 = wavefronts per category of a 64-pattern tile (more waves raise memory-level
 * parallelism; fewer reuse each matrix fragment over more patterns).  Results do not
 * depend on it. */
int plk_set_tile_shape(plk_ctx *ctx, int waves_per_category);
/* FP64 contexts with 20 or 61 states and at most four categories: form 1 runs the
 * partial-likelihood ops on the matrix cores (FP64 16x16x4 tiles, one workgroup per 16
 * patterns); form 0 the vector-ALU kernels; -1 (default) picks by state count, the faster
 * one measured (61 states: matrix cores, 20: vector ALU).  Same results to rounding (the
 * sums run in a different order); scale counts follow the same rule. */
/* Side outputs of the last plk_branch_loglik (FP64 form), as computeLikelihoodBranch
 * leaves them: lh_cat[p * ncat + c] = the category's likelihood w_c * sum_i val * node * dad
 * in the pattern's scaled frame, SAFE categories above the minimum count scaled by 2^-256
 * or zeroed (_pattern_lh_cat, tree/phylokernelnew.h:3023, 3063-3088); nptn x ncat doubles. */
int plk_get_pattern_lh_cat(plk_ctx *ctx, double *lh_cat);
/* The save_log_value = false form of the last plk_branch_loglik (FP64 form): lh[p] =
 * |sum_c lh_cat| + ptn_invar and scaling[p] = the pattern's log scaling factor, so that
 * logL_p = log(lh[p]) + scaling[p] (_pattern_lh / _pattern_scaling,
 * tree/phylokernelnew.h:3099-3107); nptn doubles each. */
int plk_get_pattern_scaling(plk_ctx *ctx, double *lh, double *scaling);
/* L'_p / L_p of every pattern at the last single-branch derivative evaluation
 * (plk_branch_derivs, plk_newton_branch): the row of G_matrix (tree/phylokernelnew.h:2518)
 * for that branch; nptn doubles. */
int plk_get_pattern_dlogl(plk_ctx *ctx, double *dlogl);
int plk_set_f64_form(plk_ctx *ctx, int form);

/* 4-state models: 1..4 = level-parallel kernel: the ops are sorted into dependency
 * levels and a 64-pattern tile runs that many independent ops of a level at once
 * (min(ncat, 4) wavefronts each), keeping partials consumed within the launch in LDS and
 * forming P(t) itself (no P(t) launch); 0 = one wavefront per category walking the ops
 * in order after a P(t) launch; -1 (default) = 4 for alignments of at most 320 tiles of
 * 64 patterns, else 0.  Results do not depend on it (partials and scale counts are
 * bit-identical in both forms). */
int plk_set_level_waves(plk_ctx *ctx, int slots);

/* Kernel timing with HIP events recorded on the context's stream around every launch
 * of the partial-traversal kernel.  kind: 0 traversal, 1 P(t), 2 derivative. */
int plk_set_timing(plk_ctx *ctx, int enable);
int plk_get_timing(plk_ctx *ctx, int kind, double *total_ms, int64_t *launches);

/* Measured device-to-device copy bandwidth ((read + written bytes)/s, best of reps) of a
 * bytes-sized hipMemcpy: the attainable HBM rate, and a copy of known size for
 * calibrating rocprof FETCH_SIZE/WRITE_SIZE. */
int plk_probe_copy(int64_t bytes, int reps, double *gbps);

/* Number of site-node partial updates (ops x patterns) executed since create. */
int64_t plk_updates(plk_ctx *ctx);
int plk_synchronize(plk_ctx *ctx);

#ifdef __cplusplus
}
#endif
#endif
