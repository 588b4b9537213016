/*
 * plk_tree.h -- PhyloTree-level host API of libplk (C ABI over the C++ PhyloTreeGPU).
 *
 * PhyloTreeGPU mirrors the part of IQ-TREE's PhyloTree that drives the likelihood
 * kernels: the unrooted binary tree, directional partial bookkeeping with one slot per
 * internal node (LM_PER_NODE, tree/phylotree.cpp:6011-6025), the post-order traversal
 * planner (computeTraversalInfo, tree/phylotree.cpp:6031-6139), invalidation
 * (clearReversePartialLh, tree/phylonode.cpp:22-33) and the Newton branch optimiser
 * (optimizeOneBranch tree/phylotree.cpp:2620-2670, minimizeNewton
 * utils/optimization.cpp:422-499, optimizeAllBranches tree/phylotree.cpp:2699-2801).
 * Every kernel call goes through the plk.h C ABI.
 *
 * Nodes 0..ntaxa-1 are leaves (one neighbour, in slot 0 of nei), internal nodes are
 * ntaxa..2*ntaxa-3; nei/len are (2*ntaxa-2) x 3 with -1 for unused slots.
 */
#ifndef PLK_TREE_H
#define PLK_TREE_H

#include <stdint.h>

#include "plk.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct plk_tree plk_tree;

int plk_tree_create(plk_tree **out, int nstates, int ncat, int ntaxa, int nptn, int precision, int scaling);
/* all_directions != 0: one slot per direction (3(T-2) slots) instead of LM_PER_NODE's one
 * per internal node, so every branch's two partials coexist (plk_tree_all_branch_derivs).
 * all_directions == 2 also reserves the scratch slots of plk_tree_nni_evaluate (6 per
 * internal branch). */
int plk_tree_create_ex(plk_tree **out, int nstates, int ncat, int ntaxa, int nptn, int precision, int scaling,
                       int all_directions);
int plk_tree_destroy(plk_tree *t);
/* Reuse a tree (and its context's device buffers) for another problem of the same state
 * count, categories, precision and scaling: ntaxa taxa and nptn patterns.  Buffers are
 * kept where they are large enough and grown otherwise; model, patterns and topology
 * must be set again.  Not for trees created with all_directions. */
int plk_tree_reshape(plk_tree *t, int ntaxa, int nptn);
plk_ctx *plk_tree_context(plk_tree *t);
int plk_tree_set_model(plk_tree *t, const double *eval, const double *evec, const double *inv_evec,
                       const double *rates, const double *props, int state_unknown, const double *tip_lh);
int plk_tree_set_patterns(plk_tree *t, const uint8_t *states, const double *freq, const double *invar);
int plk_tree_set_topology(plk_tree *t, const int32_t *nei, const double *len);
int plk_tree_set_length(plk_tree *t, int a, int b, double len);
int plk_tree_get_length(plk_tree *t, int a, int b, double *len);
int plk_tree_clear_all(plk_tree *t);

/* computeLikelihoodBranch(dad_branch = dad->node, dad) (tree/phylokernelnew.h:2658) */
int plk_tree_loglik(plk_tree *t, int dad, int node, double *logl, double *ptn_lh);
/* computeLikelihood(): branch of the first leaf (tree/phylotree.cpp:1258-1299) */
int plk_tree_compute_likelihood(plk_tree *t, double *logl);
/* computeLikelihoodDerv at the branch's current length */
int plk_tree_derivs(plk_tree *t, int dad, int node, int reuse_theta, double *df, double *ddf);
/* computeLikelihoodFromBuffer at the branch's current length */
int plk_tree_loglik_from_theta(plk_tree *t, int dad, int node, double *logl);
int plk_tree_get_partial(plk_tree *t, int dad, int node, double *partial, uint16_t *scale);
int plk_tree_optimize_one_branch(plk_tree *t, int node1, int node2, double min_bl, double max_bl,
                                 int max_nr_step, double *optx);
/* Newton solves of optimize_one/all_branches: device_resident != 0 (default) uses
 * plk_newton_branch; 0 runs minimizeNewton on the host with one plk_branch_derivs call per
 * iteration.  Both give the same lengths. */
int plk_tree_set_newton(plk_tree *t, int device_resident);
/* NNI evaluation of getBestNNIForBran (tree/phylotree.cpp:4176-4429) with NNI5 for the
 * internal branches (node1[i], node2[i]), both moves of every branch at once (tree created
 * with all_directions == 2).  Move m of a branch swaps node1's first other neighbour with
 * node2's m-th other neighbour; the move's five branches are optimised in the reference's
 * order with max_nr_step Newton steps (NNI_MAX_NR_STEP).  score[2i + m] = the move's logL,
 * lens[5(2i + m) + 0..4] = its optimised lengths (central, node1's two, node2's two, as
 * newLen[0..4]).  The tree is left unchanged. */
int plk_tree_nni_evaluate(plk_tree *t, int nbr, const int32_t *node1, const int32_t *node2, double min_bl,
                          double max_bl, int max_nr_step, double *score, double *lens);
int plk_tree_optimize_all_branches(plk_tree *t, int iterations, double tol, double min_bl, double max_bl,
                                   int max_nr_step, double *logl);
/* Invalidate every partial and run one full traversal + branch logL (the benchmark step). */
int plk_tree_full_traversal(plk_tree *t, double *logl);
/* Plan the traversal of branch (dad, node) and queue its ops on the tree's context without
 * evaluating it; returns the node references and length for plk_batch_loglik.  With
 * clear != 0 every partial is invalidated first (a full traversal). */
int plk_tree_queue_branch(plk_tree *t, int dad, int node, int clear, int *ref_a, int *ref_b, double *len);
/* Measure every tile shape (plk_set_tile_shape) on full traversals of this tree and keep
 * the fastest; returns it and its mean kernel time. */
int plk_tree_autotune(plk_tree *t, int reps, int *best_wpc, double *best_ms);
/* df/ddf of all 2T-3 branches at their current lengths (branches listed as (a, b), a < b,
 * in node order); one traversal launch + one derivative launch in all-directions mode. */
int plk_tree_all_branch_derivs(plk_tree *t, int32_t *branch_a, int32_t *branch_b, double *df, double *ddf);
/* Counters: [0] partial updates (site x node), [1] derivative evaluations, [2] traversal ops */
int plk_tree_counters(plk_tree *t, int64_t *out3);

/* ---- model-parameter loop ------------------------------------------------------------
 * The rate model is discrete Gamma with ncat mean-rate categories (median-rate with
 * plk_tree_set_gamma_median); the substitution model
 * (eigen system, frequencies) stays as plk_tree_set_model gave it.  The tree keeps a copy
 * of that model and re-uploads it with the new category rates whenever the shape changes
 * (all partials become dirty).  logL values are computeLikelihood()'s (leaf 0's branch). */

/* RateGamma::setGammaShape + computeRates (model/rategamma.cpp:173-176, 98-171): rates of
 * the plk_gamma_rates categories of shape alpha replace the model's rates (props are
 * kept).  PLK_ERR_STATE before plk_tree_set_model; plk_tree_set_model itself clears the
 * shape (rates are then the caller's own). */
int plk_tree_set_gamma_shape(plk_tree *t, double alpha);
/* current shape, 0 when none is set */
int plk_tree_get_gamma_shape(plk_tree *t, double *alpha);
/* median != 0: later shapes (plk_tree_set_gamma_shape, the searches) take their rates at
 * the category medians (plk_gamma_rates_median; RateGamma's cut_median,
 * model/rategamma.cpp:111-126) instead of the means */
int plk_tree_set_gamma_median(plk_tree *t, int median);
/* Reversible substitution model from exchangeabilities (n(n-1)/2, upper triangle row by
 * row, e.g. AC AG AT CG CT GT for DNA) and fixed frequencies pi: Q (plk_rate_matrix), its
 * eigen system (plk_eigen) and the tip vectors of seq_type (plk_tip_lh) replace the
 * model's; rates and proportions stay.  Needs plk_tree_set_model first (which clears
 * it).  The last exchangeability is the reference one (held at its value by the search). */
int plk_tree_set_gtr(plk_tree *t, int seq_type, const double *exch, const double *pi);
int plk_tree_get_gtr(plk_tree *t, double *exch);
/* ModelMarkov::optimizeParameters (model/modelmarkov.cpp:1166-1243) with fixed
 * frequencies: BFGS with forward-difference gradients (Optimization::minimizeMultiDimen /
 * dfpmin / lnsrch / derivativeFunk, utils/optimization.cpp:639-939) over the first
 * n(n-1)/2 - 1 exchangeabilities in [MIN_RATE, MAX_RATE] = [1e-4, 100], gradient tolerance
 * max(gradient_epsilon, TOL_RATE = 1e-4) (model/modelmarkov.h:30-32); every evaluation
 * re-decomposes Q on the host and runs a full traversal on the device.  Leaves the model
 * at the result; logl receives its logL. */
int plk_tree_optimize_model(plk_tree *t, double gradient_epsilon, double *logl);
/* RateGamma::optimizeParameters(gradient_epsilon, min_gamma, max_gamma)
 * (model/rategamma.cpp:215-232): Brent's method behind Optimization::minimizeOneDimen
 * (utils/optimization.cpp:174-372) over the shape, starting at the current one, with
 * tolerance max(gradient_epsilon, TOL_GAMMA_SHAPE = 0.001) (utils/tools.h:605-607 also
 * gives the reference's bounds MIN_GAMMA_SHAPE 0.02, MAX_GAMMA_SHAPE 1000).  Each
 * evaluation is a full traversal on the device.  Leaves the model at the optimum; alpha
 * (nullable) and logl receive it and its logL.  PLK_ERR_STATE without a shape. */
int plk_tree_optimize_gamma(plk_tree *t, double gradient_epsilon, double min_gamma, double max_gamma, double *alpha,
                            double *logl);
/* ModelFactory::optimizeParameters(BRLEN_OPTIMIZE, ..., logl_epsilon, gradient_epsilon)
 * (model/modelfactory.cpp:1558-1720) for the exchangeabilities (when set with
 * plk_tree_set_gtr) and a Gamma rate model (when a shape is set): for i = 2 ..
 * num_param_iterations-1, optimizeAllBranches(min(i, 3), logl_epsilon), then the
 * exchangeabilities (plk_tree_optimize_model), then the shape (plk_tree_optimize_gamma);
 * stop when an iteration's
 * logL is not above the previous one + logl_epsilon, then optimizeAllBranches(100,
 * logl_epsilon).  The reference's defaults: num_param_iterations 100, max_nr_step 100. */
int plk_tree_optimize_parameters(plk_tree *t, int num_param_iterations, double logl_epsilon, double gradient_epsilon,
                                 double min_gamma, double max_gamma, double min_bl, double max_bl, int max_nr_step,
                                 double *alpha, double *logl);

#ifdef __cplusplus
}
#endif
#endif
