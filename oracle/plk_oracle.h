/*
 * plk_oracle.h -- CPU FP64 restatement of IQ-TREE's Felsenstein likelihood kernels.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP engine
 * (tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py may load
 * it); the product path never links or calls it.
 *
 * It restates, in plain C with OpenMP, the reversible non-site-model kernels of
 * tree/phylokernelnew.h:
 *   computePartialInfo            tree/phylokernelnew.h:1013-1063
 *   computeBounds                 tree/phylokernelnew.h:1116-1147
 *   computeTraversalInfo          tree/phylotree.cpp:6031-6139, tree/phylokernelnew.h:1154-1285
 *   computePartialLikelihoodSIMD  tree/phylokernelnew.h:1636-2036 (tip-tip, tip-internal,
 *                                 internal-internal; NORM and SAFE scaling)
 *   computeLikelihoodBranchSIMD   tree/phylokernelnew.h:2658-3184
 *   computeLikelihoodBufferSIMD   tree/phylokernelnew.h:2052-2235
 *   computeLikelihoodDervSIMD     tree/phylokernelnew.h:2237-2650
 *   computeLikelihoodFromBufferSIMD tree/phylokernelnew.h:3284-3450
 *   minimizeNewton                utils/optimization.cpp:422-499
 *   optimizeOneBranch             tree/phylotree.cpp:2620-2670
 *   optimizeAllBranches           tree/phylotree.cpp:2699-2801, tree/mtree.cpp:3040-3082
 *   tip vectors                   tree/phylotreesse.cpp:354-375, model/modelsubst.cpp:183-208,
 *                                 model/modeldna.cpp:462-476, model/modelprotein.cpp:1240-1256
 *   RELL / UFBoot replicate update tree/iqtree.cpp:3852-3931
 *   NNI5 move evaluation          getBestNNIForBran tree/phylotree.cpp:4176-4429
 *   discrete-Gamma rates          RateGamma::computeRatesMean model/rategamma.cpp:98-171
 *   Brent one-dimensional search  Optimization::brent_opt / minimizeOneDimen
 *                                 utils/optimization.cpp:174-372
 *   Gamma shape optimisation      RateGamma::optimizeParameters model/rategamma.cpp:178-232
 *   model-parameter loop          ModelFactory::optimizeParameters model/modelfactory.cpp:1263-1315,
 *                                 1558-1720
 *   (BFGS for the exchangeabilities, ModelMarkov::optimizeParameters, is restated in
 *   oracle/optim_oracle.py and driven from tests/support.py)
 *
 * Memory layout follows the reference's pattern-interleaved SIMD layout with
 * ORC_V lanes: element (ptn, c, x) of a partial sits at double offset
 * (floor(ptn/V)*block + c*n + x)*V + ptn%V  (tree/phylokernelnew.h:1918-1920).
 *
 * Pinning (what checks each function; see tests/test_oracle.py):
 *   partials/logL    brute-force enumeration of internal states on 4-5 taxon trees,
 *                    two-taxon closed form, pulley principle, pattern compression
 *   scaling          deep caterpillar trees vs long-double brute force; NORM == SAFE
 *   derivatives      Richardson finite differences of logL(t)
 *   from-theta logL  equals branch logL at the same length
 *   Newton           converges to the finite-difference optimum (df ~ 0, ddf < 0)
 *   tip vectors      closed form Uinv * indicator(state set)
 *   RELL             counts = pattern frequencies reproduce the tree logL (to the float
 *                    rounding of pattern_lh); the update rule on a hand-worked sequence
 *   NNI              the move's score equals the logL of the swapped tree, built from
 *                    scratch, after the same five branch optimisations; the tree is restored
 *   Gamma rates      scipy's exact category means (golden fixtures) to the algorithms'
 *                    own accuracy; Yang (1994)'s published alpha = 0.5 table; mean 1
 *   Brent            analytic functions against their known minima and scipy's bounded
 *                    Brent; the bound and worse-than-start rules on constructed cases
 *   shape / loop     the optimum agrees with a fine grid of oracle logL(alpha) and with
 *                    scipy's bounded minimiser; the loop never lowers the logL and ends at
 *                    a stationary point (branch df ~ 0 and alpha at the grid optimum)
 */
#ifndef PLK_ORACLE_H
#define PLK_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_tree orc_tree;

/* sequence types for orc_tip_lh */
enum { ORC_SEQ_DNA = 0, ORC_SEQ_PROTEIN = 1, ORC_SEQ_CODON = 2 };

int orc_lanes(void);
/* Number of directions (tips excluded) the oracle has recomputed since create. */
long long orc_updates(const orc_tree *t);

orc_tree *orc_create(int nstates, int ncat, int ntaxa, int nptn, int safe, int nthreads);
void orc_destroy(orc_tree *t);

void orc_set_model(orc_tree *t, const double *eval, const double *evec,
                   const double *inv_evec, const double *rates, const double *props);
/* replace the eigen system only (rates, proportions and the Gamma shape stay) */
void orc_set_eigen(orc_tree *t, const double *eval, const double *evec, const double *inv_evec);
/* tip_lh: (state_unknown+1) x nstates, already multiplied by the inverse eigenvectors */
void orc_set_tip_lh(orc_tree *t, int state_unknown, const double *tip_lh);
/* states: ntaxa x nptn (row per taxon); freq, invar: nptn */
void orc_set_patterns(orc_tree *t, const uint8_t *states, const double *freq,
                      const double *invar);
/* nei, len: (2*ntaxa-2) x 3; leaves are nodes 0..ntaxa-1 with one neighbour (slot 0),
 * unused slots are -1. */
void orc_set_tree(orc_tree *t, const int *nei, const double *len);
void orc_set_length(orc_tree *t, int a, int b, double len);
double orc_get_length(const orc_tree *t, int a, int b);
void orc_clear_all(orc_tree *t);

double orc_branch_loglik(orc_tree *t, int dad, int node, double *ptn_lh);
/* returns 0, or 1 when the derivative is not finite (NORM mode error of the reference) */
int orc_branch_derivs(orc_tree *t, int dad, int node, int reuse_theta, double *df, double *ddf);
double orc_loglik_from_theta(orc_tree *t, int dad, int node);
/* partial of node's subtree seen from dad, canonical order [ptn][c][x]; scale is
 * [ptn] (NORM) or [ptn][c] (SAFE).  Computes pending partials first. */
int orc_get_partial(orc_tree *t, int dad, int node, double *out, uint16_t *scale);

double orc_optimize_one_branch(orc_tree *t, int node1, int node2, double min_bl,
                               double max_bl, int max_nr_step);
double orc_optimize_all_branches(orc_tree *t, int iterations, double tol, double min_bl,
                                 double max_bl, int max_nr_step);

void orc_tip_lh(int nstates, int seq_type, const double *inv_evec, double *out);
/* NNI5 score of one move on internal branch (node1, node2) (getBestNNIForBran); lens[5]
 * receives the five optimised lengths (central first).  The tree is restored. */
double orc_nni_eval(orc_tree *t, int node1, int node2, int move, double min_bl, double max_bl, int max_nr,
                    double *lens);
/* UFBoot RELL scoring and replicate update (IQTree::saveCurrentTree); uniform == NULL
 * scores only.  counts: nsamples x nptn. */
void orc_rell_save_tree(int nsamples, int nptn, const float *counts, const double *ptn_lh, double cur_logl,
                        double eps, const double *uniform, double *rell, double *boot_logl, int32_t *boot_counts,
                        double *boot_orig, int32_t *improved);
int orc_state_unknown(int nstates, int seq_type);

/* ---- model-parameter loop ---- */
/* discrete-Gamma mean rates of shape alpha (RateGamma::computeRatesMean) */
void orc_gamma_rates(double alpha, int ncat, double *rates);
/* median-of-category rates (RateGamma::computeRates with cut_median) */
void orc_gamma_rates_median(double alpha, int ncat, double *rates);
/* rates of later shapes at the category medians instead of the means */
void orc_set_gamma_median(orc_tree *t, int median);
typedef double (*orc_fn1)(void *ctx, double x);
/* Optimization::minimizeOneDimen + brent_opt on f(ctx, x); returns the abscissa, fx the
 * minimum, f2x the curvature estimate */
double orc_minimize_one_dimen(orc_fn1 f, void *ctx, double xmin, double xguess, double xmax, double tol, double *fx,
                              double *f2x);
/* set the Gamma shape (replaces the rates, all partials dirty); orc_set_model clears it */
void orc_set_gamma_shape(orc_tree *t, double alpha);
double orc_get_gamma_shape(const orc_tree *t);
/* RateGamma::optimizeParameters: returns the logL, the shape is left in the tree */
double orc_optimize_gamma(orc_tree *t, double gradient_epsilon, double min_gamma, double max_gamma);
/* ModelFactory::optimizeParameters with branch lengths and the Gamma shape free */
double orc_optimize_parameters(orc_tree *t, int num_iterations, double logl_epsilon, double gradient_epsilon,
                               double min_gamma, double max_gamma, double min_bl, double max_bl, int max_nr);

#ifdef __cplusplus
}
#endif
#endif
