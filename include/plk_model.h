/*
 * plk_model.h -- host FP64 model arrays and seeded synthetic alignments for libplk.
 *
 * These are the inputs IQ-TREE's kernels read from its model classes:
 *   rate matrix      ModelMarkov::decomposeRateMatrix, Q part   model/modelmarkov.cpp:1450-1514
 *   eigen system     symmetrised solve, U = Pi^-1/2 V, U^-1 = V^T Pi^1/2
 *                                                             model/modelmarkov.cpp:1519-1562
 *   discrete Gamma   RateGamma::computeRatesMean              model/rategamma.cpp:98-171, 305-470
 *   codon Q          ModelCodon (GY-style, one kappa)          model/modelcodon.cpp:366-387,
 *                                                             574-651, 857-906
 *   tip vectors      computeTipLikelihood x Uinv              tree/phylotreesse.cpp:354-375,
 *                                                             model/modelsubst.cpp:183-208
 *   random tree      MExtTree::generateYuleHarding            tree/mexttree.cpp:432-490
 *   patterns         Alignment pattern compression            alignment/alignment.cpp:1511-1560
 * All functions return PLK_OK or a negative PLK_ERR_* code (plk.h).
 */
#ifndef PLK_MODEL_H
#define PLK_MODEL_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PLK_SEQ_DNA 0
#define PLK_SEQ_PROTEIN 1
#define PLK_SEQ_CODON 2

/* STATE_UNKNOWN of a sequence type (alignment/alignment.cpp:1116-1128) */
int plk_state_unknown(int nstates, int seq_type);

/* Q = R diag(pi), zero row sums, scaled so that -sum_i pi_i Q_ii = 1.  exch holds the
 * n(n-1)/2 exchangeabilities of the upper triangle in row-major order. */
int plk_rate_matrix(int n, const double *exch, const double *pi, double *Q);

/* Reversible eigensystem of Q (row-major n x n): Q = evec diag(eval) inv_evec,
 * eigenvalues ascending. */
int plk_eigen(int n, const double *Q, const double *pi, double *eval, double *evec,
              double *inv_evec);

/* Mean-of-category discrete Gamma rates (Yang 1994), rescaled to mean 1
 * (RateGamma::computeRatesMean, model/rategamma.cpp:158-171). */
int plk_gamma_rates(double alpha, int ncat, double *rates);
/* Median-of-category rates (the cut_median branch of RateGamma::computeRates,
 * model/rategamma.cpp:111-126): the Gamma(alpha, 1/alpha) quantile at (2k+1)/(2 ncat) for
 * category k, rescaled to mean 1. */
int plk_gamma_rates_median(double alpha, int ncat, double *rates);

/* tip_partial_lh[(su+1) x n] = Uinv * tip likelihood of each state */
int plk_tip_lh(int n, int seq_type, const double *inv_evec, double *tip_lh);

/* ptn_invar for a +I model (replaces PhyloTree::computePtnInvar, tree/phylotreesse.cpp:549-612,
 * with the constant-state code of Alignment::computeConst, alignment/alignment.cpp:1321-1420):
 * ptn_invar[p] = p_invar * (sum of pi over the states every taxon of pattern p allows),
 * 0 for patterns with no common state.  states is ntaxa x nptn (plk_set_patterns' layout).
 * The matching rate model has props (1 - p_invar)/ncat and Gamma rates / (1 - p_invar)
 * (model/rategammainvar.h:69, model/rategamma.cpp:129-140). */
int plk_ptn_invar(int n, int seq_type, int ntaxa, int nptn, const uint8_t *states, const double *pi,
                  double p_invar, double *ptn_invar);

/* Exchangeabilities of the 61-state GY codon model of the standard code for (kappa,
 * omega); sense codons in AAA..TTT order. */
int plk_codon_exch(double kappa, double omega, double *exch);

/* Seeded synthetic problem: Yule-Harding tree, i.i.d. Exp(mean_len) branch lengths,
 * random parameters (DNA: GTR rates U[0.5,2]; protein: exchangeabilities Gamma(1,1);
 * codon: kappa U[1,4], omega U[0.1,1]; frequencies Dirichlet(1,...,1)), +G(alpha, ncat),
 * nsites simulated down the tree, compressed to patterns. */
typedef struct plk_problem plk_problem;
int plk_sim_create(plk_problem **out, int seq_type, int ntaxa, int nsites, double mean_len,
                   double alpha, int ncat, uint64_t seed);
/* Same, on a given unrooted topology nei ((2*ntaxa-2) x 3, leaves 0..ntaxa-1 with their
 * neighbour in slot 0) with fresh Exp(mean_len) branch lengths (unlinked partitions). */
int plk_sim_create_on_tree(plk_problem **out, int seq_type, int ntaxa, const int32_t *nei, int nsites,
                           double mean_len, double alpha, int ncat, uint64_t seed);
int plk_sim_destroy(plk_problem *p);
/* dims[0..6] = nstates, ntaxa, nsites, nptn, ncat, state_unknown, seq_type */
int plk_sim_dims(const plk_problem *p, int *dims);
/* Named arrays: "pi", "Q", "eval", "evec", "inv_evec", "rates", "props", "tip_lh",
 * "freq", "blen" (double); "nei" (int32, (2T-2) x 3); "states" (uint8, ntaxa x nptn);
 * "params" (double: alpha, kappa, omega).  *count receives the element count. */
const void *plk_sim_array(const plk_problem *p, const char *name, int64_t *count);

#ifdef __cplusplus
}
#endif
#endif
