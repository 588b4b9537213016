// Host FP64 model arrays: rate matrices, eigensystem, discrete Gamma, tip vectors.
// Everything here is O(n^3) per model and stays on the host in FP64.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "plk.h"
#include "plk_internal.h"
#include "plk_model.h"

extern "C" int plk_state_unknown(int nstates, int seq_type) {
    // alignment/alignment.cpp:1116-1128
    if (seq_type == PLK_SEQ_DNA) return 18;
    if (seq_type == PLK_SEQ_PROTEIN) return 23;
    return nstates;
}

extern "C" int plk_rate_matrix(int n, const double *exch, const double *pi, double *Q) {
    if (n < 2 || n > PLK_MAX_STATES || !exch || !pi || !Q) return plk::fail(PLK_ERR_ARG, "plk_rate_matrix: bad arguments");
    // half-matrix exchangeabilities (model/modelmarkov.cpp:1472-1484), times pi_j,
    // zero row sums and unit mean rate (1502-1514)
    std::vector<double> R((size_t)n * n, 0.0);
    for (int i = 0, k = 0; i < n; i++)
        for (int j = i + 1; j < n; j++, k++) R[i * n + j] = R[j * n + i] = exch[k];
    // zero frequencies are allowed: such a state's column is zero (no rate into it); its
    // row does not enter the normalisation and plk_eigen leaves it out
    // (model/modelmarkov.cpp:1453-1500)
    double norm = 0.0;
    for (int i = 0; i < n; i++) {
        if (!(pi[i] >= 0.0)) return plk::fail(PLK_ERR_ARG, "plk_rate_matrix: state frequencies must be >= 0");
        double rs = 0.0;
        for (int j = 0; j < n; j++) {
            double q = (i == j) ? 0.0 : R[i * n + j] * pi[j];
            Q[i * n + j] = q;
            rs += q;
        }
        Q[i * n + i] = -rs;
        norm += pi[i] * rs;
    }
    if (!(norm > 0.0)) return plk::fail(PLK_ERR_ARG, "plk_rate_matrix: zero total rate");
    for (int i = 0; i < n * n; i++) Q[i] /= norm;
    return PLK_OK;
}

// Cyclic Jacobi eigensolver for a dense symmetric matrix (row-major, overwritten).
// Rutishauser's threshold variant; converges quadratically, accurate to a few ulp of
// the largest eigenvalue, which is what the reversible-model path needs.
static bool jacobi_eigen(int n, std::vector<double> &A, std::vector<double> &w, std::vector<double> &Vm) {
    Vm.assign((size_t)n * n, 0.0);
    for (int i = 0; i < n; i++) Vm[i * n + i] = 1.0;
    w.resize(n);
    for (int sweep = 0; sweep < 100; sweep++) {
        double off = 0.0, diag = 0.0;
        for (int i = 0; i < n; i++) {
            diag += A[i * n + i] * A[i * n + i];
            for (int j = i + 1; j < n; j++) off += A[i * n + j] * A[i * n + j];
        }
        if (off <= 1e-34 * diag || off == 0.0) break;
        for (int p = 0; p < n; p++)
            for (int q = p + 1; q < n; q++) {
                double apq = A[p * n + q];
                if (apq == 0.0) continue;
                double app = A[p * n + p], aqq = A[q * n + q];
                double theta = (aqq - app) / (2.0 * apq);
                double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
                double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < n; k++) {
                    double akp = A[k * n + p], akq = A[k * n + q];
                    A[k * n + p] = c * akp - s * akq;
                    A[k * n + q] = s * akp + c * akq;
                }
                for (int k = 0; k < n; k++) {
                    double apk = A[p * n + k], aqk = A[q * n + k];
                    A[p * n + k] = c * apk - s * aqk;
                    A[q * n + k] = s * apk + c * aqk;
                }
                for (int k = 0; k < n; k++) {
                    double vkp = Vm[k * n + p], vkq = Vm[k * n + q];
                    Vm[k * n + p] = c * vkp - s * vkq;
                    Vm[k * n + q] = s * vkp + c * vkq;
                }
            }
    }
    for (int i = 0; i < n; i++) w[i] = A[i * n + i];
    return true;
}

static int eigen_full(int n, const double *Q, const double *pi, double *eval, double *evec, double *inv_evec);

constexpr double kZeroFreq = 1e-10;  // ZERO_FREQ (utils/eigendecomposition.h:24)

extern "C" int plk_eigen(int n, const double *Q, const double *pi, double *eval, double *evec, double *inv_evec) {
    if (n < 2 || n > PLK_MAX_STATES || !Q || !pi || !eval || !evec || !inv_evec)
        return plk::fail(PLK_ERR_ARG, "plk_eigen: bad arguments");
    std::vector<int> nz;
    for (int i = 0; i < n; i++) {
        if (!(pi[i] >= 0.0)) return plk::fail(PLK_ERR_ARG, "plk_eigen: state frequencies must be >= 0");
        if (pi[i] > kZeroFreq) nz.push_back(i);
    }
    const int m = (int)nz.size();
    if (m == n) return eigen_full(n, Q, pi, eval, evec, inv_evec);
    if (m < 2) return plk::fail(PLK_ERR_ARG, "plk_eigen: fewer than two states with non-zero frequency");
    // States with frequency <= ZERO_FREQ (model/modelmarkov.cpp:1453-1500, 1563-1597): the
    // system of the other states, their frequencies renormalised and their rows of Q
    // rescaled to one substitution per unit time, is decomposed alone; eigenvalue k of it
    // goes to the k-th non-zero state's position, its eigenvectors fill the non-zero rows
    // and columns, and a zero-frequency state gets eigenvalue 0 and unit vectors.
    std::vector<double> qs((size_t)m * m), ps(m), ev(m), U((size_t)m * m), Ui((size_t)m * m);
    double psum = 0.0;
    for (int a = 0; a < m; a++) psum += pi[nz[a]];
    for (int a = 0; a < m; a++) ps[a] = pi[nz[a]] / psum;
    double norm = 0.0;
    for (int a = 0; a < m; a++) {
        double rs = 0.0;
        for (int b = 0; b < m; b++) {
            qs[a * m + b] = a == b ? 0.0 : Q[nz[a] * n + nz[b]];
            rs += qs[a * m + b];
        }
        qs[a * m + a] = -rs;
        norm += ps[a] * rs;
    }
    if (!(norm > 0.0)) return plk::fail(PLK_ERR_ARG, "plk_eigen: zero total rate among the non-zero states");
    for (double &v : qs) v /= norm;
    int rc = eigen_full(m, qs.data(), ps.data(), ev.data(), U.data(), Ui.data());
    if (rc) return rc;
    for (int i = 0; i < n * n; i++) evec[i] = inv_evec[i] = 0.0;
    for (int i = 0; i < n; i++) {
        eval[i] = 0.0;
        evec[i * n + i] = inv_evec[i * n + i] = 1.0;
    }
    for (int a = 0; a < m; a++) {
        const int i = nz[a];
        eval[i] = std::isnan(ev[a]) ? 0.0 : ev[a];
        evec[i * n + i] = inv_evec[i * n + i] = 0.0;
        for (int b = 0; b < m; b++) {
            evec[i * n + nz[b]] = U[a * m + b];
            inv_evec[i * n + nz[b]] = Ui[a * m + b];
        }
    }
    return PLK_OK;
}

static int eigen_full(int n, const double *Q, const double *pi, double *eval, double *evec, double *inv_evec) {
    // S = Pi^1/2 Q Pi^-1/2 (model/modelmarkov.cpp:1520), symmetric for reversible Q
    std::vector<double> s(n), S((size_t)n * n), w, Vm;
    for (int i = 0; i < n; i++) {
        if (!(pi[i] > 0.0)) return plk::fail(PLK_ERR_ARG, "plk_eigen: state frequencies must be > 0");
        s[i] = std::sqrt(pi[i]);
    }
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) S[i * n + j] = s[i] * Q[i * n + j] / s[j];
    double asym = 0.0;
    for (int i = 0; i < n; i++)
        for (int j = i + 1; j < n; j++) {
            asym = std::max(asym, std::fabs(S[i * n + j] - S[j * n + i]));
            double m = 0.5 * (S[i * n + j] + S[j * n + i]);
            S[i * n + j] = S[j * n + i] = m;
        }
    if (asym > 0.1) return plk::fail(PLK_ERR_ARG, "plk_eigen: rate matrix is not reversible for these frequencies");
    jacobi_eigen(n, S, w, Vm);
    // ascending eigenvalues, like SelfAdjointEigenSolver (model/modelmarkov.cpp:1535)
    std::vector<int> order(n);
    for (int i = 0; i < n; i++) order[i] = i;
    std::sort(order.begin(), order.end(), [&](int a, int b) { return w[a] < w[b]; });
    for (int k = 0; k < n; k++) {
        int col = order[k];
        eval[k] = w[col];
        for (int i = 0; i < n; i++) {
            evec[i * n + k] = Vm[i * n + col] / s[i];          // U = Pi^-1/2 V
            inv_evec[k * n + i] = Vm[i * n + col] * s[i];      // U^-1 = V^T Pi^1/2
        }
    }
    return PLK_OK;
}

// ---- discrete Gamma (model/rategamma.cpp:98-171) -----------------------------------
// The reference computes the category boundaries with published algorithms, restated
// here from their descriptions: Pike & Hill (1966) log-gamma (Stirling series),
// Bhattacharjee (1970) AS 32 incomplete gamma ratio, Odeh & Evans (1974) AS 70 normal
// percentage points, Best & Roberts (1975) AS 91 chi-square percentage points.

static double ln_gamma(double alpha) {
    double x = alpha, f = 0.0, z;
    if (x < 7.0) {
        f = 1.0;
        z = x - 1.0;
        while (++z < 7.0) f *= z;
        x = z;
        f = -std::log(f);
    }
    z = 1.0 / (x * x);
    return f + (x - 0.5) * std::log(x) - x + 0.918938533204673 +
           (((-0.000595238095238 * z + 0.000793650793651) * z - 0.002777777777778) * z + 0.083333333333333) / x;
}

static double incomplete_gamma(double x, double alpha, double ln_gamma_alpha) {
    const double accurate = 1e-8, overflow = 1e30;
    if (x == 0.0) return 0.0;
    if (x < 0.0 || alpha <= 0.0) return -1.0;
    double factor = std::exp(alpha * std::log(x) - x - ln_gamma_alpha);
    if (x <= 1.0 || x < alpha) {
        // series expansion
        double gin = 1.0, term = 1.0, rn = alpha;
        do {
            rn += 1.0;
            term *= x / rn;
            gin += term;
        } while (term > accurate);
        return gin * factor / alpha;
    }
    // continued fraction
    double a = 1.0 - alpha, b = a + x + 1.0, term = 0.0;
    double pn[6] = {1.0, x, x + 1.0, x * b, 0.0, 0.0};
    double gin = pn[2] / pn[3];
    for (;;) {
        a += 1.0;
        b += 2.0;
        term += 1.0;
        double an = a * term;
        for (int i = 0; i < 2; i++) pn[i + 4] = b * pn[i + 2] - an * pn[i];
        if (pn[5] != 0.0) {
            double rn = pn[4] / pn[5];
            double dif = std::fabs(gin - rn);
            if (dif <= accurate && dif <= accurate * rn) return 1.0 - factor * gin;
            gin = rn;
        }
        for (int i = 0; i < 4; i++) pn[i] = pn[i + 2];
        if (std::fabs(pn[4]) >= overflow)
            for (int i = 0; i < 4; i++) pn[i] /= overflow;
    }
}

static double point_normal(double prob) {
    const double a0 = -0.322232431088, a1 = -1.0, a2 = -0.342242088547, a3 = -0.0204231210245;
    const double a4 = -0.453642210148e-4, b0 = 0.0993484626060, b1 = 0.588581570495;
    const double b2 = 0.531103462366, b3 = 0.103537752850, b4 = 0.0038560700634;
    double p1 = prob < 0.5 ? prob : 1.0 - prob;
    if (p1 < 1e-20) return -9999.0;
    double y = std::sqrt(std::log(1.0 / (p1 * p1)));
    double z = y + ((((y * a4 + a3) * y + a2) * y + a1) * y + a0) / ((((y * b4 + b3) * y + b2) * y + b1) * y + b0);
    return prob < 0.5 ? -z : z;
}

static double point_chi2(double prob, double v) {
    const double e = 0.5e-6, aa = 0.6931471805;
    if (prob < 0.000002 || prob > 0.999998 || v <= 0.0) return -1.0;
    double g = ln_gamma(v / 2.0);
    double xx = v / 2.0, c = xx - 1.0, ch;
    if (v < -1.24 * std::log(prob)) {
        ch = std::pow(prob * xx * std::exp(g + xx * aa), 1.0 / xx);
        if (ch - e < 0.0) return ch;
    } else if (v <= 0.32) {
        ch = 0.4;
        double a = std::log(1.0 - prob), q;
        do {
            q = ch;
            double p1 = 1.0 + ch * (4.67 + ch);
            double p2 = ch * (6.73 + ch * (6.66 + ch));
            double t = -0.5 + (4.67 + 2.0 * ch) / p1 - (6.73 + ch * (13.32 + 3.0 * ch)) / p2;
            ch -= (1.0 - std::exp(a + g + 0.5 * ch + c * aa) * p2 / p1) / t;
        } while (std::fabs(q / ch - 1.0) - 0.01 > 0.0);
    } else {
        double x = point_normal(prob);
        double p1 = 0.222222 / v;
        ch = v * std::pow(x * std::sqrt(p1) + 1.0 - p1, 3.0);
        if (ch > 2.2 * v + 6.0) ch = -2.0 * (std::log(1.0 - prob) - c * std::log(0.5 * ch) + g);
    }
    // seven-term Taylor refinement (AS 91 step 4)
    double q;
    do {
        q = ch;
        double p1 = 0.5 * ch;
        double t = incomplete_gamma(p1, xx, g);
        if (t < 0.0) return -1.0;
        double p2 = prob - t;
        t = p2 * std::exp(xx * aa + g + p1 - c * std::log(ch));
        double b = t / ch, a = 0.5 * t - b * c;
        double s1 = (210 + a * (140 + a * (105 + a * (84 + a * (70 + 60 * a))))) / 420;
        double s2 = (420 + a * (735 + a * (966 + a * (1141 + 1278 * a)))) / 2520;
        double s3 = (210 + a * (462 + a * (707 + 932 * a))) / 2520;
        double s4 = (252 + a * (672 + 1182 * a) + c * (294 + a * (889 + 1740 * a))) / 5040;
        double s5 = (84 + 264 * a + c * (175 + 606 * a)) / 2520;
        double s6 = (120 + c * (346 + 127 * c)) / 5040;
        ch += t * (1 + 0.5 * t * s1 - b * c * (s1 - b * (s2 - b * (s3 - b * (s4 - b * (s5 - b * s6))))));
    } while (std::fabs(q / ch - 1.0) > e);
    return ch;
}

extern "C" int plk_gamma_rates(double alpha, int ncat, double *rates) {
    if (!(alpha > 0.0) || ncat < 1 || ncat > PLK_MAX_CATS || !rates) return plk::fail(PLK_ERR_ARG, "plk_gamma_rates: bad arguments");
    if (ncat == 1) {
        rates[0] = 1.0;
        return PLK_OK;
    }
    // computeRatesMean: cut points of Gamma(alpha, 1/alpha) from chi2 quantiles (Eq. 9),
    // category means from incomplete-gamma(alpha+1) differences (Eq. 10)
    double lnga1 = ln_gamma(alpha + 1.0);
    std::vector<double> freqK(ncat);
    for (int i = 0; i < ncat - 1; i++) freqK[i] = point_chi2((i + 1.0) / ncat, 2.0 * alpha) / (2.0 * alpha);
    for (int i = 0; i < ncat - 1; i++) freqK[i] = incomplete_gamma(freqK[i] * alpha, alpha + 1.0, lnga1);
    rates[0] = freqK[0] * ncat;
    rates[ncat - 1] = (1.0 - freqK[ncat - 2]) * ncat;
    for (int i = 1; i < ncat - 1; i++) rates[i] = (freqK[i] - freqK[i - 1]) * ncat;
    // keep sum(rates) == ncat (model/rategamma.cpp:133-140)
    double sum = 0.0;
    for (int i = 0; i < ncat; i++) sum += rates[i];
    if (sum != (double)ncat)
        for (int i = 0; i < ncat; i++) rates[i] *= ncat / sum;
    return PLK_OK;
}

extern "C" int plk_gamma_rates_median(double alpha, int ncat, double *rates) {
    if (!(alpha > 0.0) || ncat < 1 || ncat > PLK_MAX_CATS || !rates)
        return plk::fail(PLK_ERR_ARG, "plk_gamma_rates_median: bad arguments");
    if (ncat == 1) {
        rates[0] = 1.0;
        return PLK_OK;
    }
    double sum = 0.0;
    for (int k = 0; k < ncat; k++) {
        const double q = point_chi2((2.0 * k + 1.0) / (2.0 * ncat), 2.0 * alpha) / (2.0 * alpha);
        rates[k] = std::fabs(q);
        sum += rates[k];
    }
    for (int k = 0; k < ncat; k++) rates[k] = rates[k] * ncat / sum;
    return PLK_OK;
}

extern "C" int plk_tip_lh(int n, int seq_type, const double *inv_evec, double *tip_lh) {
    if (n < 2 || n > PLK_MAX_STATES || !inv_evec || !tip_lh) return plk::fail(PLK_ERR_ARG, "plk_tip_lh: bad arguments");
    int su = plk_state_unknown(n, seq_type);
    std::vector<double> lik(n);
    for (int s = 0; s <= su; s++) {
        std::fill(lik.begin(), lik.end(), 0.0);
        if (s < n) {
            lik[s] = 1.0;
        } else if (seq_type == PLK_SEQ_DNA && s < 18) {
            // ambiguity bitmask cstate = s - 3 (model/modeldna.cpp:462-476)
            int cs = s - n + 1;
            for (int j = 0; j < n; j++)
                if (cs & (1 << j)) lik[j] = 1.0;
        } else if (seq_type == PLK_SEQ_PROTEIN && s < n + 3) {
            // B = N|D, Z = Q|E, J = I|L (model/modelprotein.cpp:1240-1256)
            static const int ambi[6] = {2, 3, 5, 6, 9, 10};
            lik[ambi[(s - n) * 2]] = 1.0;
            lik[ambi[(s - n) * 2 + 1]] = 1.0;
        } else {
            std::fill(lik.begin(), lik.end(), 1.0);
        }
        // ModelSubst::multiplyWithInvEigenvector (model/modelsubst.cpp:183-196)
        for (int i = 0; i < n; i++) {
            double v = 0.0;
            for (int j = 0; j < n; j++) v += inv_evec[i * n + j] * lik[j];
            tip_lh[s * n + i] = v;
        }
    }
    return PLK_OK;
}

// Invariant-site mass per pattern.  Alignment::computeConst (alignment/alignment.cpp:1321-1420)
// intersects, over the taxa, the set of states each observed state allows
// (Alignment::getAppearance, :5567-5612) and codes the result as const_char: a single state,
// STATE_UNKNOWN when every state is allowed, a DNA ambiguity code, protein B/Z/J, or
// STATE_UNKNOWN + 1 when the intersection is empty.  PhyloTree::computePtnInvar
// (tree/phylotreesse.cpp:549-612) then sets p_invar times the frequency of that state (times
// the sum of the allowed states' frequencies for an ambiguity code, 1 for STATE_UNKNOWN) and 0
// for non-constant patterns.
extern "C" int plk_ptn_invar(int n, int seq_type, int ntaxa, int nptn, const uint8_t *states, const double *pi,
                             double p_invar, double *ptn_invar) {
    if (n < 2 || n > PLK_MAX_STATES || ntaxa < 1 || nptn < 0 || (nptn > 0 && (!states || !ptn_invar)) || !pi ||
        !(p_invar >= 0.0 && p_invar < 1.0))
        return plk::fail(PLK_ERR_ARG, "plk_ptn_invar: bad arguments");
    const int su = plk_state_unknown(n, seq_type);
    // allowed-state sets as bit masks over up to 64 states (61 codons fit one word)
    std::vector<uint64_t> allow(256, 0);
    const uint64_t all = (n == 64) ? ~0ull : ((1ull << n) - 1);
    for (int s = 0; s < 256; s++) {
        if (s == su) allow[s] = all;
        else if (s < n) allow[s] = 1ull << s;
        else if (seq_type == PLK_SEQ_DNA && s < su) allow[s] = (uint64_t)(s - (n - 1)) & all;
        else if (seq_type == PLK_SEQ_PROTEIN && s < n + 3) {
            static const int ambi[6] = {2, 3, 5, 6, 9, 10};  // B = N|D, Z = Q|E, J = I|L
            allow[s] = (1ull << ambi[(s - n) * 2]) | (1ull << ambi[(s - n) * 2 + 1]);
        } else allow[s] = ~0ull;  // not a valid state code: checked below
    }
    for (int p = 0; p < nptn; p++) {
        uint64_t m = all;
        for (int t = 0; t < ntaxa; t++) {
            const uint8_t s = states[(size_t)t * nptn + p];
            if (s > su) return plk::fail(PLK_ERR_ARG, "plk_ptn_invar: state code above STATE_UNKNOWN");
            m &= allow[s];
        }
        double v = 0.0;
        if (m == all) {
            v = p_invar;  // const_char == STATE_UNKNOWN
        } else if (m != 0) {
            for (int x = 0; x < n; x++)
                if (m & (1ull << x)) v += pi[x];
            v *= p_invar;
        }
        ptn_invar[p] = v;
    }
    return PLK_OK;
}

// Standard genetic code over codons AAA..TTT (A,C,G,T order), '*' = stop
// (alignment/alignment.cpp:44 uses the same NCBI table 1 in this order).
namespace plk {
const char kStandardGeneticCode[65] = "KNKNTTTTRSRSIIMIQHQHPPPPRRRRLLLLEDEDAAAAGGGGVVVV*Y*YSSSS*CWCLFLF";
}
static const char *const kStandardCode = plk::kStandardGeneticCode;

extern "C" int plk_codon_exch(double kappa, double omega, double *exch) {
    if (!(kappa > 0.0) || !(omega > 0.0) || !exch) return plk::fail(PLK_ERR_ARG, "plk_codon_exch: bad arguments");
    int sense[64], n = 0;
    for (int c = 0; c < 64; c++)
        if (kStandardCode[c] != '*') sense[n++] = c;
    // single-nucleotide changes only (CA_MULTI_NT -> 0), transitions are index distance 2
    // in A,C,G,T; kappa on transitions, omega on non-synonymous (CK_ONE_KAPPA)
    for (int a = 0, k = 0; a < n; a++)
        for (int b = a + 1; b < n; b++, k++) {
            int ci = sense[a], cj = sense[b], diffs = 0, ts = 0;
            for (int sh = 4; sh >= 0; sh -= 2) {
                int n1 = (ci >> sh) & 3, n2 = (cj >> sh) & 3;
                if (n1 != n2) {
                    diffs++;
                    ts = std::abs(n1 - n2) == 2;
                }
            }
            double r = 0.0;
            if (diffs == 1) {
                bool syn = kStandardCode[ci] == kStandardCode[cj];
                r = 1.0;
                if (syn) {
                    if (ts) r *= kappa;
                } else {
                    r *= ts ? omega * kappa : omega;
                }
            }
            exch[k] = r;
        }
    return PLK_OK;
}
