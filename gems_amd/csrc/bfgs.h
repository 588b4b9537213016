// Multi-dimensional minimisation for the substitution-model parameters: the quasi-Newton
// method of Davidon-Fletcher-Powell / Broyden-Fletcher-Goldfarb-Shanno with a
// backtracking line search (Numerical Recipes, 2nd ed., 10.7 and 9.7), on gradients by
// forward differences, as IQ-TREE's Optimization::minimizeMultiDimen runs it.  Host code;
// every function evaluation is a likelihood on the device, so the evaluator returns a
// status.  Vectors are 0-based here.
#pragma once

#include <algorithm>
#include <cmath>
#include <vector>

#include "plk.h"

namespace plk {

// fixBound (utils/optimization.cpp:149-157): clamp x into [lower, upper]
inline void fix_bound(std::vector<double> &x, const std::vector<double> &lo, const std::vector<double> &hi) {
    for (size_t i = 0; i < x.size(); i++) {
        if (x[i] < lo[i]) x[i] = lo[i];
        else if (x[i] > hi[i]) x[i] = hi[i];
    }
}

// Optimization::derivativeFunk (utils/optimization.cpp:916-939): f(x) and the forward
// difference quotients with step ERROR_X = 1e-4 relative (absolute when x_i = 0), the
// step taken as the representable difference (x + h) - x.
template <class F>
int forward_gradient(F &&f, std::vector<double> &x, double &fx, std::vector<double> &g) {
    const double kErrorX = 1.0e-4;
    const size_t n = x.size();
    std::vector<double> h(n);
    int rc = f(x, fx);
    if (rc) return rc;
    g.assign(n, 0.0);
    for (size_t d = 0; d < n; d++) {
        const double t = x[d];
        h[d] = kErrorX * std::fabs(t);
        if (h[d] == 0.0) h[d] = kErrorX;
        x[d] = t + h[d];
        h[d] = x[d] - t;
        if ((rc = f(x, g[d]))) return rc;
        x[d] = t;
    }
    for (size_t d = 0; d < n; d++) g[d] = (g[d] - fx) / h[d];
    return PLK_OK;
}

// Optimization::lnsrch (utils/optimization.cpp:639-718): from xold along p, accept the
// first step length (1, then quadratic / cubic backtracking, at least a tenth of the last)
// that decreases f by the fraction ALF = 1e-4 of the slope; steps are clamped into the
// bounds; a step below TOLX = 1e-7 relative returns xold with check = 1.
template <class F>
int line_search(F &&f, const std::vector<double> &xold, double fold, const std::vector<double> &g,
                std::vector<double> &p, std::vector<double> &x, double &fx, double stpmax, int &check,
                const std::vector<double> &lo, const std::vector<double> &hi) {
    const double kAlf = 1.0e-4, kTolX = 1.0e-7;
    const size_t n = xold.size();
    check = 0;
    double sum = 0.0;
    for (size_t i = 0; i < n; i++) sum += p[i] * p[i];
    sum = std::sqrt(sum);
    if (sum > stpmax)
        for (size_t i = 0; i < n; i++) p[i] *= stpmax / sum;
    double slope = 0.0;
    for (size_t i = 0; i < n; i++) slope += g[i] * p[i];
    double test = 0.0;
    for (size_t i = 0; i < n; i++) test = std::max(test, std::fabs(p[i]) / std::max(std::fabs(xold[i]), 1.0));
    const double alamin = kTolX / test;
    double alam = 1.0, alam2 = 0.0, f2 = 0.0, fold2 = 0.0, tmplam = 0.0;
    bool first = true;
    x.assign(n, 0.0);
    for (;;) {
        for (size_t i = 0; i < n; i++) x[i] = xold[i] + alam * p[i];
        fix_bound(x, lo, hi);
        int rc = f(x, fx);
        if (rc) return rc;
        if (alam < alamin) {
            x = xold;
            check = 1;
            return PLK_OK;
        }
        if (fx <= fold + kAlf * alam * slope) return PLK_OK;
        if (first) {
            tmplam = -slope / (2.0 * (fx - fold - slope));
        } else {
            const double rhs1 = fx - fold - alam * slope, rhs2 = f2 - fold2 - alam2 * slope;
            const double a = (rhs1 / (alam * alam) - rhs2 / (alam2 * alam2)) / (alam - alam2);
            const double b = (-alam2 * rhs1 / (alam * alam) + alam * rhs2 / (alam2 * alam2)) / (alam - alam2);
            if (a == 0.0) {
                tmplam = -slope / (2.0 * b);
            } else {
                const double disc = b * b - 3.0 * a * slope;
                if (disc < 0.0) tmplam = 0.5 * alam;
                else if (b <= 0.0) tmplam = (-b + std::sqrt(disc)) / (3.0 * a);
                else tmplam = -slope / (b + std::sqrt(disc));
            }
            if (tmplam > 0.5 * alam) tmplam = 0.5 * alam;
        }
        alam2 = alam;
        f2 = fx;
        fold2 = fold;
        alam = std::max(tmplam, 0.1 * alam);
        first = false;
    }
}

// Optimization::dfpmin (utils/optimization.cpp:781-906): BFGS with an inverse-Hessian
// estimate started at the identity, at most ITMAX = 200 iterations, stopping when a step
// moves no coordinate by more than TOLX = 4 EPS (EPS = 3e-8) relative or when the scaled
// gradient max_i |g_i| max(|x_i|, 1) / max(|f|, 1) falls below gtol.  The update is
// skipped when fac^2 <= EPS |dg|^2 |dx|^2.  Returns the minimum in fret, x at it.
template <class F>
int bfgs_minimize(F &&f, std::vector<double> &x, const std::vector<double> &lo, const std::vector<double> &hi,
                  double gtol, double &fret) {
    const int kItMax = 200;
    const double kEps = 3.0e-8, kTolX = 4 * kEps, kStpMx = 100.0;
    const size_t n = x.size();
    std::vector<double> g, dg(n), hdg(n), pnew, xi(n);
    std::vector<std::vector<double>> hessin(n, std::vector<double>(n, 0.0));
    double fp = 0.0;
    int rc = forward_gradient(f, x, fp, g);
    if (rc) return rc;
    double sum = 0.0;
    for (size_t i = 0; i < n; i++) {
        xi[i] = -g[i];
        sum += x[i] * x[i];
        hessin[i][i] = 1.0;
    }
    const double stpmax = kStpMx * std::max(std::sqrt(sum), (double)n);
    fret = fp;
    for (int its = 1; its <= kItMax; its++) {
        int check = 0;
        if ((rc = line_search(f, x, fp, g, xi, pnew, fret, stpmax, check, lo, hi))) return rc;
        fp = fret;
        double test = 0.0;
        for (size_t i = 0; i < n; i++) {
            xi[i] = pnew[i] - x[i];
            x[i] = pnew[i];
        }
        for (size_t i = 0; i < n; i++) test = std::max(test, std::fabs(xi[i]) / std::max(std::fabs(x[i]), 1.0));
        if (test < kTolX) return PLK_OK;
        dg = g;
        double fdummy = 0.0;
        if ((rc = forward_gradient(f, x, fdummy, g))) return rc;
        test = 0.0;
        const double den = std::max(std::fabs(fret), 1.0);
        for (size_t i = 0; i < n; i++) test = std::max(test, std::fabs(g[i]) * std::max(std::fabs(x[i]), 1.0) / den);
        if (test < gtol) return PLK_OK;
        for (size_t i = 0; i < n; i++) dg[i] = g[i] - dg[i];
        for (size_t i = 0; i < n; i++) {
            hdg[i] = 0.0;
            for (size_t j = 0; j < n; j++) hdg[i] += hessin[i][j] * dg[j];
        }
        double fac = 0.0, fae = 0.0, sumdg = 0.0, sumxi = 0.0;
        for (size_t i = 0; i < n; i++) {
            fac += dg[i] * xi[i];
            fae += dg[i] * hdg[i];
            sumdg += dg[i] * dg[i];
            sumxi += xi[i] * xi[i];
        }
        if (fac * fac > kEps * sumdg * sumxi) {
            fac = 1.0 / fac;
            const double fad = 1.0 / fae;
            for (size_t i = 0; i < n; i++) dg[i] = fac * xi[i] - fad * hdg[i];
            for (size_t i = 0; i < n; i++)
                for (size_t j = 0; j < n; j++)
                    hessin[i][j] += fac * xi[i] * xi[j] - fad * hdg[i] * hdg[j] + fae * dg[i] * dg[j];
        }
        for (size_t i = 0; i < n; i++) {
            xi[i] = 0.0;
            for (size_t j = 0; j < n; j++) xi[i] -= hessin[i][j] * g[j];
        }
    }
    return PLK_OK;
}

}  // namespace plk
