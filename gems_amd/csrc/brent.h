// One-dimensional minimisation for the model-parameter loop: Brent's method (parabolic
// interpolation with golden-section fall-back, R. P. Brent, "Algorithms for Minimization
// without Derivatives", 1973) behind the bracketing rule of IQ-TREE's
// Optimization::minimizeOneDimen.  Host code; every function evaluation is a likelihood
// on the device, so the evaluator returns a status.
#pragma once

#include <cmath>

#include "plk.h"

namespace plk {

// Optimization::brent_opt (utils/optimization.cpp:174-310).  The iterate sequence has to
// be the reference's for the optimum to match it, so the constants (100 iterations,
// golden fraction 0.3819660, absolute tolerance floor 1e-10) and the acceptance tests
// are the same.  f(x, &fx) evaluates; [lo, hi] brackets the minimum with x0 inside and
// f(x0) <= f(lo), f(hi).  On return xopt/fopt hold the best point and f2 the curvature
// of the parabola through the last three points.
template <class F>
int brent_minimize(F &&f, double lo_in, double x0, double hi_in, double tol, double f_lo, double f_x0, double f_hi,
                   double &xopt, double &fopt, double &f2) {
    const int kMaxIter = 100;
    const double kGold = 0.3819660, kFloor = 1e-10;
    double lo = std::min(lo_in, hi_in), hi = std::max(lo_in, hi_in);
    // x: best so far; w: second best; v: previous w
    double x = x0, fx = f_x0, w, fw, v, fv;
    if (f_lo < f_hi) {
        w = lo_in, fw = f_lo, v = hi_in, fv = f_hi;
    } else {
        w = hi_in, fw = f_hi, v = lo_in, fv = f_lo;
    }
    double step = 0.0, prev_step = 0.0;
    auto curvature = [&]() {
        const double xw = x - w, wv = w - v, vx = v - x;
        return 2.0 * (fv * xw + fx * wv + fw * vx) / (v * v * xw + x * x * wv + w * w * vx);
    };
    for (int it = 0; it < kMaxIter; it++) {
        const double mid = 0.5 * (lo + hi);
        const double tol1 = tol * std::fabs(x) + kFloor, tol2 = 2.0 * tol1;
        if (std::fabs(x - mid) <= tol2 - 0.5 * (hi - lo)) break;
        bool golden = true;
        if (std::fabs(prev_step) > tol1) {
            // parabola through (x, fx), (w, fw), (v, fv)
            const double r = (x - w) * (fx - fv);
            double q = (x - v) * (fx - fw);
            double p = (x - v) * q - (x - w) * r;
            q = 2.0 * (q - r);
            if (q > 0.0) p = -p;
            q = std::fabs(q);
            const double older = prev_step;
            prev_step = step;
            if (!(std::fabs(p) >= std::fabs(0.5 * q * older) || p <= q * (lo - x) || p >= q * (hi - x))) {
                golden = false;
                step = p / q;
                const double u = x + step;
                if (u - lo < tol2 || hi - u < tol2) step = mid - x >= 0.0 ? std::fabs(tol1) : -std::fabs(tol1);
            }
        }
        if (golden) {
            prev_step = x >= mid ? lo - x : hi - x;
            step = kGold * prev_step;
        }
        const double u = std::fabs(step) >= tol1 ? x + step : x + (step >= 0.0 ? std::fabs(tol1) : -std::fabs(tol1));
        double fu = 0.0;
        int rc = f(u, fu);
        if (rc) return rc;
        if (fu <= fx) {
            if (u >= x) lo = x;
            else hi = x;
            v = w, fv = fw;
            w = x, fw = fx;
            x = u, fx = fu;
        } else {
            if (u < x) lo = u;
            else hi = u;
            if (fu <= fw || w == x) {
                v = w, fv = fw;
                w = u, fw = fu;
            } else if (fu <= fv || v == x || v == w) {
                v = u, fv = fu;
            }
        }
    }
    xopt = x;
    fopt = fx;
    f2 = curvature();
    return PLK_OK;
}

// Optimization::minimizeOneDimen (utils/optimization.cpp:322-372): try the bracket
// xguess -/+ 50 * tol * xguess; if either end is lower than xguess fall back to the
// whole interval [xmin, xmax]; minimise with brent_minimize and return xguess instead
// when that did worse than f(xguess).
template <class F>
int minimize_one_dimen(F &&f, double xmin, double xguess, double xmax, double tol, double &xopt, double &fopt,
                       double &f2) {
    xguess = std::min(std::max(xguess, xmin), xmax);
    const double eps = xguess * tol * 50.0;
    double a = std::max(xguess - eps, xmin), b = xguess, c = std::min(xguess + eps, xmax);
    double fa, fb, fc;
    int rc = f(b, fb);
    if (rc) return rc;
    fa = fb;
    fc = fb;
    if (a < b && (rc = f(a, fa))) return rc;
    if (b < c && (rc = f(c, fc))) return rc;
    if (fa < fb || fc < fb) {
        if (a != xmin && (rc = f(xmin, fa))) return rc;
        if (c != xmax && (rc = f(xmax, fc))) return rc;
        a = xmin;
        c = xmax;
    }
    rc = brent_minimize(f, a, b, c, tol, fa, fb, fc, xopt, fopt, f2);
    if (rc) return rc;
    if (fopt > fb) {
        // worse than the start: evaluate there again (leaves the model at xguess)
        if ((rc = f(b, fopt))) return rc;
        xopt = b;
    }
    return PLK_OK;
}

}  // namespace plk
