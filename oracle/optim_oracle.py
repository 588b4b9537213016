"""optim_oracle.py -- restatement of IQ-TREE's multi-dimensional minimiser.

TEST INFRASTRUCTURE ONLY (the checker): tests/ load it; the product never does.

Optimization::minimizeMultiDimen (utils/optimization.cpp:750-779) runs dfpmin
(781-906): BFGS (Numerical Recipes, 2nd ed., 10.7) with the backtracking line search
lnsrch (639-718) on forward-difference gradients from derivativeFunk (916-939), steps
clamped into the bounds by fixBound (149-157).  Restated here with 0-based vectors; the
arithmetic is written in the reference's order so that the same function values give the
same iterates.

Pinning (tests/test_oracle_params.py): quadratic bowls and a Rosenbrock valley against
their known minima and scipy's BFGS, a bound-clamped minimum, and the first evaluations
(f(x), then x + h e_d with h = 1e-4 |x_d|).
"""
import math

ERROR_X = 1.0e-4


def fix_bound(x, lower, upper):
    for i in range(len(x)):
        if x[i] < lower[i]:
            x[i] = lower[i]
        elif x[i] > upper[i]:
            x[i] = upper[i]


def derivative_funk(f, x):
    """(f(x), forward differences); x is restored."""
    n = len(x)
    fx = f(list(x))
    h = [0.0] * n
    dfx = [0.0] * n
    for d in range(n):
        t = x[d]
        h[d] = ERROR_X * abs(t)
        if h[d] == 0.0:
            h[d] = ERROR_X
        x[d] = t + h[d]
        h[d] = x[d] - t
        dfx[d] = f(list(x))
        x[d] = t
    for d in range(n):
        dfx[d] = (dfx[d] - fx) / h[d]
    return fx, dfx


def lnsrch(f, xold, fold, g, p, stpmax, lower, upper):
    """Returns (x, f(x), check); p is scaled in place when longer than stpmax."""
    ALF, TOLX = 1.0e-4, 1.0e-7
    n = len(xold)
    s = math.sqrt(sum(v * v for v in p))
    if s > stpmax:
        for i in range(n):
            p[i] *= stpmax / s
    slope = 0.0
    for i in range(n):
        slope += g[i] * p[i]
    test = 0.0
    for i in range(n):
        test = max(test, abs(p[i]) / max(abs(xold[i]), 1.0))
    alamin = TOLX / test
    alam, alam2, f2, fold2, tmplam = 1.0, 0.0, 0.0, 0.0, 0.0
    first = True
    while True:
        x = [xold[i] + alam * p[i] for i in range(n)]
        fix_bound(x, lower, upper)
        fx = f(list(x))
        if alam < alamin:
            return list(xold), fx, 1
        if fx <= fold + ALF * alam * slope:
            return x, fx, 0
        if first:
            tmplam = -slope / (2.0 * (fx - fold - slope))
        else:
            rhs1 = fx - fold - alam * slope
            rhs2 = f2 - fold2 - alam2 * slope
            a = (rhs1 / (alam * alam) - rhs2 / (alam2 * alam2)) / (alam - alam2)
            b = (-alam2 * rhs1 / (alam * alam) + alam * rhs2 / (alam2 * alam2)) / (alam - alam2)
            if a == 0.0:
                tmplam = -slope / (2.0 * b)
            else:
                disc = b * b - 3.0 * a * slope
                if disc < 0.0:
                    tmplam = 0.5 * alam
                elif b <= 0.0:
                    tmplam = (-b + math.sqrt(disc)) / (3.0 * a)
                else:
                    tmplam = -slope / (b + math.sqrt(disc))
            if tmplam > 0.5 * alam:
                tmplam = 0.5 * alam
        alam2 = alam
        f2 = fx
        fold2 = fold
        alam = max(tmplam, 0.1 * alam)
        first = False


def dfpmin(f, x, lower, upper, gtol):
    """BFGS from x (updated in place); returns the minimum."""
    ITMAX, EPS, STPMX = 200, 3.0e-8, 100.0
    TOLX = 4 * EPS
    n = len(x)
    fp, g = derivative_funk(f, x)
    xi = [-v for v in g]
    s = sum(v * v for v in x)
    hessin = [[1.0 if i == j else 0.0 for j in range(n)] for i in range(n)]
    stpmax = STPMX * max(math.sqrt(s), float(n))
    fret = fp
    for _ in range(ITMAX):
        pnew, fret, _check = lnsrch(f, x, fp, g, xi, stpmax, lower, upper)
        fp = fret
        for i in range(n):
            xi[i] = pnew[i] - x[i]
            x[i] = pnew[i]
        test = 0.0
        for i in range(n):
            test = max(test, abs(xi[i]) / max(abs(x[i]), 1.0))
        if test < TOLX:
            return fret
        dg = list(g)
        _f, g = derivative_funk(f, x)
        test = 0.0
        den = max(abs(fret), 1.0)
        for i in range(n):
            test = max(test, abs(g[i]) * max(abs(x[i]), 1.0) / den)
        if test < gtol:
            return fret
        for i in range(n):
            dg[i] = g[i] - dg[i]
        hdg = [0.0] * n
        for i in range(n):
            for j in range(n):
                hdg[i] += hessin[i][j] * dg[j]
        fac = fae = sumdg = sumxi = 0.0
        for i in range(n):
            fac += dg[i] * xi[i]
            fae += dg[i] * hdg[i]
            sumdg += dg[i] * dg[i]
            sumxi += xi[i] * xi[i]
        if fac * fac > EPS * sumdg * sumxi:
            fac = 1.0 / fac
            fad = 1.0 / fae
            for i in range(n):
                dg[i] = fac * xi[i] - fad * hdg[i]
            for i in range(n):
                for j in range(n):
                    hessin[i][j] += fac * xi[i] * xi[j] - fad * hdg[i] * hdg[j] + fae * dg[i] * dg[j]
        for i in range(n):
            xi[i] = 0.0
            for j in range(n):
                xi[i] -= hessin[i][j] * g[j]
    return fret
