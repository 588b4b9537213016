"""The oracle's model-parameter loop, pinned without a GPU: discrete-Gamma rates against
scipy's exact category means and Yang (1994)'s table, Brent's method against analytic
minima and scipy, the shape optimum against a fine grid of oracle logL(alpha), and the
branch/shape alternation against the stationarity it must end at."""
import ctypes
import json
import os

import numpy as np
import pytest
import scipy.optimize
import scipy.stats

import gems_amd as G
import support as S

HERE = os.path.dirname(os.path.abspath(__file__))
FIX = json.load(open(os.path.join(HERE, "golden", "model_fixtures.json")))


def orc_rates(alpha, k):
    r = np.zeros(k)
    S.oracle_lib().orc_gamma_rates(alpha, k, S._dp(r))
    return r


@pytest.mark.parametrize("case", FIX["gamma"], ids=lambda c: f"a{c['alpha']}k{c['ncat']}")
def test_gamma_rates_vs_scipy(case):
    r = orc_rates(case["alpha"], case["ncat"])
    # the published algorithms stop at ~1e-6 relative accuracy
    np.testing.assert_allclose(r, case["rates"], rtol=2e-5, atol=1e-9)
    assert abs(r.mean() - 1.0) < 1e-12


def test_gamma_rates_yang_table():
    # Yang (1994), J. Mol. Evol. 39:306, alpha = 0.5, four categories, mean rates
    np.testing.assert_allclose(orc_rates(0.5, 4), [0.0334, 0.2519, 0.8203, 2.8944], atol=6e-5)


@pytest.mark.parametrize("alpha", [0.02, 0.11, 0.5, 1.7, 7.5, 99.0, 1000.0])
@pytest.mark.parametrize("k", [2, 4, 8])
def test_gamma_rates_same_as_library_host(alpha, k):
    r = np.zeros(k)
    assert G.lib().plk_gamma_rates(alpha, k, r.ctypes.data_as(ctypes.POINTER(ctypes.c_double))) == 0
    # the same published algorithms, written apart: rounding-level differences, which the
    # cut-point iterations amplify to ~1e-11 at the largest shapes
    np.testing.assert_allclose(orc_rates(alpha, k), r, rtol=1e-10, atol=1e-300)


@pytest.mark.parametrize("alpha", [0.05, 0.2, 0.5, 1.0, 3.0, 20.0, 200.0])
@pytest.mark.parametrize("k", [2, 4, 6, 8])
def test_median_rates_vs_scipy_and_library(alpha, k):
    """cut_median rates: Gamma(alpha, 1/alpha) quantiles at (2i+1)/(2k), mean 1; the
    published chi-square routine is accurate to ~1e-6 relative."""
    o = np.zeros(k)
    S.oracle_lib().orc_gamma_rates_median(alpha, k, S._dp(o))
    q = scipy.stats.gamma.ppf((2 * np.arange(k) + 1) / (2.0 * k), a=alpha, scale=1.0 / alpha)
    np.testing.assert_allclose(o, q * k / q.sum(), rtol=2e-5)
    assert abs(o.mean() - 1.0) < 1e-12 and np.all(np.diff(o) > 0)
    r = np.zeros(k)
    assert G.lib().plk_gamma_rates_median(alpha, k, r.ctypes.data_as(ctypes.POINTER(ctypes.c_double))) == 0
    np.testing.assert_allclose(o, r, rtol=1e-10)


@pytest.mark.parametrize("fn,xmin,guess,xmax,xstar", [
    (lambda x: (x - 2.0) ** 2 + 1.0, 0.02, 1.0, 1000.0, 2.0),
    (lambda x: x - np.log(x), 0.02, 5.0, 1000.0, 1.0),
    (lambda x: (np.log(x) - np.log(30.0)) ** 2, 0.02, 31.0, 1000.0, 30.0),
    (lambda x: x, 0.5, 1.0, 3.0, 0.5),             # minimum on the lower bound
    (lambda x: -x, 0.5, 1.0, 3.0, 3.0),            # on the upper bound
    (lambda x: abs(x - 1.0), 0.02, 1.0, 1000.0, 1.0),  # at the starting point
])
def test_brent_known_minima(fn, xmin, guess, xmax, xstar):
    tol = 1e-3
    x, fx, _f2, calls = S.minimize_one_dimen(fn, xmin, guess, xmax, tol)
    assert abs(x - xstar) <= 4 * (tol * abs(xstar) + 1e-10) + 1e-9
    assert fx == pytest.approx(fn(x), abs=1e-15)
    # the search evaluates the guess first, then its bracket guess -/+ 50 tol guess
    eps = guess * tol * 50
    assert calls[0] == guess
    assert calls[1] == pytest.approx(max(guess - eps, xmin)) and calls[2] == pytest.approx(min(guess + eps, xmax))
    ref = scipy.optimize.minimize_scalar(fn, bounds=(xmin, xmax), method="bounded", options={"xatol": 1e-9})
    assert abs(x - float(ref.x)) <= 4 * (tol * abs(xstar) + 1e-10) + 1e-8


def test_brent_curvature_of_a_parabola():
    x, fx, f2, _ = S.minimize_one_dimen(lambda x: 3.0 * (x - 4.0) ** 2 - 7.0, 0.1, 1.0, 50.0, 1e-4)
    assert x == pytest.approx(4.0, rel=1e-3) and fx == pytest.approx(-7.0, abs=1e-5)
    assert f2 == pytest.approx(6.0, rel=1e-6)  # three points on a parabola give its exact curvature


@pytest.fixture(scope="module")
def dna_problem():
    # sequences simulated down the tree under Gamma(alpha = 0.6), so logL(alpha) has an
    # interior maximum
    return G.Problem(G.SEQ_DNA, 10, 3000, mean_len=0.15, alpha=0.6, ncat=4, seed=11)


def oracle_logl_at(o, alpha):
    o.set_gamma_shape(alpha)
    return o.compute_likelihood()


def test_gamma_optimum_matches_grid_and_scipy(dna_problem):
    o = S.Oracle(dna_problem, nthreads=2)
    o.set_gamma_shape(1.0)
    start = o.compute_likelihood()
    alpha, lnl = o.optimize_gamma(gradient_epsilon=1e-4)
    assert lnl >= start
    assert lnl == pytest.approx(oracle_logl_at(o, alpha), abs=1e-9)
    ref = scipy.optimize.minimize_scalar(lambda a: -oracle_logl_at(o, a), bounds=(0.02, 1000.0), method="bounded",
                                         options={"xatol": 1e-7})
    assert alpha == pytest.approx(ref.x, rel=3e-3)
    grid = ref.x * np.linspace(0.97, 1.03, 61)
    best = max(oracle_logl_at(o, a) for a in grid)
    assert lnl >= best - 1e-4
    o.close()


def test_gamma_optimum_independent_of_start(dna_problem):
    o = S.Oracle(dna_problem, nthreads=2)
    res = []
    for start in (0.05, 1.0, 40.0):
        o.set_gamma_shape(start)
        res.append(o.optimize_gamma(gradient_epsilon=1e-4))
    for a, l in res[1:]:
        assert a == pytest.approx(res[0][0], rel=5e-3)
        assert l == pytest.approx(res[0][1], abs=1e-3)
    o.close()


def test_parameter_loop_ends_stationary(dna_problem):
    o = S.Oracle(dna_problem, nthreads=2)
    o.set_gamma_shape(2.0)
    start = o.compute_likelihood()
    only_branches = S.Oracle(dna_problem, nthreads=2)
    only_branches.set_gamma_shape(2.0)
    lb = only_branches.optimize_all(iterations=100, tol=0.01)
    alpha, lnl = o.optimize_parameters(logl_epsilon=0.01, gradient_epsilon=1e-4)
    assert lnl >= start and lnl >= lb - 1e-6
    # the shape is optimal for the final lengths ...
    a2, l2 = o.optimize_gamma(gradient_epsilon=1e-4)
    assert a2 == pytest.approx(alpha, rel=5e-3) and l2 == pytest.approx(lnl, abs=0.02)
    # ... and no branch has a Newton step left that would gain more than the loop's epsilon
    # (df^2 / 2|ddf|), unless it is pressed against the lower length bound
    p = dna_problem
    for u in range(p.ntaxa, 2 * p.ntaxa - 2):
        for v in p.nei[u]:
            df, ddf, st = o.derivs(u, int(v))
            assert st == 0 and ddf < 0
            t = o.get_length(u, int(v))
            assert df * df / (2 * abs(ddf)) < 0.01 or (df < 0 and t < 1e-4)
    only_branches.close()
    o.close()


# ---- BFGS (ModelMarkov::optimizeParameters) -----------------------------------------

def test_bfgs_quadratic_bowl_and_scipy():
    a = np.array([1.0, 4.0, 0.5, 2.0, 8.0])
    c = np.array([0.7, 2.0, 5.0, 1.3, 0.2])
    fn = lambda x: float(np.sum(a * (np.asarray(x) - c) ** 2) + 3.0)  # noqa: E731
    x = [1.0] * 5
    fret = S.optim_oracle().dfpmin(fn, x, [1e-4] * 5, [100.0] * 5, 1e-6)
    # forward differences with h = 1e-4 |x| shift the stationary point by ~h / 2
    np.testing.assert_allclose(x, c, rtol=2e-3, atol=1e-4)
    ref = scipy.optimize.minimize(fn, np.ones(5), method="BFGS")
    assert fret == pytest.approx(fn(x)) and fret <= ref.fun + 1e-6


def test_bfgs_rosenbrock_valley():
    fn = lambda x: 100.0 * (x[1] - x[0] ** 2) ** 2 + (1.0 - x[0]) ** 2  # noqa: E731
    x = [0.5, 1.5]
    fret = S.optim_oracle().dfpmin(fn, x, [-10.0, -10.0], [10.0, 10.0], 1e-8)
    # forward-difference gradients stop the search along the curved valley floor a few
    # percent short of (1, 1)
    np.testing.assert_allclose(x, [1.0, 1.0], atol=0.06)
    # (fret can be the value of the last rejected trial point, as in the reference's lnsrch
    # when the step underflows)
    assert fn(x) < 1e-3 and fret < 1e-2


def test_bfgs_minimum_on_a_bound():
    fn = lambda x: float(sum((v + 1.0) ** 2 for v in x))  # noqa: E731
    x = [2.0, 3.0, 0.5]
    S.optim_oracle().dfpmin(fn, x, [0.0] * 3, [10.0] * 3, 1e-8)
    np.testing.assert_allclose(x, [0.0, 0.0, 0.0], atol=1e-6)


def test_bfgs_first_evaluations_are_forward_differences():
    calls = []

    def fn(x):
        calls.append(list(x))
        return float(sum(v * v for v in x))

    x = [2.0, 0.0, 5.0]
    S.optim_oracle().dfpmin(fn, x, [-10.0] * 3, [10.0] * 3, 1e-8)
    assert calls[0] == [2.0, 0.0, 5.0]
    assert calls[1] == [2.0 + 2e-4, 0.0, 5.0]
    assert calls[2] == [2.0, 1e-4, 5.0]      # h = ERROR_X when x_d == 0
    assert calls[3] == [2.0, 0.0, 5.0 + 5e-4]


def true_exchangeabilities(p):
    """Relative exchangeabilities of the simulating Q (Q_ij = r_ij pi_j s), GT = 1."""
    Q, pi = p.Q, p.pi
    e = np.array([Q[i, j] / pi[j] for i in range(4) for j in range(i + 1, 4)])
    return e / e[-1]


def test_gtr_search_recovers_the_simulating_rates():
    p = G.Problem(G.SEQ_DNA, 12, 6000, mean_len=0.2, alpha=0.8, ncat=4, seed=17)
    o = S.Oracle(p, nthreads=2)
    o.set_gamma_shape(0.8)
    o.set_gtr(np.ones(6), p.pi)
    start = o.compute_likelihood()
    ex, lnl = o.optimize_model(gradient_epsilon=1e-4)
    assert lnl > start
    np.testing.assert_allclose(ex / ex[-1], true_exchangeabilities(p), rtol=0.25)
    # a stationary point: scipy's bounded quasi-Newton from there gains next to nothing
    f = lambda x: o.gtr_objective(list(x))  # noqa: E731
    ref = scipy.optimize.minimize(f, ex[:5], method="L-BFGS-B", bounds=[(1e-4, 100.0)] * 5)
    assert -ref.fun <= lnl + 0.05
    o.close()
