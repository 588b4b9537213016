"""+I on the GPU: non-zero ptn_invar through every kernel family that reads it, against
the oracle (FP64: |dlogL|/|logL| <= 1e-10, scale counts bit-exact; FP32 form: 1e-6).
The FP32 form renormalises every pattern, so its sum with the (unscaled) invariant
mass is formed in log space (logL) and in the true frame (theta, derivatives)."""
import sys

import numpy as np
import pytest

import gems_amd as G
import support as S

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel(a, b):
    return abs(a - b) / max(abs(b), 1e-300)


def invar_problem(n, T, sites, seed, tree=None, p_inv=0.3, frac=0.3, mean_len=0.2):
    rng = np.random.default_rng(seed)
    p = S.make_problem(n, T, sites, rng, tree=tree, ambiguous=0.03, mean_len=mean_len)
    S.plant_constant_columns(p, frac, rng)
    return S.apply_pinvar(p, p_inv)


CASES = [(4, 32, 4000, 0), (4, 32, 4000, 1), (20, 24, 2000, 0), (20, 24, 2000, 1), (61, 12, 400, 1)]
IDS = ["dna-norm", "dna-safe", "protein-norm", "protein-safe", "codon-safe"]


@pytest.mark.parametrize("precision", [G.FP64, G.FP32], ids=["fp64", "fp32"])
@pytest.mark.parametrize("case", CASES, ids=IDS)
def test_invar_loglik_and_derivs(case, precision):
    n, T, sites, safe = case
    p = invar_problem(n, T, sites, seed=n + T + safe)
    say("invar", case, "precision", precision, "nptn", p.nptn, "invariant patterns", int((p.invar > 0).sum()))
    g = G.Tree(p, precision=precision, scaling=safe)
    o = S.Oracle(p, safe=bool(safe), nthreads=4)
    tol = 1e-10 if precision == G.FP64 else 1e-6
    a, b = 0, int(p.nei[0, 0])
    for (u, v) in [(a, b), (int(p.nei[T, 0]), T), (int(p.nei[T + 3, 1]), T + 3)]:
        lg, plg = g.loglik(u, v, with_patterns=True)
        lo, plo = o.loglik(u, v, with_patterns=True)
        say("  branch", (u, v), "logL", lg, lo, "rel", rel(lg, lo))
        assert rel(lg, lo) <= tol, (u, v, lg, lo)
        np.testing.assert_allclose(plg, plo, rtol=tol, atol=tol)
        dfg, ddfg, _ = g.derivs(u, v)
        dfo, ddfo, _ = o.derivs(u, v)
        dtol = 1e-9 if precision == G.FP64 else 2e-4
        say("  df", dfg, dfo, "ddf", ddfg, ddfo)
        assert abs(dfg - dfo) <= dtol * max(1.0, abs(dfo))
        assert abs(ddfg - ddfo) <= dtol * max(1.0, abs(ddfo))
        assert rel(g.loglik_from_theta(u, v), lo) <= tol
    # the invariant term matters here (the comparison is not vacuous)
    q = invar_problem(n, T, sites, seed=n + T + safe)
    q.invar = None
    assert abs(S.Oracle(q, safe=bool(safe)).loglik(a, b) - o.loglik(a, b)) > 1.0


@pytest.mark.parametrize("precision,safe", [(G.FP64, 0), (G.FP64, 1), (G.FP32, 0)], ids=["fp64-norm", "fp64-safe", "fp32"])
def test_invar_forced_scaling(precision, safe):
    """A 300-taxon caterpillar (rescaling forced on most patterns) with constant columns:
    FP64 scale counts equal the oracle's and are zero on every pattern with invariant
    mass; logL, per-pattern logL and derivatives match the oracle."""
    T = 300
    p = invar_problem(4, T, 150, seed=7, tree=S.caterpillar_tree(T, 1.5))
    say("invar caterpillar precision", precision, "safe", safe)
    g = G.Tree(p, precision=precision, scaling=safe)
    o = S.Oracle(p, safe=bool(safe))
    tol = 1e-10 if precision == G.FP64 else 1e-6
    a, b = 0, int(p.nei[0, 0])
    lg, plg = g.loglik(a, b, with_patterns=True)
    lo, plo = o.loglik(a, b, with_patterns=True)
    assert rel(lg, lo) <= tol
    np.testing.assert_allclose(plg, plo, rtol=tol, atol=1e-9)
    dfg, ddfg, _ = g.derivs(a, b)
    dfo, ddfo, _ = o.derivs(a, b)
    dtol = 1e-9 if precision == G.FP64 else 2e-4
    assert abs(dfg - dfo) <= dtol * max(1.0, abs(dfo))
    assert abs(ddfg - ddfo) <= dtol * max(1.0, abs(ddfo))
    if precision == G.FP64:
        inv = p.invar > 0
        for v in range(T, 2 * T - 2, 23):
            d = int(p.nei[v, 0])
            _, sg = g.partial(d, v, safe=bool(safe))
            _, so = o.partial(d, v, safe=bool(safe))
            assert np.array_equal(sg, so)
            assert not sg.reshape(p.nptn, -1)[inv].any()
        _, so = o.partial(a, b, safe=bool(safe))
        assert so.max() >= 1


@pytest.mark.parametrize("precision", [G.FP64, G.FP32], ids=["fp64", "fp32"])
def test_invar_newton(precision):
    """Device-resident Newton with invariant mass reaches the oracle's optimum."""
    p = invar_problem(20, 16, 1500, seed=4)
    g = G.Tree(p, precision=precision)
    o = S.Oracle(p)
    u, v = int(p.nei[20, 0]), 20
    tg = g.optimize_one(u, v)
    to = o.optimize_one(u, v)
    say("newton", precision, tg, to)
    assert abs(tg - to) <= (1e-6 if precision == G.FP64 else 1e-4) * max(1.0, to)
    lg, lo = g.loglik(u, v), o.loglik(u, v)
    assert rel(lg, lo) <= (1e-10 if precision == G.FP64 else 1e-6)
