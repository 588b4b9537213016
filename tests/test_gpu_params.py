"""Model-parameter loop through the C ABI: the Gamma-shape search
(plk_tree_optimize_gamma, RateGamma::optimizeParameters) and the branch/shape
alternation (plk_tree_optimize_parameters, ModelFactory::optimizeParameters), every
evaluation a device traversal, against the oracle's restatement on the same inputs."""
import numpy as np
import pytest

import gems_amd as G
import support as S

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def rel(a, b):
    return abs(a - b) / max(abs(b), 1e-300)


CASES = [
    ("dna-f64", G.SEQ_DNA, 16, 4000, G.FP64, 0),
    ("dna-f64-safe", G.SEQ_DNA, 16, 4000, G.FP64, 1),
    ("protein-f64", G.SEQ_PROTEIN, 12, 1500, G.FP64, 0),
    ("dna-f32", G.SEQ_DNA, 16, 4000, G.FP32, 0),
    ("protein-f32", G.SEQ_PROTEIN, 12, 1500, G.FP32, 0),
]


def logl_tol(prec):
    return 1e-10 if prec == G.FP64 else 2e-6


def same_tol(prec):
    """Two device evaluations of one logL (fused into the traversal or from cached partials)."""
    return 1e-13 if prec == G.FP64 else 1e-7


# The search's stopping test compares bracket widths built from steps of exactly
# tol * |x|, so rounding-level differences in the logL (or in the starting shape) can
# decide whether it takes one more step; two correct runs then stop at different points
# of the final bracket.  Parity is therefore checked in parts: logL(alpha) against the
# oracle to the kernels' tolerance, the search itself step for step against the oracle's
# Brent driven by the device's own logL, and the end points to the search tolerance.
ALPHA_END = 3e-3  # relative: two end points of one final bracket (2 tol, tol = 1e-3)


@pytest.mark.parametrize("name,seq,T,sites,prec,safe", CASES, ids=[c[0] for c in CASES])
def test_optimize_gamma_matches_oracle(name, seq, T, sites, prec, safe):
    p = G.Problem(seq, T, sites, mean_len=0.15, alpha=0.7, seed=31)
    say("optimize_gamma", name, "nptn", p.nptn)
    g = G.Tree(p, precision=prec, scaling=safe)
    o = S.Oracle(p, safe=bool(safe), nthreads=4)

    def f_dev(x):
        g.set_gamma_shape(x)
        return -g.compute_likelihood()

    for start in (1.0, 0.1):
        g.set_gamma_shape(start)
        o.set_gamma_shape(start)
        ag, lg = g.optimize_gamma(gradient_epsilon=1e-4)
        ao, lo = o.optimize_gamma(gradient_epsilon=1e-4)
        say(f"  start {start}: alpha gpu {ag:.10f} oracle {ao:.10f}  logL gpu {lg:.10f} oracle {lo:.10f}")
        assert g.gamma_shape() == ag
        # the model is left at the optimum: a fresh likelihood agrees
        assert rel(g.full_traversal(), lg) <= same_tol(prec)
        # logL(alpha) on the device = the oracle's, at both end points
        for a in (ag, ao):
            o.set_gamma_shape(a)
            assert rel(f_dev(a), -o.compute_likelihood()) <= logl_tol(prec)
        # the device search = the oracle's Brent on the device's logL, step for step (the
        # oracle is built with fused multiply-adds, so its step arithmetic may round 1 ulp
        # apart)
        xb, fb, _f2, calls = S.minimize_one_dimen(f_dev, 0.02, start, 1000.0, 1e-3)
        xb_final = -f_dev(xb)
        say(f"  oracle Brent on device logL: {xb:.12f} after {len(calls)} evaluations")
        assert rel(xb, ag) <= 1e-12 and rel(xb_final, lg) <= same_tol(prec)
        # end points
        assert rel(ag, ao) <= ALPHA_END and abs(lg - lo) <= 1e-3
    g.close()
    o.close()


@pytest.mark.parametrize("name,seq,T,sites,prec,safe", CASES[::2], ids=[c[0] for c in CASES[::2]])
def test_optimize_parameters_matches_oracle(name, seq, T, sites, prec, safe):
    p = G.Problem(seq, T, sites, mean_len=0.15, alpha=0.7, seed=37)
    say("optimize_parameters", name, "nptn", p.nptn)
    g = G.Tree(p, precision=prec, scaling=safe)
    o = S.Oracle(p, safe=bool(safe), nthreads=4)
    g.set_gamma_shape(2.0)
    o.set_gamma_shape(2.0)
    start = g.compute_likelihood()
    ag, lg = g.optimize_parameters(logl_epsilon=0.01, gradient_epsilon=1e-4)
    ao, lo = o.optimize_parameters(logl_epsilon=0.01, gradient_epsilon=1e-4)
    say(f"  alpha gpu {ag:.10f} oracle {ao:.10f}  logL gpu {lg:.10f} oracle {lo:.10f}")
    assert lg >= start
    assert rel(ag, ao) <= 2 * ALPHA_END and abs(lg - lo) <= 0.01  # logl_epsilon
    # the alternation = ModelFactory::optimizeParameters' rule driven through the C ABI's
    # own pieces (same device arithmetic, so the same decisions)
    h = G.Tree(p, precision=prec, scaling=safe)
    h.set_gamma_shape(2.0)
    cur = h.compute_likelihood()
    for i in range(2, 100):
        h.optimize_all(iterations=min(i, 3), tol=0.01)
        _a, new = h.optimize_gamma(gradient_epsilon=1e-4)
        if new > cur + 0.01:
            cur = new
        else:
            cur = h.optimize_all(iterations=100, tol=0.01)
            break
    assert h.gamma_shape() == ag and rel(cur, lg) <= same_tol(prec)
    for u in range(p.ntaxa, 2 * p.ntaxa - 2):
        for v in p.nei[u]:
            assert g.get_length(u, int(v)) == h.get_length(u, int(v))
    h.close()
    g.close()
    o.close()


@pytest.mark.parametrize("prec", [G.FP64, G.FP32], ids=["f64", "f32"])
def test_optimize_gamma_median_rates(prec):
    """The cut_median rate model through the same search: logL(alpha) and the search
    against the oracle's median-rate restatement."""
    p = G.Problem(G.SEQ_DNA, 16, 4000, mean_len=0.15, alpha=0.7, seed=41)
    say("optimize_gamma median", prec, "nptn", p.nptn)
    g = G.Tree(p, precision=prec)
    o = S.Oracle(p, nthreads=4)
    g.set_gamma_median(True)
    o.set_gamma_median(True)
    g.set_gamma_shape(1.0)
    o.set_gamma_shape(1.0)
    assert rel(g.compute_likelihood(), o.compute_likelihood()) <= logl_tol(prec)
    ag, lg = g.optimize_gamma(gradient_epsilon=1e-4)
    ao, lo = o.optimize_gamma(gradient_epsilon=1e-4)
    say(f"  alpha gpu {ag:.10f} oracle {ao:.10f}  logL gpu {lg:.10f} oracle {lo:.10f}")

    def f_dev(x):
        g.set_gamma_shape(x)
        return -g.compute_likelihood()

    xb, _fb, _f2, _calls = S.minimize_one_dimen(f_dev, 0.02, 1.0, 1000.0, 1e-3)
    assert rel(xb, ag) <= 1e-12
    assert rel(ag, ao) <= ALPHA_END and abs(lg - lo) <= 1e-3
    # a different rate model: the mean-rate optimum differs
    g.set_gamma_median(False)
    g.set_gamma_shape(1.0)
    am, _ = g.optimize_gamma(gradient_epsilon=1e-4)
    assert abs(am - ag) > 1e-3 * ag
    g.close()
    o.close()


def test_gamma_needs_a_shape_and_model():
    p = G.Problem(G.SEQ_DNA, 6, 300, seed=3)
    g = G.Tree(p)
    with pytest.raises(G.PlkError):
        g.optimize_gamma()  # rates are the caller's own until a shape is set
    with pytest.raises(G.PlkError):
        g.set_gamma_shape(-1.0)
    g.set_gamma_shape(0.5)
    assert g.gamma_shape() == 0.5
    g.set_model(p)  # a new model clears the shape
    assert g.gamma_shape() == 0.0
    g.close()


@pytest.mark.parametrize("prec", [G.FP64, G.FP32], ids=["f64", "f32"])
def test_optimize_model_matches_oracle(prec):
    """Exchangeabilities by BFGS (plk_tree_optimize_model): the device search step for
    step against the oracle's dfpmin driven by the device's own logL, logL(exch) against the
    oracle, and the end points against the oracle's own search."""
    p = G.Problem(G.SEQ_DNA, 14, 5000, mean_len=0.2, alpha=0.8, seed=43)
    say("optimize_model", prec, "nptn", p.nptn)
    g = G.Tree(p, precision=prec)
    o = S.Oracle(p, nthreads=4)
    for t in (g, o):
        t.set_gamma_shape(0.8)
        t.set_gtr(np.ones(6), p.pi)
    assert rel(g.compute_likelihood(), o.compute_likelihood()) <= logl_tol(prec)
    eg, lg = g.optimize_model(gradient_epsilon=1e-4)
    eo, lo = o.optimize_model(gradient_epsilon=1e-4)
    say(f"  exch gpu {np.round(eg, 5)} oracle {np.round(eo, 5)}  logL gpu {lg:.8f} oracle {lo:.8f}")

    def f_dev(x):
        g.set_gtr(np.r_[x, 1.0], p.pi)
        return -g.compute_likelihood()

    x = [1.0] * 5
    g.set_gtr(np.ones(6), p.pi)
    S.optim_oracle().dfpmin(f_dev, x, [1e-4] * 5, [100.0] * 5, 1e-4)
    np.testing.assert_allclose(x, eg[:5], rtol=1e-9)
    o.set_gtr(eg, p.pi)
    assert rel(f_dev(list(eg[:5])), -o.compute_likelihood()) <= logl_tol(prec)
    if prec == G.FP64:
        np.testing.assert_allclose(eg, eo, rtol=1e-4)
        assert abs(lg - lo) <= 1e-4
    else:
        # FP32 logL carries ~1e-8 relative rounding, which forward differences with steps of
        # 1e-4 relative turn into gradient noise: the search stops a little short
        np.testing.assert_allclose(eg, eo, rtol=5e-2)
        assert lo - 0.5 <= lg <= lo + 1e-3
    g.close()
    o.close()


def test_optimize_parameters_with_model_matches_oracle():
    p = G.Problem(G.SEQ_DNA, 12, 3000, mean_len=0.2, alpha=0.8, seed=47)
    say("optimize_parameters with exchangeabilities", "nptn", p.nptn)
    g = G.Tree(p)
    o = S.Oracle(p, nthreads=4)
    for t in (g, o):
        t.set_gamma_shape(1.5)
        t.set_gtr(np.ones(6), p.pi)
    start = g.compute_likelihood()
    ag, lg = g.optimize_parameters(logl_epsilon=0.01, gradient_epsilon=1e-4)
    eo, lo = o.optimize_parameters_model(logl_epsilon=0.01, gradient_epsilon=1e-4)
    eg = g.gtr()
    say(f"  alpha gpu {ag:.6f} oracle {o.gamma_shape():.6f}  logL gpu {lg:.6f} oracle {lo:.6f}")
    assert lg > start
    assert abs(lg - lo) <= 0.01 and rel(ag, o.gamma_shape()) <= 2 * ALPHA_END
    np.testing.assert_allclose(eg, eo, rtol=3e-2)
    g.close()
    o.close()
