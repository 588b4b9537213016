"""FP64 matrix-core traversal (20 and 61 states, ncat <= 4) against the vector-ALU form
and the oracle: logL |d|/|logL| <= 1e-10, partials within 1e-12 (20 states) / 1e-9 (61
states) of the pattern maximum, scale counts equal to the oracle's, NORM and SAFE,
forced rescaling and ragged pattern counts.  The 61-state bound: an eigen-space entry
U^-1 (a o b) of the codon model sums 61 terms of alternating sign that are up to ~1e5
times the result, so a different summation order (the matrix core's) moves it by
1e-11 to 2e-10 of the pattern maximum; the logL stays within 1e-10."""
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


@pytest.mark.parametrize("safe", [0, 1], ids=["norm", "safe"])
@pytest.mark.parametrize("n,T,sites,ncat", [(20, 24, 2000, 4), (20, 16, 900, 2), (61, 12, 600, 4), (61, 10, 300, 3)],
                         ids=["protein-c4", "protein-c2", "codon-c4", "codon-c3"])
def test_mfma_form_against_vector_form_and_oracle(n, T, sites, ncat, safe):
    rng = np.random.default_rng(n + T + safe)
    p = S.make_problem(n, T, sites, rng, ncat=ncat, mean_len=0.3, ambiguous=0.03)
    say("mfma", n, T, p.nptn, "ncat", ncat, "safe", safe)
    gm = G.Tree(p, scaling=safe)
    gm.set_f64_form(1)
    gv = G.Tree(p, scaling=safe)
    gv.set_f64_form(0)
    o = S.Oracle(p, safe=bool(safe), nthreads=4)
    ptol = 1e-12 if n == 20 else 1e-9
    for (u, v) in [(0, int(p.nei[0, 0])), (int(p.nei[T, 0]), T), (int(p.nei[T + 2, 1]), T + 2)]:
        lm, plm = gm.loglik(u, v, with_patterns=True)
        lv = gv.loglik(u, v)
        lo, plo = o.loglik(u, v, with_patterns=True)
        say("  branch", (u, v), "mfma-oracle", rel(lm, lo), "mfma-valu", rel(lm, lv))
        assert rel(lm, lo) <= 1e-10 and rel(lm, lv) <= 1e-10
        np.testing.assert_allclose(plm, plo, rtol=1e-10, atol=1e-10)
        dm, dv = gm.derivs(u, v)[:2], gv.derivs(u, v)[:2]
        do = o.derivs(u, v)[:2]
        for i in range(2):
            assert abs(dm[i] - do[i]) <= 1e-9 * max(1.0, abs(do[i]))
    for w in range(T, 2 * T - 2, 3):
        d = int(p.nei[w, 0])
        pm, sm = gm.partial(d, w, safe=bool(safe))
        po, so = o.partial(d, w, safe=bool(safe))
        assert np.array_equal(sm, so)
        scale = np.abs(po).max(axis=(1, 2), keepdims=True)
        worst = float((np.abs(pm - po) / (scale + 1e-300)).max())
        assert worst <= ptol, (w, worst)


@pytest.mark.parametrize("safe", [0, 1], ids=["norm", "safe"])
def test_mfma_forced_scaling_protein(safe):
    """A 100-taxon protein caterpillar (2^-256 rescaling on most patterns, repeatedly
    along the path): scale counts equal to the oracle's, logL and derivatives."""
    T = 100
    rng = np.random.default_rng(12)
    p = S.make_problem(20, T, 120, rng, tree=S.caterpillar_tree(T, 1.5), ambiguous=0.05)
    g = G.Tree(p, scaling=safe)
    g.set_f64_form(1)
    o = S.Oracle(p, safe=bool(safe))
    a, b = 0, int(p.nei[0, 0])
    assert rel(g.loglik(a, b), o.loglik(a, b)) <= 1e-10
    top = 0
    for w in range(T, 2 * T - 2, 9):
        d = int(p.nei[w, 0])
        _, sg = g.partial(d, w, safe=bool(safe))
        _, so = o.partial(d, w, safe=bool(safe))
        assert np.array_equal(sg, so)
        top = max(top, int(so.max()))
    assert top >= 2
    dg, do = g.derivs(a, b)[:2], o.derivs(a, b)[:2]
    for i in range(2):
        assert abs(dg[i] - do[i]) <= 1e-9 * max(1.0, abs(do[i]))


@pytest.mark.parametrize("sites", [1, 15, 16, 17, 63, 64, 65, 130])
def test_mfma_ragged_patterns(sites):
    rng = np.random.default_rng(sites)
    p = S.make_problem(20, 9, sites, rng, compress=False)
    g = G.Tree(p)
    g.set_f64_form(1)
    o = S.Oracle(p)
    a, b = 0, int(p.nei[0, 0])
    assert rel(g.loglik(a, b), o.loglik(a, b)) <= 1e-10
