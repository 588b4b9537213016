"""GPU parity of the UFBoot RELL scoring (plk_rell_*, through the C ABI) against the
oracle's saveCurrentTree loop on the same per-pattern logL and the same uniform deviates:
scores to 1e-12 relative (FP64 sums in different fixed orders), replicate state
(best logL, tie counts, originating tree score, improved flags) identical."""
import sys

import numpy as np
import pytest

import gems_amd as G
import support as S

pytestmark = pytest.mark.gpu


def branches(p):
    out = []
    for u in range(2 * p.ntaxa - 2):
        for k in range(3):
            v = int(p.nei[u, k])
            if v > u:
                out.append((u, v))
    return out


@pytest.mark.parametrize("prec", [G.FP64, G.FP32], ids=["fp64", "fp32"])
def test_rell_matches_oracle(prec):
    p = G.Problem(G.SEQ_DNA, 16, 3000, seed=5)
    print("rell", prec, p.nptn, flush=True)
    rng = np.random.default_rng(11)
    nsites = int(round(p.freq.sum()))
    ns = 257
    # nonparametric bootstrap of the sites, as per-pattern counts (createBootstrapAlignment)
    counts = rng.multinomial(nsites, p.freq / p.freq.sum(), size=ns).astype(np.float32)
    g = G.Tree(p, precision=prec)
    dev = G.Rell(counts)
    orc = S.OracleRell(counts)
    br = branches(p)
    for k in range(6):
        u, v = br[k]
        if k:
            g.set_length(u, v, 0.02 + 0.05 * k)  # a different tree for every round
        logl, lp = g.loglik(u, v, with_patterns=True)
        rd = dev.eval(g.context())
        uni = rng.random(ns)
        ro, imp_o = orc.save_tree(lp, logl, eps=0.5, uniform=uni)
        np.testing.assert_allclose(rd, ro, rtol=1e-12, atol=0)
        imp_d, nimp = dev.save_tree(g.context(), logl, uni, eps=0.5)
        assert np.array_equal(imp_d, imp_o) and nimp == int(imp_o.sum())
        lo, cnt, orig = dev.state()
        assert np.array_equal(lo, orc.logl) and np.array_equal(cnt, orc.best_counts)
        assert np.array_equal(orig, orc.orig)
    # the original alignment as a replicate scores the tree logL
    r1 = G.Rell(p.freq.reshape(1, -1).astype(np.float32))
    logl, lp = g.loglik(*br[0], with_patterns=True)
    assert abs(r1.eval(g.context())[0] - logl) <= float(np.sum(np.abs(p.freq * lp))) * 2.0 ** -24


def test_rell_rejects_mismatch():
    p = G.Problem(G.SEQ_DNA, 8, 300, seed=2)
    g = G.Tree(p)
    r = G.Rell(np.ones((4, p.nptn + 1), dtype=np.float32))
    with pytest.raises(G.PlkError):
        r.eval(g.context())  # pattern counts differ
    r2 = G.Rell(np.ones((4, p.nptn), dtype=np.float32))
    with pytest.raises(G.PlkError):
        r2.eval(g.context())  # no logL evaluated yet on this context
    g.full_traversal()
    assert r2.eval(g.context()).shape == (4,)
