"""Parity at the measured workloads' full sizes (the seeded problems of
gems_amd/workloads.py), against the oracle.  Tolerances: FP64 |dlogL|/|logL| <= 1e-10 and
derivatives 1e-9 (relative, or absolute below 1); FP32 form 1e-6 on logL and about ten
times the measured error on derivatives (stated per test)."""
import sys

import numpy as np
import pytest

import gems_amd as G
from gems_amd import workloads as W
import support as S

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel(a, b):
    return abs(a - b) / max(abs(b), 1e-300)


def derr(g, o):
    return max(abs(g[i] - o[i]) / max(1.0, abs(o[i])) for i in range(2))


def test_config1_all_branch_derivatives():
    """DNA, 32 taxa x 10,000 sites: df/ddf of all 61 branches (one all-directions launch)
    against the oracle branch by branch.  FP32 bound 1e-4 (measured ~1.2e-5)."""
    p = W.config1()
    o = S.Oracle(p, nthreads=8)
    for prec, tol in ((G.FP64, 1e-9), (G.FP32, 1e-4)):
        g = G.Tree(p, precision=prec, all_directions=True)
        br, df, ddf = g.all_branch_derivs()
        assert len(br) == 2 * p.ntaxa - 3
        worst = 0.0
        for (u, v), a, b in zip(br, df, ddf):
            do = o.derivs(u, v)[:2]
            worst = max(worst, derr((a, b), do))
        say("config1 precision", prec, "branches", len(br), "max deriv err", worst)
        assert worst <= tol
        g.close()


@pytest.mark.parametrize("prec", [G.FP64, G.FP32], ids=["fp64", "fp32"])
def test_config2_full_size(prec):
    """Protein, 100 taxa x 50,000 sites: logL and per-pattern logL at two branches, df/ddf
    at both.  FP32 derivative bound 1e-5 (measured ~7e-7)."""
    p = W.config2()
    say("config2", prec, "patterns", p.nptn)
    o = S.Oracle(p, nthreads=16)
    g = G.Tree(p, precision=prec)
    tol, dtol = (1e-10, 1e-9) if prec == G.FP64 else (1e-6, 1e-5)
    for (u, v) in [(0, int(p.nei[0, 0])), (100, int(p.nei[100, 0]))]:
        lg, plg = g.loglik(u, v, with_patterns=True)
        lo, plo = o.loglik(u, v, with_patterns=True)
        say("  branch", (u, v), "rel", rel(lg, lo))
        assert rel(lg, lo) <= tol
        np.testing.assert_allclose(plg, plo, rtol=tol * 10, atol=tol * 10)
        e = derr(g.derivs(u, v)[:2], o.derivs(u, v)[:2])
        say("  deriv err", e)
        assert e <= dtol


@pytest.mark.parametrize("prec", [G.FP64, G.FP32], ids=["fp64", "fp32"])
def test_config3_full_size(prec):
    """300 unlinked partitions on 64 taxa in one launch per traversal: every partition's
    logL equals evaluating it alone (bit for bit) and the oracle's."""
    probs = W.config3()
    trees = [G.Tree(p, precision=prec) for p in probs]
    alone = [t.full_traversal() for t in trees]
    b = G.Batch(trees)
    br = [(0, int(p.nei[0, 0])) for p in probs]
    vals, tot = b.loglik(br, clear=True)
    assert np.array_equal(vals, np.array(alone))
    tol = 1e-10 if prec == G.FP64 else 1e-6
    worst = 0.0
    for p, v in zip(probs, vals):
        o = S.Oracle(p, nthreads=4)
        worst = max(worst, rel(v, o.loglik(0, int(p.nei[0, 0]))))
        o.close()
    say("config3", prec, "partitions", len(probs), "worst rel", worst)
    assert worst <= tol
    b.close()


@pytest.mark.parametrize("prec", [G.FP64, G.FP32], ids=["fp64", "fp32"])
def test_config4_full_size(prec):
    """Codon, 48 taxa x 5,000 codons, SAFE scaling in FP64 (the reference's rule for 61
    states): logL, per-pattern logL and df/ddf at two branches.  FP32 derivative bound 1e-5."""
    p = W.config4()
    say("config4", prec, "patterns", p.nptn)
    o = S.Oracle(p, safe=True, nthreads=16)
    g = G.Tree(p, precision=prec, scaling=G.SAFE)
    tol, dtol = (1e-10, 1e-9) if prec == G.FP64 else (1e-6, 1e-5)
    for (u, v) in [(0, int(p.nei[0, 0])), (48, int(p.nei[48, 1]))]:
        lg, plg = g.loglik(u, v, with_patterns=True)
        lo, plo = o.loglik(u, v, with_patterns=True)
        say("  branch", (u, v), "rel", rel(lg, lo))
        assert rel(lg, lo) <= tol
        np.testing.assert_allclose(plg, plo, rtol=tol * 10, atol=tol * 10)
        e = derr(g.derivs(u, v)[:2], o.derivs(u, v)[:2])
        say("  deriv err", e)
        assert e <= dtol
    if prec == G.FP64:
        _, sg = g.partial(0, int(p.nei[0, 0]), safe=True)
        _, so = o.partial(0, int(p.nei[0, 0]), safe=True)
        assert np.array_equal(sg, so)


def test_config5_full_size_requests():
    """Serving at the largest request sizes: a 512-taxon x 100,000-site DNA request (Newton
    on one branch) and a 512-taxon x 100,000-site protein request (logL), among smaller
    ones, give the same results alone and in the mix, and their logL matches the oracle."""
    specs, _ = W.config5_specs(6)
    big = [(G.SEQ_DNA, 512, 100000, 1, 9001), (G.SEQ_PROTEIN, 512, 100000, 0, 9002)]
    probs = [W.config5_problem(sp) for sp in big + specs]
    kinds = [sp[3] for sp in big + specs]
    say("config5 patterns", [p.nptn for p in probs[:2]])

    def requests():
        return [G.Request(p, kind=G.REQ_LOGL if k == 0 else G.REQ_OPT_BRANCH, precision=G.FP64,
                          opt_branch=(p.ntaxa, int(p.nei[p.ntaxa, 0]))) for p, k in zip(probs, kinds)]

    alone = requests()
    s1 = G.Server(nworkers=1)
    for r in alone:
        s1.submit(r)
        s1.drain()
    s1.close()
    mixed = requests()
    s = G.Server(nworkers=8)
    for r in mixed:
        s.submit(r)
    s.drain()
    st, cn = s.stats(), s.counters()
    say("mixed", st, cn)
    s.close()
    assert st["failed"] == 0 and cn["batched"] > cn["launches"]
    for a, m in zip(alone, mixed):
        ra, rm = a.result(), m.result()
        assert ra["status"] == 0 and rm["status"] == 0
        assert (ra["logl_before"], ra["logl_after"], ra["opt_len"]) == (rm["logl_before"], rm["logl_after"], rm["opt_len"])
    for p, r in zip(probs[:2], mixed[:2]):
        o = S.Oracle(p, nthreads=16)
        lo = o.loglik(0, int(p.nei[0, 0]))
        say("  oracle", p.n, lo, r.result()["logl_before"], rel(r.result()["logl_before"], lo))
        assert rel(r.result()["logl_before"], lo) <= 1e-10
        o.close()
