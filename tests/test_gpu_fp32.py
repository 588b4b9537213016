"""FP32 form (FP32 partials with per-pattern power-of-two exponents, matrix cores for
20/61 states) against the FP64 CPU oracle.  Target |dlogL|/|logL| <= 1e-6; derivative
errors are asserted loosely and printed."""
import sys

import numpy as np
import pytest

import gems_amd as G
import support as S

pytestmark = pytest.mark.gpu

LOGL_TOL = 1e-6


def say(*a):
    print(*a, flush=True)


def branches(p, limit=None):
    out = []
    for u in range(2 * p.ntaxa - 2):
        for k in range(3):
            v = int(p.nei[u, k])
            if v > u:
                out.append((u, v))
    return out[:limit] if limit else out


def rel(a, b):
    return abs(a - b) / max(abs(b), 1e-300)


CASES = [
    (G.SEQ_DNA, 32, 10000, 0.1, 0.5),
    (G.SEQ_PROTEIN, 40, 5000, 0.2, 0.8),
    (G.SEQ_CODON, 12, 600, 0.1, 0.5),
]


@pytest.mark.parametrize("case", CASES, ids=["dna", "protein", "codon"])
def test_fp32_loglik(case):
    seq, T, sites, ml, alpha = case
    say("fp32 loglik", case)
    p = G.Problem(seq, T, sites, mean_len=ml, alpha=alpha, seed=21)
    g = G.Tree(p, precision=G.FP32)
    o = S.Oracle(p, safe=(seq == G.SEQ_CODON), nthreads=4)
    worst = 0.0
    for (u, v) in branches(p, 5) + [(0, int(p.nei[0, 0]))]:
        lg, plg = g.loglik(u, v, with_patterns=True)
        lo, plo = o.loglik(u, v, with_patterns=True)
        worst = max(worst, rel(lg, lo))
        assert rel(lg, lo) <= LOGL_TOL, (u, v, lg, lo)
        assert np.max(np.abs(plg - plo)) < 1e-3
    say(f"  worst |dlogL|/|logL| = {worst:.2e}")


@pytest.mark.parametrize("case", CASES, ids=["dna", "protein", "codon"])
def test_fp32_derivatives(case):
    seq, T, sites, ml, alpha = case
    say("fp32 derivs", case)
    p = G.Problem(seq, T, min(sites, 4000), mean_len=ml, alpha=alpha, seed=4)
    g = G.Tree(p, precision=G.FP32)
    o = S.Oracle(p, safe=(seq == G.SEQ_CODON))
    for (u, v) in branches(p, 4):
        dg, ddg, rc = g.derivs(u, v)
        assert rc == 0
        do, ddo, _ = o.derivs(u, v)
        e1 = abs(dg - do) / max(1.0, abs(do))
        e2 = abs(ddg - ddo) / max(1.0, abs(ddo))
        say(f"  branch {u}-{v}: df {dg:.6f} vs {do:.6f} ({e1:.1e}), ddf {ddg:.4f} vs {ddo:.4f} ({e2:.1e})")
        # FP32 partials (7 digits) carried into an FP64 derivative sum: measured up to
        # ~1.2e-5 relative (DNA, 32 taxa x 10k sites), ~7e-7 (protein); bound ~10x that
        assert e1 < 2e-4 and e2 < 2e-4
        assert rel(g.loglik_from_theta(u, v), o.loglik_from_theta(u, v)) <= LOGL_TOL


def test_fp32_partials_real_space():
    """FP32 partial * 2^e equals U . (oracle eigen-space partial) * 2^(256 s)."""
    p = G.Problem(G.SEQ_PROTEIN, 20, 800, mean_len=0.3, seed=8)
    g = G.Tree(p, precision=G.FP32)
    o = S.Oracle(p)
    a, b = 0, int(p.nei[0, 0])
    g.loglik(a, b)
    import ctypes
    for v in range(p.ntaxa, 2 * p.ntaxa - 2, 5):
        dad = int(p.nei[v, 0])
        pg, _ = np.zeros(1), None
        out = np.zeros(p.nptn * p.n * p.ncat)
        # use the tree layer to orient and fetch; exponents via the context
        G._check(G.lib().plk_tree_get_partial(g.h, dad, v, G._d(out), None))
        ex = np.zeros(p.nptn, dtype=np.int32)
        G._check(G.lib().plk_get_exponents(g.context(), v - p.ntaxa, ex.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))))
        real_g = out.reshape(p.nptn, p.ncat, p.n) * np.ldexp(1.0, ex)[:, None, None]
        po, so = o.partial(dad, v)
        real_o = np.einsum("xi,pci->pcx", p.evec, po) * np.ldexp(1.0, 256 * so.astype(np.int64))[:, None, None]
        scale = np.abs(real_o).max(axis=(1, 2), keepdims=True)
        assert np.all(np.abs(real_g - real_o) <= 2e-5 * scale)


def test_fp32_deep_tree_exponents():
    T = 400
    rng = np.random.default_rng(6)
    pp = S.make_problem(4, T, 120, rng, tree=S.caterpillar_tree(T, 2.0))
    say("fp32 deep caterpillar")
    g = G.Tree(pp, precision=G.FP32)
    o = S.Oracle(pp)
    a, b = 0, int(pp.nei[0, 0])
    assert rel(g.loglik(a, b), o.loglik(a, b)) <= LOGL_TOL


def test_fp32_newton():
    p = G.Problem(G.SEQ_PROTEIN, 12, 2000, mean_len=0.2, seed=3)
    g = G.Tree(p, precision=G.FP32)
    o = S.Oracle(p)
    lg = g.optimize_all(iterations=5, tol=1e-4)
    lo = o.optimize_all(iterations=5, tol=1e-4)
    say(f"fp32 newton logL {lg} vs {lo}")
    assert rel(lg, lo) <= 1e-6
    for (u, v) in branches(p):
        assert abs(g.get_length(u, v) - o.get_length(u, v)) <= 1e-3 * max(1e-2, o.get_length(u, v))


def test_fp32_ragged():
    for nsites in (1, 63, 65):
        rng = np.random.default_rng(nsites)
        pp = S.make_problem(20, 7, nsites, rng, ambiguous=0.2, compress=False)
        g = G.Tree(pp, precision=G.FP32)
        o = S.Oracle(pp)
        for (u, v) in branches(pp, 3):
            assert rel(g.loglik(u, v), o.loglik(u, v)) <= LOGL_TOL


def test_fp32_tile_shapes_identical():
    p = G.Problem(G.SEQ_PROTEIN, 30, 3000, mean_len=0.2, seed=12)
    g = G.Tree(p, precision=G.FP32)
    vals = []
    for w in (0, 1, 2, 4):
        g.set_tile_shape(w)
        vals.append(g.full_traversal())
    assert vals[0] == vals[1] == vals[2] == vals[3], vals
    best, ms = g.autotune(reps=2)
    say("autotuned waves/category", best, "ms", ms)
    assert g.full_traversal() == vals[0]
    # codon: pattern-block shape against the category-wave shape
    pc = G.Problem(G.SEQ_CODON, 10, 400, seed=2)
    gc = G.Tree(pc, precision=G.FP32)
    gc.set_tile_shape(1)
    a = gc.full_traversal()
    gc.set_tile_shape(0)
    assert gc.full_traversal() == a


@pytest.mark.parametrize("seq,ncat", [(G.SEQ_CODON, 6), (G.SEQ_PROTEIN, 8), (G.SEQ_DNA, 8)], ids=["codon6", "protein8", "dna8"])
def test_fp32_many_categories(seq, ncat):
    """Five to eight rate categories (the wide workgroup shapes) against the oracle."""
    say("fp32 many categories", seq, ncat)
    p = G.Problem(seq, 10, 400, mean_len=0.15, alpha=0.6, ncat=ncat, seed=29)
    g = G.Tree(p, precision=G.FP32)
    o = S.Oracle(p, safe=(seq == G.SEQ_CODON), nthreads=4)
    for (u, v) in branches(p, 3) + [(0, int(p.nei[0, 0]))]:
        lg, lo = g.loglik(u, v), o.loglik(u, v)
        assert rel(lg, lo) <= LOGL_TOL, (u, v, lg, lo)
    g.close()
    o.close()


@pytest.mark.parametrize("sites", [16600, 16640], ids=["257tiles", "258tiles"])
def test_fp32_split_units(sites):
    """Tile counts just above the 256 compute units: the last tiles run as single-block
    units.  Per-pattern logL is the same as with whole tiles (another tile shape), the total
    differs only by summation order, and both match the oracle."""
    p = G.Problem(G.SEQ_PROTEIN, 60, sites, mean_len=1.0, seed=5)
    say("fp32 split units", sites, "nptn", p.nptn, "tiles", (p.nptn + 63) // 64)
    o = S.Oracle(p, nthreads=4)
    a, b = 0, int(p.nei[0, 0])
    lo, plo = o.loglik(a, b, with_patterns=True)
    res = {}
    for shape in (1, 2):
        g = G.Tree(p, precision=G.FP32)
        g.set_tile_shape(shape)
        res[shape] = g.loglik(a, b, with_patterns=True)
        g.close()
    (l1, p1), (l2, p2) = res[1], res[2]
    assert np.array_equal(p1, p2)
    assert rel(l1, l2) <= 1e-12
    assert rel(l1, lo) <= LOGL_TOL and np.max(np.abs(p1 - plo)) < 1e-3
    o.close()
