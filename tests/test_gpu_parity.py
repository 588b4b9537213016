"""GPU parity: the HIP engine (through the C ABI) against the CPU oracle on the same
seeded inputs.  Tolerance: |dlogL|/|logL| <= 1e-10 (FP64 form), scale counts bit-exact."""
import sys

import numpy as np
import pytest

import gems_amd as G
import support as S

pytestmark = pytest.mark.gpu

TOL = 1e-10


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


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
    (G.SEQ_DNA, 32, 10000, 0.1, 0.5),     # DNA GTR+G4, 32 taxa, 10k sites
    (G.SEQ_PROTEIN, 24, 3000, 0.2, 0.8),
    (G.SEQ_CODON, 12, 600, 0.1, 0.5),
]


@pytest.mark.parametrize("safe", [0, 1])
@pytest.mark.parametrize("case", CASES, ids=["dna", "protein", "codon"])
def test_loglik_parity(case, safe):
    seq, T, sites, ml, alpha = case
    say("case", case, "safe", safe)
    p = G.Problem(seq, T, sites, mean_len=ml, alpha=alpha, seed=11)
    g = G.Tree(p, scaling=safe)
    o = S.Oracle(p, safe=bool(safe), nthreads=4)
    for (u, v) in branches(p, 6) + [(0, int(p.nei[0, 0]))]:
        lg, plg = g.loglik(u, v, with_patterns=True)
        lo, plo = o.loglik(u, v, with_patterns=True)
        assert rel(lg, lo) <= TOL, (u, v, lg, lo)
        np.testing.assert_allclose(plg, plo, rtol=1e-10, atol=1e-10)


@pytest.mark.parametrize("safe", [0, 1])
def test_partials_and_scale_counts(safe):
    T = 150
    p = G.Problem(G.SEQ_DNA, T, 700, mean_len=0.6, alpha=0.5, seed=3)
    say("partials safe", safe, "nptn", p.nptn)
    g = G.Tree(p, scaling=safe)
    o = S.Oracle(p, safe=bool(safe))
    a, b = 0, int(p.nei[0, 0])
    assert rel(g.loglik(a, b), o.loglik(a, b)) <= TOL
    # several internal directions
    checked = 0
    for v in range(T, 2 * T - 2, 17):
        dad = int(p.nei[v, 0])
        pg, sg = g.partial(dad, v, safe=bool(safe))
        po, so = o.partial(dad, v, safe=bool(safe))
        assert np.array_equal(sg, so)
        scale = np.abs(po).max(axis=(1, 2), keepdims=True)
        assert np.all(np.abs(pg - po) <= 1e-12 * scale + 1e-300)
        checked += 1
    assert checked > 3


@pytest.mark.parametrize("safe,ncat", [(0, 4), (1, 4), (0, 2), (1, 6)])
def test_codon_partials_and_scale_counts(safe, ncat):
    """61 states: the vector-ALU kernels (LDS-staged for ncat <= 4, scalar loads above) and
    the matrix-core form (ncat <= 4) against the oracle's partials and scale counts.  The
    vector kernels run the oracle's sums in its order; eigen-space values of 61 states
    cancel heavily, so host rounding alone puts ~1e-12 to 1e-11 of the pattern's largest
    value between them, and the matrix core's summation order up to ~2e-10."""
    rng = np.random.default_rng(9)
    pp = S.make_problem(61, 16, 300, rng, ncat=ncat, mean_len=0.4, ambiguous=0.05)
    say("codon partials safe", safe, "ncat", ncat, "nptn", pp.nptn)
    g = G.Tree(pp, scaling=safe)
    o = S.Oracle(pp, safe=bool(safe))
    a, b = 0, int(pp.nei[0, 0])
    assert rel(g.loglik(a, b), o.loglik(a, b)) <= TOL
    gv = G.Tree(pp, scaling=safe)
    gv.set_f64_form(0)  # the vector-ALU kernels (LDS-staged for ncat <= 4)
    for v in range(16, 30, 3):
        dad = int(pp.nei[v, 0])
        po, so = o.partial(dad, v, safe=bool(safe))
        scale = np.abs(po).max(axis=(1, 2), keepdims=True)
        # the matrix-core form (default for ncat <= 4) sums in another order: 1e-9
        for tree, bound in ((g, 1e-10 if ncat > 4 else 1e-9), (gv, 1e-10)):
            pg, sg = tree.partial(dad, v, safe=bool(safe))
            assert np.array_equal(sg, so)
            worst = float((np.abs(pg - po) / (scale + 1e-300)).max())
            say("node", v, "max |diff| / pattern max", worst)
            assert worst <= bound


@pytest.mark.parametrize("safe", [0, 1])
def test_forced_scaling_caterpillar(safe):
    T = 300
    rng = np.random.default_rng(5)
    pp = S.make_problem(4, T, 150, rng, tree=S.caterpillar_tree(T, 1.5), ambiguous=0.05)
    say("caterpillar safe", safe)
    g = G.Tree(pp, scaling=safe)
    o = S.Oracle(pp, safe=bool(safe))
    a, b = 0, int(pp.nei[0, 0])
    lg, lo = g.loglik(a, b), o.loglik(a, b)
    assert rel(lg, lo) <= TOL
    _, sg = g.partial(a, b, safe=bool(safe))
    _, so = o.partial(a, b, safe=bool(safe))
    assert so.max() >= 1 and np.array_equal(sg, so)
    df_g, ddf_g, _ = g.derivs(a, b)
    df_o, ddf_o, _ = o.derivs(a, b)
    assert abs(df_g - df_o) <= 1e-9 * max(1, abs(df_o))
    assert abs(ddf_g - ddf_o) <= 1e-9 * max(1, abs(ddf_o))


@pytest.mark.parametrize("case", CASES, ids=["dna", "protein", "codon"])
def test_derivatives_and_theta(case):
    seq, T, sites, ml, alpha = case
    say("derivs", case)
    p = G.Problem(seq, T, min(sites, 3000), mean_len=ml, alpha=alpha, seed=5)
    safe = 1 if seq == G.SEQ_CODON else 0
    g = G.Tree(p, scaling=safe)
    o = S.Oracle(p, safe=bool(safe))
    for (u, v) in branches(p, 5):
        dg, ddg, rc = g.derivs(u, v)
        assert rc == 0
        do, ddo, _ = o.derivs(u, v)
        assert abs(dg - do) <= 1e-9 * max(1.0, abs(do)), (dg, do)
        assert abs(ddg - ddo) <= 1e-9 * max(1.0, abs(ddo)), (ddg, ddo)
        # Newton re-evaluation from the resident theta buffer at a new length
        t0 = g.get_length(u, v)
        for t in (t0 * 0.5, t0 * 2 + 0.01):
            g.set_length(u, v, t)
            o.set_length(u, v, t)
            dg2, ddg2, _ = g.derivs(u, v, reuse=True)
            do2, ddo2, _ = o.derivs(u, v, reuse=True)
            assert abs(dg2 - do2) <= 1e-9 * max(1.0, abs(do2))
            assert rel(g.loglik_from_theta(u, v), o.loglik_from_theta(u, v)) <= TOL


@pytest.mark.parametrize("nsites", [1, 2, 63, 64, 65, 130])
def test_ragged_pattern_counts(nsites):
    rng = np.random.default_rng(nsites)
    pp = S.make_problem(20, 9, nsites, rng, ambiguous=0.2, compress=False)
    say("ragged", nsites)
    g = G.Tree(pp)
    o = S.Oracle(pp)
    for (u, v) in branches(pp, 4):
        assert rel(g.loglik(u, v), o.loglik(u, v)) <= TOL


def test_newton_optimize_all_parity():
    p = G.Problem(G.SEQ_DNA, 16, 3000, seed=9)
    say("newton")
    g = G.Tree(p)
    o = S.Oracle(p)
    lg = g.optimize_all(iterations=5, tol=1e-6)
    lo = o.optimize_all(iterations=5, tol=1e-6)
    assert rel(lg, lo) <= 1e-9
    for (u, v) in branches(p):
        assert abs(g.get_length(u, v) - o.get_length(u, v)) <= 1e-7 * max(1e-3, o.get_length(u, v))


def test_deterministic_repeat():
    p = G.Problem(G.SEQ_PROTEIN, 30, 4000, seed=2)
    g = G.Tree(p)
    a = g.full_traversal()
    b = g.full_traversal()
    assert a == b


def test_invalid_ops_rejected():
    import ctypes
    p = G.Problem(G.SEQ_DNA, 8, 100, seed=1)
    g = G.Tree(p)
    ctx = g.context()
    op = (G.plk_op * 1)()
    op[0].out_slot = 99
    op[0].child[0] = -1
    op[0].child[1] = -2
    rc = G.lib().plk_update_partials(ctx, op, 1)
    assert rc == -1
    v = ctypes.c_double()
    assert G.lib().plk_branch_loglik(ctx, -1, -2, 0.1, ctypes.byref(v), None) == -1
    assert G.lib().plk_branch_loglik(ctx, 0, 1000, 0.1, ctypes.byref(v), None) == -1
    # the context still works afterwards
    o = S.Oracle(p)
    assert rel(g.loglik(0, int(p.nei[0, 0])), o.loglik(0, int(p.nei[0, 0]))) <= TOL


@pytest.mark.parametrize("prec", [G.FP64, G.FP32], ids=["fp64", "fp32"])
@pytest.mark.parametrize("safe", [0, 1])
def test_all_branch_derivs(prec, safe):
    if prec == G.FP32 and safe:
        pytest.skip("the FP32 form has no SAFE mode")
    p = G.Problem(G.SEQ_DNA, 20, 2000, seed=13)
    say("all-branch derivs", prec, safe)
    g = G.Tree(p, precision=prec, scaling=safe, all_directions=True)
    g.full_traversal()
    br, df, ddf = g.all_branch_derivs()
    assert len(br) == 2 * p.ntaxa - 3
    single = G.Tree(p, precision=prec, scaling=safe)
    o = S.Oracle(p, safe=bool(safe))
    tol = 1e-9 if prec == G.FP64 else 2e-4  # FP32: ~10x the measured 1.2e-5
    for (u, v), d1, d2 in zip(br, df, ddf):
        s1, s2, _ = single.derivs(u, v)
        assert d1 == s1 and d2 == s2, (u, v)
        o1, o2, _ = o.derivs(u, v)
        assert abs(d1 - o1) <= tol * max(1.0, abs(o1)) and abs(d2 - o2) <= tol * max(1.0, abs(o2))
    # after changing a length the pass recomputes only what depends on it
    u, v = br[5]
    g.set_length(u, v, 0.3)
    single.set_length(u, v, 0.3)
    br2, df2, _ = g.all_branch_derivs()
    for (a, b), d1 in zip(br2, df2):
        assert d1 == single.derivs(a, b)[0]
    assert g.loglik(*br[0]) == single.loglik(*br[0])


NEWTON_CASES = [
    (G.FP64, 0, G.SEQ_DNA, 2000),
    (G.FP64, 1, G.SEQ_PROTEIN, 1500),
    (G.FP32, 0, G.SEQ_PROTEIN, 1500),
    (G.FP64, 1, G.SEQ_CODON, 300),
]


@pytest.mark.parametrize("case", NEWTON_CASES, ids=["dna-f64", "protein-f64-safe", "protein-f32", "codon-f64-safe"])
def test_device_newton_matches_host_loop(case):
    """The device-resident Newton solve takes the same steps as minimizeNewton driven from
    the host one derivative call at a time: identical lengths, evaluation counts and logL."""
    prec, safe, seq, sites = case
    p = G.Problem(seq, 14, sites, seed=21)
    say("device newton", prec, safe, seq)
    dev = G.Tree(p, precision=prec, scaling=safe)
    host = G.Tree(p, precision=prec, scaling=safe)
    host.set_newton(False)
    o = S.Oracle(p, safe=bool(safe)) if prec == G.FP64 else None
    br = branches(p)
    for (u, v) in br[:6]:
        a = dev.optimize_one(u, v)
        b = host.optimize_one(u, v)
        assert a == b, (u, v, a, b)
        if o is not None:  # the oracle's minimizeNewton on the same sequence of solves
            # both stop once a step is below xacc (= min_bl, 1e-6) from an iterate that
            # brackets the same root, so they agree to 2 * xacc
            t_orc = o.optimize_one(u, v)
            assert abs(a - t_orc) <= 2e-6, (u, v, a, t_orc)
    assert dev.counters()[1] == host.counters()[1]  # derivative evaluations
    assert dev.loglik(*br[0]) == host.loglik(*br[0])


def test_device_newton_abi_errors():
    import ctypes
    p = G.Problem(G.SEQ_DNA, 8, 400, seed=3)
    g = G.Tree(p)
    g.full_traversal()
    ctx = g.context()
    u, v = branches(p)[0]
    # node references of the branch as the tree layer passes them
    ra, rb = -(u + 1) if u < p.ntaxa else u - p.ntaxa, -(v + 1) if v < p.ntaxa else v - p.ntaxa
    out, ev = ctypes.c_double(), ctypes.c_int()
    L = G.lib()
    assert L.plk_newton_branch(ctx, ra, rb, 1e-6, 0.1, 1e-7, 1e-6, 100, ctypes.byref(out), None) == -1  # x2 < x1
    assert L.plk_newton_branch(ctx, ra, rb, 1e-6, 0.1, 10.0, 0.0, 100, ctypes.byref(out), None) == -1  # xacc
    # max_nr = 0: the reference loop never runs and reports the iteration limit
    rc = L.plk_newton_branch(ctx, ra, rb, 1e-6, 0.1, 10.0, 1e-6, 0, ctypes.byref(out), ctypes.byref(ev))
    assert rc == -4 and b"Maximum number of iterations" in L.plk_last_error()
    rc = L.plk_newton_branch(ctx, ra, rb, 1e-6, 0.1, 10.0, 1e-6, 100, ctypes.byref(out), ctypes.byref(ev))
    assert rc == 0 and 1e-6 <= out.value <= 10.0 and ev.value >= 1


SCALE_RULE_CASES = [(4, 300, 150, 1.5, 0), (4, 300, 150, 1.5, 1), (20, 100, 120, 1.5, 0), (20, 100, 120, 1.5, 1),
                    (61, 40, 60, 1.0, 1)]


@pytest.mark.parametrize("case", SCALE_RULE_CASES, ids=lambda c: f"n{c[0]}-{'safe' if c[4] else 'norm'}")
def test_scale_counts_follow_rule(case):
    """The device's scale counts on internal directions of a deep caterpillar equal the
    counts derived from the reference's rule (tree/phylokernelnew.h:1643-1644,
    1757-1760, 1929-1936, 1997-2034) by support.derived_scale_counts, which tracks the
    true magnitudes with explicit exponents and shares no arithmetic with the oracle."""
    n, T, sites, blen, safe = case
    rng = np.random.default_rng(5)
    pp = S.make_problem(n, T, sites, rng, tree=S.caterpillar_tree(T, blen), ambiguous=0.05)
    say("scale rule", case, "nptn", pp.nptn)
    g = G.Tree(pp, scaling=safe)
    memo = {}
    checked = top = 0
    for v in list(range(T, 2 * T - 2, 7)) + [2 * T - 3]:
        for k in range(3):
            d = int(pp.nei[v, k])
            _, sg = g.partial(d, v, safe=bool(safe))
            sg = sg.reshape(pp.nptn, -1).astype(np.int64) if safe else sg.astype(np.int64)
            dc, amb = S.derived_scale_counts(pp, d, v, bool(safe), memo=memo)
            assert amb.sum() <= 1
            assert np.array_equal(sg[~amb], dc[~amb]), (v, d)
            top = max(top, int(sg.max()))
            checked += 1
    assert checked >= 3 * (T - 2) // 7 and top >= 2


@pytest.mark.parametrize("safe", [0, 1], ids=["norm", "safe"])
@pytest.mark.parametrize("seq", [G.SEQ_DNA, G.SEQ_PROTEIN], ids=["dna", "protein"])
def test_pattern_side_outputs(seq, safe):
    """_pattern_lh_cat times the pattern's scaling equals the true per-category likelihood
    of an independent real-space pruning (1e-10 of the pattern's likelihood); the
    save_log_value = false form gives
    log(lh) + scaling = the per-pattern logL; L'/L per pattern (the G_matrix row) equals
    a central finite difference of the per-pattern logL and sums, weighted, to df."""
    p = G.Problem(seq, 10, 400, mean_len=0.2, alpha=0.6, seed=21)
    g = G.Tree(p, scaling=safe)
    a, b = 0, int(p.nei[0, 0])
    lg, plg = g.loglik(a, b, with_patterns=True)
    cat = g.pattern_lh_cat()
    lh, sc = g.pattern_scaling()
    np.testing.assert_allclose(np.log(lh) + sc, plg, rtol=1e-13, atol=1e-12)
    ref = S.pruning_lh_cat(p, a, b)  # ncat x nptn
    # relative to the pattern's likelihood: a category far below the others is an
    # eigen-space sum with cancellation, exact only to ~1e-16 of the pattern total
    err = np.abs((cat * np.exp(sc)[:, None]).T - ref) / ref.sum(axis=0)[None, :]
    assert float(err.max()) <= 1e-10
    u, v = int(p.nei[p.ntaxa, 0]), p.ntaxa
    df, _, _ = g.derivs(u, v)
    dl = g.pattern_dlogl()
    assert abs(float(np.dot(dl, p.freq)) - df) <= 1e-9 * max(1.0, abs(df))
    t0 = g.get_length(u, v)
    h = 1e-5 * max(t0, 1e-3)
    o = S.Oracle(p, safe=bool(safe))
    o.set_length(u, v, t0 + h)
    _, lp = o.loglik(u, v, with_patterns=True)
    o.set_length(u, v, t0 - h)
    _, lm = o.loglik(u, v, with_patterns=True)
    fd = (lp - lm) / (2 * h)
    np.testing.assert_allclose(dl, fd, rtol=1e-5, atol=1e-6)


@pytest.mark.parametrize("precision", [G.FP64, G.FP32], ids=["fp64", "fp32"])
@pytest.mark.parametrize("n,zeros", [(4, [2]), (20, [3, 17])], ids=["dna", "protein"])
def test_zero_frequency_states_loglik(n, zeros, precision):
    """A model with zero-frequency states (decomposed as model/modelmarkov.cpp:1453-1597
    does) on an alignment that never shows them: the device logL equals the oracle's and
    the oracle's equals the independent expm pruning on the full Q."""
    import ctypes
    rng = np.random.default_rng(31 + n)
    p = S.make_problem(n, 12, 800, rng)
    pi = p.pi.copy()
    pi[zeros] = 0.0
    pi /= pi.sum()
    exch = rng.uniform(0.5, 2.0, n * (n - 1) // 2) if n == 4 else rng.gamma(1.0, 1.0, n * (n - 1) // 2)
    Q = S.rate_matrix(exch, pi)
    ev, U, Ui = np.zeros(n), np.zeros(n * n), np.zeros(n * n)
    dp = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))  # noqa: E731
    Qc = np.ascontiguousarray(Q).ravel()
    assert G.lib().plk_eigen(n, dp(Qc), dp(pi), dp(ev), dp(U), dp(Ui)) == 0
    st = np.array(p.states)
    for z in zeros:
        st[st == z] = (z + 1) % n
    p.pi, p.Q, p.eval, p.evec, p.inv_evec = pi, Q, ev, U.reshape(n, n), Ui.reshape(n, n)
    p.states = np.ascontiguousarray(st)
    p.tip_lh = np.stack([p.inv_evec @ S.tip_likelihood(p, s) for s in range(p.su + 1)])
    g = G.Tree(p, precision=precision)
    o = S.Oracle(p)
    a, b = 0, int(p.nei[0, 0])
    lo = o.loglik(a, b)
    _, ref = S.pruning_loglik(p, a, b)
    assert rel(lo, ref) <= 1e-10
    assert rel(g.loglik(a, b), lo) <= (1e-10 if precision == G.FP64 else 1e-6)


def test_failure_paths_zero_likelihood_pattern():
    """A pattern whose likelihood is exactly 0 (a tip shows a state of zero frequency),
    forced from the data so every kernel runs normally: the branch logL takes the
    reference's per-pattern fallback (log 2^-1024 for the non-finite pattern,
    tree/phylokernelnew.h:3167-3184) as the oracle does; the derivative is non-finite, so
    NORM scaling fails with the reference's "use -safe" error (2594-2596) and SAFE warns
    and returns df = ddf = 0 (2646-2649)."""
    import ctypes
    n, z = 4, 2
    rng = np.random.default_rng(77)
    p = S.make_problem(n, 10, 300, rng)
    pi = p.pi.copy()
    pi[z] = 0.0
    pi /= pi.sum()
    Q = S.rate_matrix(rng.uniform(0.5, 2.0, 6), pi)
    ev, U, Ui = np.zeros(n), np.zeros(n * n), np.zeros(n * n)
    dp = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))  # noqa: E731
    Qc = np.ascontiguousarray(Q).ravel()
    assert G.lib().plk_eigen(n, dp(Qc), dp(pi), dp(ev), dp(U), dp(Ui)) == 0
    st = np.array(p.states)
    st[st == z] = 1
    st[0, 5] = z  # pattern 5: taxon 0 shows the unreachable state
    p.pi, p.Q, p.eval, p.evec, p.inv_evec = pi, Q, ev, U.reshape(n, n), Ui.reshape(n, n)
    p.states = np.ascontiguousarray(st)
    p.tip_lh = np.stack([p.inv_evec @ S.tip_likelihood(p, s) for s in range(p.su + 1)])
    a, b = 0, int(p.nei[0, 0])
    o = S.Oracle(p)
    lo, plo = o.loglik(a, b, with_patterns=True)
    for safe in (0, 1):
        g = G.Tree(p, scaling=safe)
        lg, plg = g.loglik(a, b, with_patterns=True)
        assert np.isfinite(lg) and rel(lg, lo) <= 1e-10
        assert plg[5] == 4 * S.LOG_SCALING_THRESHOLD and plo[5] == plg[5]
        assert abs(lg - float(np.dot(plg, p.freq))) <= 1e-9 * abs(lg)
        df, ddf, rc = g.derivs(a, b)
        if safe:
            assert rc == 0 and df == 0.0 and ddf == 0.0
        else:
            assert rc == -4 and b"use SAFE scaling" in G.lib().plk_last_error()
