"""Pins the CPU oracle (oracle/plk_oracle.c) against checks that do not share its
arithmetic: brute-force enumeration, real-space pruning with scipy.linalg.expm,
closed forms, invariants and finite differences."""
import numpy as np
import pytest

import support as S


def branches(p):
    out = []
    for u in range(2 * p.ntaxa - 2):
        for k in range(3):
            v = int(p.nei[u, k])
            if v > u:
                out.append((u, v))
    return out


@pytest.mark.parametrize("n,T,sites", [(4, 4, 300), (4, 5, 300), (20, 4, 60), (20, 5, 30)])
def test_brute_force(n, T, sites):
    rng = np.random.default_rng(n * 100 + T)
    p = S.make_problem(n, T, sites, rng, ambiguous=0.15)
    o = S.Oracle(p)
    a, b = 0, int(p.nei[0, 0])
    lp_bf, tot_bf = S.brute_force_loglik(p, a, b)
    tot, lp = o.loglik(a, b, with_patterns=True)
    assert abs(tot - tot_bf) <= 1e-12 * abs(tot_bf)
    np.testing.assert_allclose(lp, lp_bf, rtol=1e-12, atol=1e-12)


def test_codon_pruning():
    rng = np.random.default_rng(7)
    p = S.make_problem(61, 6, 40, rng)
    o = S.Oracle(p, safe=True)
    a, b = 0, int(p.nei[0, 0])
    _, ref = S.pruning_loglik(p, a, b)
    assert abs(o.loglik(a, b) - ref) <= 1e-11 * abs(ref)


@pytest.mark.parametrize("n", [4, 20])
def test_pulley_and_pruning(n):
    rng = np.random.default_rng(3 + n)
    p = S.make_problem(n, 12, 400, rng, ambiguous=0.05)
    o = S.Oracle(p)
    vals = [o.loglik(u, v) for (u, v) in branches(p)]
    _, ref = S.pruning_loglik(p, 0, int(p.nei[0, 0]))
    assert np.ptp(vals) <= 1e-12 * abs(ref)
    assert abs(vals[0] - ref) <= 1e-12 * abs(ref)


def test_three_taxon_closed_form():
    # star tree: L = sum_c w_c sum_x pi_x prod_leaf P_c(t_leaf)[x, s_leaf]
    import scipy.linalg
    rng = np.random.default_rng(11)
    p = S.make_problem(4, 3, 50, rng)
    o = S.Oracle(p)
    lp = np.zeros(p.nptn)
    center = 3
    for c in range(len(p.rates)):
        acc = np.tile(p.pi, (p.nptn, 1))
        for leaf in range(3):
            t = p.blen[leaf, 0]
            P = scipy.linalg.expm(p.Q * p.rates[c] * t)
            acc *= P[:, p.states[leaf]].T
        lp += p.props[c] * acc.sum(axis=1)
    ref = float(np.dot(np.log(lp), p.freq))
    assert abs(o.loglik(center, 0) - ref) <= 1e-12 * abs(ref)


def test_compression_invariance():
    rng = np.random.default_rng(5)
    p1 = S.make_problem(4, 6, 500, np.random.default_rng(5), compress=False)
    p2 = S.make_problem(4, 6, 500, np.random.default_rng(5), compress=True)
    assert p2.nptn < p1.nptn
    l1 = S.Oracle(p1).loglik(0, int(p1.nei[0, 0]))
    l2 = S.Oracle(p2).loglik(0, int(p2.nei[0, 0]))
    assert abs(l1 - l2) <= 1e-11 * abs(l1)


@pytest.mark.parametrize("n", [4, 20])
def test_scaling_deep_tree(n):
    """A long caterpillar forces 2^-256 rescaling many times; logL must equal the
    renormalising real-space pruning, and NORM must equal SAFE."""
    T = 300 if n == 4 else 100
    rng = np.random.default_rng(17)
    p = S.make_problem(n, T, 80, rng, tree=S.caterpillar_tree(T, 1.5))
    on = S.Oracle(p, safe=False)
    os_ = S.Oracle(p, safe=True)
    a, b = 0, int(p.nei[0, 0])
    ln = on.loglik(a, b)
    ls = os_.loglik(a, b)
    _, ref = S.pruning_loglik(p, a, b)
    assert abs(ln - ref) <= 1e-11 * abs(ref)
    assert abs(ls - ref) <= 1e-11 * abs(ref)
    # scaling really happened
    part, sc = on.partial(a, b)
    assert sc.max() >= 1
    part_s, sc_s = os_.partial(a, b, safe=True)
    assert sc_s.max() >= 1


@pytest.mark.parametrize("n", [4, 20])
def test_derivatives_finite_difference(n):
    rng = np.random.default_rng(23 + n)
    p = S.make_problem(n, 8, 300, rng)
    o = S.Oracle(p)
    for (u, v) in branches(p)[:4]:
        t0 = o.get_length(u, v)
        df, ddf, st = o.derivs(u, v)
        assert st == 0
        h = 1e-4 * max(t0, 1e-3)

        def f(t):
            o.set_length(u, v, t)
            return o.loglik(u, v)

        # Richardson-extrapolated central differences
        d1h = (f(t0 + h) - f(t0 - h)) / (2 * h)
        d1h2 = (f(t0 + h / 2) - f(t0 - h / 2)) / h
        d1 = (4 * d1h2 - d1h) / 3
        f0 = f(t0)
        d2h = (f(t0 + h) - 2 * f0 + f(t0 - h)) / h ** 2
        d2h2 = (f(t0 + h / 2) - 2 * f0 + f(t0 - h / 2)) / (h / 2) ** 2
        d2 = (4 * d2h2 - d2h) / 3
        o.set_length(u, v, t0)
        assert abs(df - d1) <= 1e-6 * max(1.0, abs(df))
        assert abs(ddf - d2) <= 1e-4 * max(1.0, abs(ddf))


def test_from_theta_equals_branch():
    rng = np.random.default_rng(29)
    p = S.make_problem(20, 7, 100, rng)
    o = S.Oracle(p)
    u, v = branches(p)[3]
    o.derivs(u, v)
    for t in (0.01, 0.3, 2.0):
        o.lib.orc_set_length(o.h, u, v, t)
        # from-theta uses the current length; the theta buffer stays valid for this branch
        lt = o.loglik_from_theta(u, v)
        lb = o.loglik(u, v)
        assert abs(lt - lb) <= 1e-12 * abs(lb)
        o.derivs(u, v)


def test_newton_reaches_stationary_point():
    rng = np.random.default_rng(31)
    p = S.make_problem(4, 10, 800, rng)
    o = S.Oracle(p)
    l0 = o.loglik(0, int(p.nei[0, 0]))
    l1 = o.optimize_all(iterations=20, tol=1e-6)
    assert l1 >= l0
    for (u, v) in branches(p):
        t = o.get_length(u, v)
        df, ddf, _ = o.derivs(u, v)
        if 1e-5 < t < 9.0:
            assert abs(df) < 0.05, (u, v, t, df)


def test_tip_vectors_closed_form():
    rng = np.random.default_rng(2)
    for n, st in [(4, S.SEQ_DNA), (20, S.SEQ_PROTEIN)]:
        p = S.make_problem(n, 4, 10, rng, seq_type=st)
        o = S.Oracle(p)
        for s in range(o.tip_lh.shape[0]):
            ref = p.inv_evec @ S.tip_likelihood(p, s)
            np.testing.assert_allclose(o.tip_lh[s], ref, rtol=0, atol=1e-14)


def test_planted_mistake_is_caught():
    """Transposing P in the independent pruning reference must break agreement: the
    checks above are sensitive to an operand-order mistake."""
    rng = np.random.default_rng(41)
    p = S.make_problem(4, 5, 100, rng)
    o = S.Oracle(p)
    a, b = 0, int(p.nei[0, 0])
    good = o.loglik(a, b)
    bad = p.__class__(**dict(p.__dict__))
    bad.Q = p.Q.T.copy()  # wrong orientation of the rate matrix
    _, ref_bad = S.pruning_loglik(bad, a, b)
    assert abs(good - ref_bad) > 1e-6 * abs(good)


def test_rell_pattern_frequencies_reproduce_logl():
    """A replicate whose counts are the pattern frequencies is the original alignment:
    its RELL score is the tree logL, up to the float rounding of each pattern logL."""
    rng = np.random.default_rng(31)
    p = S.make_problem(4, 8, 400, rng)
    o = S.Oracle(p)
    a, b = 0, int(p.nei[0, 0])
    tot, lp = o.loglik(a, b, with_patterns=True)
    freq = np.asarray(p.freq, dtype=np.float64)
    counts = np.stack([freq, 2 * freq, np.zeros_like(freq)]).astype(np.float32)
    r = S.OracleRell(counts)
    rell, _ = r.save_tree(lp, tot)
    bound = float(np.sum(np.abs(freq * lp))) * 2.0 ** -24
    assert abs(rell[0] - tot) <= bound
    assert abs(rell[1] - 2 * tot) <= 2 * bound
    assert rell[2] == 0.0


def test_rell_update_rule_hand_worked():
    """saveCurrentTree's replicate rule (tree/iqtree.cpp:3917-3931) on a sequence of
    scores worked out by hand: one replicate, one pattern of count 1, eps = 0.5."""
    r = S.OracleRell(np.ones((1, 1), dtype=np.float32))
    seq = [
        # (pattern logL, uniform, expected improved, best logL, count, orig)
        (-100.0, 0.99, 1, -100.0, 1, -100.0),   # first tree always wins (best = -DBL_MAX)
        (-99.75, 0.40, 1, -99.75, 2, -99.75),   # tie within eps: 0.40 <= 1/2 -> count 2
        (-99.5, 0.40, 0, -99.75, 2, -99.75),    # tie: 0.40 > 1/3 -> kept
        (-99.5, 0.25, 1, -99.5, 3, -99.5),      # tie: 0.25 <= 1/3 -> count 3, max taken
        (-98.0, 0.99, 1, -98.0, 1, -98.0),      # clearly better -> count reset to 1
        (-98.6, 0.00, 0, -98.0, 1, -98.0),      # clearly worse (below best - eps)
    ]
    for lh, u, imp, best, cnt, orig in seq:
        _, improved = r.save_tree(np.array([lh]), lh, eps=0.5, uniform=np.array([u]))
        assert improved[0] == imp and r.logl[0] == best and r.best_counts[0] == cnt and r.orig[0] == orig, lh


def _swap(nei, blen, n1, n2, move):
    """The NNI move of getBestNNIForBran on copies of (nei, blen): node1's first other
    neighbour trades places (slot and branch length) with node2's move-th other one."""
    nei, blen = nei.copy(), blen.copy()
    a_slot = [k for k in range(3) if nei[n1, k] != n2][0]
    x_slot = [k for k in range(3) if nei[n2, k] != n1][move]
    A, X = int(nei[n1, a_slot]), int(nei[n2, x_slot])
    la, lx = blen[n1, a_slot], blen[n2, x_slot]
    ka = [k for k in range(3) if nei[A, k] == n1][0]
    kx = [k for k in range(3) if nei[X, k] == n2][0]
    nei[n1, a_slot], blen[n1, a_slot] = X, lx
    nei[X, kx], blen[X, kx] = n1, lx
    nei[n2, x_slot], blen[n2, x_slot] = A, la
    nei[A, ka], blen[A, ka] = n2, la
    return nei, blen


def test_nni_eval_equals_swapped_tree_from_scratch():
    """orc_nni_eval's score is the logL of the swapped tree, built from scratch, after the
    same five branch optimisations in the same order; the tree is left as it was."""
    rng = np.random.default_rng(41)
    p = S.make_problem(4, 9, 300, rng)
    o = S.Oracle(p)
    a, b = 0, int(p.nei[0, 0])
    before = o.loglik(a, b)
    T = p.ntaxa
    checked = 0
    for n1 in range(T, 2 * T - 2):
        for n2 in [int(x) for x in p.nei[n1] if x > n1]:
            for move in (0, 1):
                score, lens = o.nni_eval(n1, n2, move)
                nei, blen = _swap(p.nei, p.blen, n1, n2, move)
                q = S.Problem(**{**p.__dict__, "nei": nei, "blen": blen})
                f = S.Oracle(q)
                order = [(n1, int(y)) for y in nei[n1] if y != n2] + [(n1, n2)] + [(n2, int(z)) for z in nei[n2] if z != n1]
                got = [f.optimize_one(u, v, max_nr=10) for (u, v) in order]
                assert np.array_equal(np.array([got[2], got[0], got[1], got[3], got[4]]), lens)
                assert f.loglik_from_theta(*order[-1]) == score
                checked += 1
    assert checked >= 8
    assert o.loglik(a, b) == before


def _scale_cases():
    # (states, taxa, sites, caterpillar branch length, safe)
    return [(4, 300, 150, 1.5, False), (4, 300, 150, 1.5, True), (20, 100, 120, 1.5, False),
            (20, 100, 120, 1.5, True), (61, 40, 60, 1.0, True)]


def _scale_problem(n, T, sites, blen, seed=5):
    rng = np.random.default_rng(seed)
    return S.make_problem(n, T, sites, rng, tree=S.caterpillar_tree(T, blen), ambiguous=0.05)


def _count_mismatches(o, p, safe, memo, rule="reference"):
    """Compare o's scale counts with the rule-derived ones on every internal direction."""
    T = p.ntaxa
    o.loglik(0, int(p.nei[0, 0]))
    bad = total = amb_total = top = 0
    for v in range(T, 2 * T - 2):
        for k in range(3):
            d = int(p.nei[v, k])
            _, so = o.partial(d, v, safe=safe)
            so = so.reshape(p.nptn, -1).astype(np.int64) if safe else so.astype(np.int64)
            dc, amb = S.derived_scale_counts(p, d, v, safe, rule=rule, memo=memo)
            bad += int(np.sum((so != dc)[~amb]))
            amb_total += int(amb.sum())
            total += so.size
            top = max(top, int(so.max()))
    return bad, total, amb_total, top


@pytest.mark.parametrize("case", _scale_cases(), ids=lambda c: f"n{c[0]}-{'safe' if c[4] else 'norm'}")
def test_scale_counts_follow_rule(case):
    """Every scale count the oracle stores, on every direction of a deep caterpillar (many
    forced 2^-256 rescalings), equals the count derived from the reference's rule by
    support.derived_scale_counts, which tracks the true magnitudes with explicit
    exponents and shares no arithmetic with the oracle.  The planted variants of the
    rule (NORM maximum of the last category only, a threshold of 2^-255, the maximum
    taken before the U^-1 projection) each give different counts on the same problem,
    so a mistake of any of these kinds in the oracle would fail this test."""
    n, T, sites, blen, safe = case
    p = _scale_problem(n, T, sites, blen)
    o = S.Oracle(p, safe=safe)
    memo = {}
    bad, total, amb, top = _count_mismatches(o, p, safe, memo)
    assert total > 0 and amb <= total * 1e-4
    assert bad == 0
    assert top >= 2  # rescaling happened more than once along a path
    for rule in ("last_category", "threshold_2^-255", "before_projection"):
        if rule == "last_category" and safe:
            continue
        diff = 0
        for v in range(T, 2 * T - 2):
            for k in range(3):
                d = int(p.nei[v, k])
                ref, _ = S.derived_scale_counts(p, d, v, safe, memo=memo)
                var, _ = S.derived_scale_counts(p, d, v, safe, rule=rule, memo=memo)
                diff += int(np.sum(ref != var))
        assert diff > 0, rule


def test_scale_counts_tip_cases():
    """Tip-tip directions store 0 (no check); a tip-internal direction stores the
    internal child's count plus at most one (tree/phylokernelnew.h:1643-1644, 1757-1760);
    internal-internal the children's sum plus at most one (1929-1936)."""
    p = _scale_problem(4, 120, 100, 1.5, seed=8)
    for safe in (False, True):
        o = S.Oracle(p, safe=safe)
        o.loglik(0, int(p.nei[0, 0]))
        T = p.ntaxa
        for v in range(T, 2 * T - 2):
            for k in range(3):
                d = int(p.nei[v, k])
                kids = [int(x) for x in p.nei[v] if x >= 0 and x != d]
                _, s = o.partial(d, v, safe=safe)
                s = s.astype(np.int64)
                base = np.zeros_like(s)
                for ch in kids:
                    if ch >= T:
                        _, sc = o.partial(v, ch, safe=safe)
                        base += sc.astype(np.int64)
                if all(ch < T for ch in kids):
                    assert not s.any()
                else:
                    inc = s - base
                    assert inc.min() >= 0 and inc.max() <= 1


def test_planted_category_max_in_oracle_is_caught(tmp_path):
    """The oracle rebuilt with a planted mistake (NORM's maximum reset per category, so
    its check sees the last category only) fails the rule comparison."""
    import shutil
    import subprocess
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    so = tmp_path / "libplk_oracle_planted.so"
    subprocess.run([cc, "-O2", "-fPIC", "-fopenmp", "-std=c11", "-DORC_V=4", "-DORC_PLANT_CATEGORY_MAX",
                    "-march=x86-64-v3", "-shared", "-o", str(so), str(S.ORACLE_DIR / "plk_oracle.c"), "-lm"],
                   check=True)
    lib = S.load_oracle(so)
    p = _scale_problem(4, 300, 150, 1.5)
    bad, total, _, _ = _count_mismatches(S.Oracle(p, safe=False, lib=lib), p, False, {})
    assert bad > 0
    # the logL is blind to where the power-of-two scaling happens
    good = S.Oracle(p, safe=False).loglik(0, int(p.nei[0, 0]))
    planted = S.Oracle(p, safe=False, lib=lib).loglik(0, int(p.nei[0, 0]))
    assert abs(good - planted) <= 1e-11 * abs(good)


# ---------------------------------------------------------------------------------------
# +I: invariant-site mass (ptn_invar)


def _invar_problem(n, T, sites, seed, tree=None, p_inv=0.3, frac=0.3, ambiguous=0.05):
    rng = np.random.default_rng(seed)
    p = S.make_problem(n, T, sites, rng, tree=tree, ambiguous=ambiguous)
    S.plant_constant_columns(p, frac, rng)
    return S.apply_pinvar(p, p_inv)


@pytest.mark.parametrize("n,T,sites", [(4, 4, 300), (4, 5, 200), (20, 4, 60)])
def test_invar_brute_force(n, T, sites):
    """logL with ptn_invar (lh = sum_c w_c L_c + invar, tree/phylokernelnew.h:3097) against
    brute-force enumeration with the invariant term added."""
    p = _invar_problem(n, T, sites, seed=n + T)
    assert (p.invar > 0).sum() >= 5
    o = S.Oracle(p)
    a, b = 0, int(p.nei[0, 0])
    lp_bf, tot_bf = S.brute_force_loglik(p, a, b)
    tot, lp = o.loglik(a, b, with_patterns=True)
    assert abs(tot - tot_bf) <= 1e-12 * abs(tot_bf)
    np.testing.assert_allclose(lp, lp_bf, rtol=1e-12, atol=1e-12)
    # the invariant term changes the answer (the check is not vacuous)
    q = _invar_problem(n, T, sites, seed=n + T)
    q.invar = None
    assert abs(S.brute_force_loglik(q, a, b)[1] - tot_bf) > 1e-3


@pytest.mark.parametrize("safe", [False, True])
def test_invar_deep_tree_and_gate(safe):
    """A deep caterpillar with constant columns: patterns with invariant mass are never
    rescaled (the ptn_invar == 0 gate of tree/phylokernelnew.h:1999, 2020), their
    partials may underflow and the invariant mass carries them; the oracle's logL equals
    the renormalising pruning with the invariant term, and its scale counts equal the
    rule-derived ones (zero on every pattern with invariant mass)."""
    T = 200
    p = _invar_problem(4, T, 150, seed=3, tree=S.caterpillar_tree(T, 1.5))
    o = S.Oracle(p, safe=safe)
    a, b = 0, int(p.nei[0, 0])
    lp_ref, ref = S.pruning_loglik(p, a, b)
    tot, lp = o.loglik(a, b, with_patterns=True)
    assert abs(tot - ref) <= 1e-10 * abs(ref)
    np.testing.assert_allclose(lp, lp_ref, rtol=1e-10, atol=1e-11)  # the all-unknown column has logL 0
    bad, total, amb, top = _count_mismatches(o, p, safe, {})
    assert bad == 0 and total > 0 and top >= 2
    inv = p.invar > 0
    for v in range(T, 2 * T - 2, 7):
        d = int(p.nei[v, 0])
        _, sc = o.partial(d, v, safe=safe)
        sc = sc.reshape(p.nptn, -1)
        assert not sc[inv].any()


def test_ptn_invar_library_equals_definition():
    """plk_ptn_invar (the computeConst/computePtnInvar coding) equals the definition
    p_inv * sum_x pi_x prod_t [taxon t allows x] on DNA with ambiguity codes, protein with
    B/Z/J and codons, and the hand-worked columns below."""
    import gems_amd as G
    for n, seed in ((4, 1), (20, 2), (61, 3)):
        rng = np.random.default_rng(seed)
        p = S.make_problem(n, 9, 400, rng, ambiguous=0.1)
        S.plant_constant_columns(p, 0.4, rng)
        want = S.ptn_invar_identity(p, 0.25)
        got = G.ptn_invar(n, p.seq_type, p.states, p.pi, 0.25)
        np.testing.assert_allclose(got, want, rtol=1e-15, atol=0)
        assert (got > 0).sum() >= 0.3 * p.nptn and (got == 0).sum() > 0
    pi = np.array([0.1, 0.2, 0.3, 0.4])
    # columns: A,A | A,R(A|G = mask 5, code 8) | R,R | A,C | -,- | Y(C|T = mask 10, code 13),C
    cols = np.array([[0, 0], [0, 8], [8, 8], [0, 1], [18, 18], [13, 1]], dtype=np.uint8).T
    got = G.ptn_invar(4, S.SEQ_DNA, cols, pi, 0.5)
    np.testing.assert_array_equal(got, 0.5 * np.array([0.1, 0.1, 0.1 + 0.3, 0.0, 1.0, 0.2]))
    # protein: B (N|D) with N; Z with Z; J with K (nothing in common)
    pi20 = np.full(20, 0.05)
    cols = np.array([[20, 2], [21, 21], [22, 11]], dtype=np.uint8).T
    got = G.ptn_invar(20, S.SEQ_PROTEIN, cols, pi20, 0.2)
    np.testing.assert_allclose(got, 0.2 * np.array([0.05, 0.10, 0.0]), rtol=1e-15)
