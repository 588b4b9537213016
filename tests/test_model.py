"""Host model code of the engine (gems_amd/csrc/model.cpp, sim.cpp) against scipy
fixtures, closed forms and invariants.  CPU only: no kernel is launched."""
import ctypes
import json
import pathlib

import numpy as np
import pytest

import gems_amd as G
import scipy.linalg
import support as S

FIX = json.loads((pathlib.Path(__file__).parent / "golden" / "model_fixtures.json").read_text())
L = G.lib()


def _d(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def eigen(Q, pi):
    n = len(pi)
    ev, U, Ui = np.zeros(n), np.zeros(n * n), np.zeros(n * n)
    assert L.plk_eigen(n, _d(np.ascontiguousarray(Q.ravel())), _d(pi), _d(ev), _d(U), _d(Ui)) == 0
    return ev, U.reshape(n, n), Ui.reshape(n, n)


@pytest.mark.parametrize("case", FIX["gamma"], ids=lambda c: f"a{c['alpha']}k{c['ncat']}")
def test_gamma_rates_vs_scipy(case):
    r = np.zeros(case["ncat"])
    assert L.plk_gamma_rates(case["alpha"], case["ncat"], _d(r)) == 0
    # AS 91 / AS 32 are accurate to ~1e-6 relative (their own stopping rules)
    np.testing.assert_allclose(r, case["rates"], rtol=2e-5, atol=1e-9)
    assert abs(r.mean() - 1.0) < 1e-12


def test_gamma_known_answer():
    r = np.zeros(4)
    L.plk_gamma_rates(0.5, 4, _d(r))
    np.testing.assert_allclose(r, [0.0334, 0.2519, 0.8203, 2.8944], atol=6e-5)


@pytest.mark.parametrize("case", FIX["pmatrix"], ids=lambda c: f"n{c['n']}t{c['t']}")
def test_rate_matrix_and_pmatrix(case):
    n = case["n"]
    pi = np.array(case["pi"])
    exch = np.array(case["exch"])
    Q = np.zeros(n * n)
    assert L.plk_rate_matrix(n, _d(exch), _d(pi), _d(Q)) == 0
    np.testing.assert_allclose(Q, case["Q"], rtol=1e-13, atol=1e-15)
    Q = Q.reshape(n, n)
    ev, U, Ui = eigen(Q, pi)
    np.testing.assert_allclose(U @ np.diag(ev) @ Ui, Q, atol=1e-13)
    np.testing.assert_allclose(Ui @ U, np.eye(n), atol=1e-13)
    assert np.all(np.diff(ev) >= 0) and abs(ev[-1]) < 1e-12
    P = U @ np.diag(np.exp(ev * case["t"])) @ Ui
    np.testing.assert_allclose(P.ravel(), case["P"], atol=2e-14)


def test_pmatrix_invariants():
    rng = np.random.default_rng(3)
    n = 20
    pi = rng.dirichlet(np.ones(n))
    exch = rng.gamma(1.0, 1.0, 190)
    Q = np.zeros(n * n)
    L.plk_rate_matrix(n, _d(exch), _d(pi), _d(Q))
    Q = Q.reshape(n, n)
    assert abs(-(pi * np.diag(Q)).sum() - 1.0) < 1e-13
    np.testing.assert_allclose(Q.sum(1), 0, atol=1e-14)
    ev, U, Ui = eigen(Q, pi)
    P = lambda t: U @ np.diag(np.exp(ev * t)) @ Ui
    np.testing.assert_allclose(P(0.0), np.eye(n), atol=1e-13)
    np.testing.assert_allclose(P(0.3) @ P(0.4), P(0.7), atol=1e-13)
    np.testing.assert_allclose(P(0.5).sum(1), 1, atol=1e-13)
    D = pi[:, None] * P(0.5)
    np.testing.assert_allclose(D, D.T, atol=1e-14)


def test_jc_closed_form():
    n = 4
    pi = np.full(4, 0.25)
    Q = np.zeros(16)
    L.plk_rate_matrix(4, _d(np.ones(6)), _d(pi), _d(Q))
    ev, U, Ui = eigen(Q.reshape(4, 4), pi)
    for t in (0.01, 0.2, 1.0):
        P = U @ np.diag(np.exp(ev * t)) @ Ui
        assert abs(P[0, 0] - (0.25 + 0.75 * np.exp(-4 * t / 3))) < 1e-14
        assert abs(P[0, 1] - (0.25 - 0.25 * np.exp(-4 * t / 3))) < 1e-14


def test_codon_structure():
    ex = np.zeros(61 * 60 // 2)
    assert L.plk_codon_exch(2.0, 0.3, _d(ex)) == 0
    vals = set(np.round(ex, 12))
    assert vals <= {0.0, 1.0, 2.0, 0.3, 0.6}
    # single-nucleotide neighbours of the 61 sense codons in the standard code: 263 pairs
    assert int((ex > 0).sum()) == 263
    import support as S
    ref, _ = S.codon_exch(2.0, 0.3)
    np.testing.assert_allclose(ex, ref, rtol=0, atol=0)


def test_tip_lh():
    rng = np.random.default_rng(1)
    for n, st in [(4, 0), (20, 1), (61, 2)]:
        Ui = rng.normal(size=(n, n))
        su = L.plk_state_unknown(n, st)
        out = np.zeros((su + 1) * n)
        assert L.plk_tip_lh(n, st, _d(np.ascontiguousarray(Ui.ravel())), _d(out)) == 0
        out = out.reshape(su + 1, n)
        import support as S
        prob = S.Problem(n=n, seq_type=st)
        for s in range(su + 1):
            np.testing.assert_allclose(out[s], Ui @ S.tip_likelihood(prob, s), atol=1e-13)


@pytest.mark.parametrize("seq,T,sites", [(0, 32, 2000), (1, 20, 500), (2, 8, 200)])
def test_simulator(seq, T, sites):
    p = G.Problem(seq, T, sites, seed=5)
    assert p.ntaxa == T and p.nsites == sites
    assert abs(p.freq.sum() - sites) < 1e-9
    assert p.states.shape == (T, p.nptn) and p.states.max() < p.n
    # tree: leaves degree 1, internal degree 3, symmetric adjacency
    for v in range(2 * T - 2):
        deg = (p.nei[v] >= 0).sum()
        assert deg == (1 if v < T else 3)
        for k in range(3):
            x = p.nei[v, k]
            if x >= 0:
                kk = list(p.nei[x]).index(v)
                assert p.blen[x, kk] == p.blen[v, k] > 0
    # same seed -> same problem
    q = G.Problem(seq, T, sites, seed=5)
    assert np.array_equal(p.states, q.states) and np.array_equal(p.blen, q.blen)
    assert abs(p.pi.sum() - 1) < 1e-12


def _lib_eigen(Q, pi):
    n = len(pi)
    ev, U, Ui = np.zeros(n), np.zeros(n * n), np.zeros(n * n)
    Qc = np.ascontiguousarray(Q, dtype=np.float64).ravel()
    pic = np.ascontiguousarray(pi, dtype=np.float64)
    dp = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))  # noqa: E731
    assert G.lib().plk_eigen(n, dp(Qc), dp(pic), dp(ev), dp(U), dp(Ui)) == 0, G.lib().plk_last_error()
    return ev, U.reshape(n, n), Ui.reshape(n, n)


@pytest.mark.parametrize("n,zeros", [(4, [2]), (20, [3, 17]), (61, [0, 30, 60])])
def test_zero_frequency_states(n, zeros):
    """States of zero frequency (model/modelmarkov.cpp:1453-1500, 1563-1597): the other
    states' system is decomposed alone, the zero states get eigenvalue 0 and unit vectors.
    Then U exp(L t) U^-1 is expm(Q t) on the non-zero states (scipy), the identity on the
    zero ones, and no non-zero state moves into a zero one."""
    rng = np.random.default_rng(n)
    pi = rng.dirichlet(np.ones(n))
    pi[zeros] = 0.0
    pi /= pi.sum()
    exch = S.codon_exch(2.0, 0.5)[0] if n == 61 else rng.uniform(0.5, 2.0, n * (n - 1) // 2)
    Q = S.rate_matrix(exch, pi)
    ev, U, Ui = _lib_eigen(Q, pi)
    nz = [i for i in range(n) if i not in zeros]
    Qn = Q[np.ix_(nz, nz)]
    for t in (0.05, 0.7):
        P = U @ np.diag(np.exp(ev * t)) @ Ui
        np.testing.assert_allclose(P[np.ix_(nz, nz)], scipy.linalg.expm(Qn * t), atol=1e-12)
        np.testing.assert_allclose(P[np.ix_(zeros, range(n))], np.eye(n)[zeros], atol=1e-14)
        np.testing.assert_allclose(P[np.ix_(nz, zeros)], 0.0, atol=1e-14)
    assert all(ev[z] == 0.0 for z in zeros)
