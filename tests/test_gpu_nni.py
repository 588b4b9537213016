"""Many-branch Newton (plk_newton_multi) and batched NNI5 evaluation
(plk_tree_nni_evaluate), through the C ABI, against the single-branch device Newton and
the oracle's getBestNNIForBran restatement (orc_nni_eval)."""
import numpy as np
import pytest

import gems_amd as G
import support as S

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def internal_branches(p):
    T = p.ntaxa
    return [(u, int(v)) for u in range(T, 2 * T - 2) for v in p.nei[u] if v > u]


def rel(a, b):
    return abs(a - b) / max(abs(b), 1e-300)


@pytest.mark.parametrize("prec,safe", [(G.FP64, 0), (G.FP64, 1), (G.FP32, 0)], ids=["f64", "f64-safe", "f32"])
def test_newton_multi_matches_single(prec, safe):
    """Every branch solved in one lockstep batch gives the lengths the one-branch device
    solve gives (same derivative arithmetic, same step logic)."""
    p = G.Problem(G.SEQ_DNA, 14, 2000, seed=23)
    say("newton multi", prec, safe)
    g = G.Tree(p, precision=prec, scaling=safe, all_directions=True)
    br, _, _ = g.all_branch_derivs()  # every direction computed
    ctx = g.context()
    T = p.ntaxa

    def ref(frm, to):
        return -(to + 1) if to < T else 3 * (to - T) + [int(x) for x in p.nei[to]].index(frm)

    refs = [(ref(v, u), ref(u, v)) for (u, v) in br]
    guesses = [g.get_length(u, v) for (u, v) in br]
    out, logl, st = g.newton_multi(refs, guesses)
    assert np.all(st == 0)
    import ctypes
    for i, (a, b) in enumerate(refs):
        one = ctypes.c_double()
        assert G.lib().plk_newton_branch(ctx, a, b, 1e-6, guesses[i], 10.0, 1e-6, 100, ctypes.byref(one), None) == 0
        if safe:
            assert abs(out[i] - one.value) <= 1e-9 * max(1.0, one.value)
        else:
            assert out[i] == one.value, (i, out[i], one.value)
        # logl is the logL at the answer
        lt = ctypes.c_double()
        assert G.lib().plk_loglik_from_theta(ctx, out[i], ctypes.byref(lt)) == 0
        assert rel(logl[i], lt.value) <= 1e-12
    # eval-only mode returns the logL at the guesses (from theta in FP64; the FP32 form's
    # branch kernel sums in FP32, so it agrees to the FP32 form's 1e-6 bar)
    _, l0, _ = g.newton_multi(refs, guesses, max_nr=-1)
    lg = g.loglik(*br[0])
    assert rel(l0[0], lg) <= (1e-12 if prec == G.FP64 else 1e-6)


@pytest.mark.parametrize("prec", [G.FP64, G.FP32], ids=["f64", "f32"])
def test_nni_matches_oracle(prec):
    p = G.Problem(G.SEQ_DNA, 16, 3000, seed=29)
    say("nni", prec, "nptn", p.nptn)
    g = G.Tree(p, precision=prec, all_directions=2)
    before = g.loglik(0, int(p.nei[0, 0]))
    ib = internal_branches(p)
    scores, lens = g.nni_evaluate(ib)
    o = S.Oracle(p)
    tol = 1e-9 if prec == G.FP64 else 1e-6
    for i, (u, v) in enumerate(ib):
        for m in (0, 1):
            so, lo = o.nni_eval(u, v, m)
            assert rel(scores[i, m], so) <= tol, (u, v, m, scores[i, m], so)
            assert np.all(np.abs(lens[i, m] - lo) <= (2e-6 if prec == G.FP64 else 1e-3) * np.maximum(1.0, lo)), (u, v, m)
    # the tree itself is unchanged
    assert g.loglik(0, int(p.nei[0, 0])) == before
    for (u, v) in ib[:3]:
        assert g.get_length(u, v) == p.blen[u][list(p.nei[u]).index(v)]


def test_nni_protein_safe_batch():
    p = G.Problem(G.SEQ_PROTEIN, 12, 800, seed=31)
    g = G.Tree(p, scaling=1, all_directions=2)
    ib = internal_branches(p)
    scores, _ = g.nni_evaluate(ib)
    o = S.Oracle(p, safe=True)
    for i, (u, v) in enumerate(ib[:4]):
        for m in (0, 1):
            assert rel(scores[i, m], o.nni_eval(u, v, m)[0]) <= 1e-9
