"""Level-parallel 4-state kernel (plk_set_level_waves) against the category-split kernel
and the oracle.  FP64: logL, partials and scale counts bit-identical to the
category-split kernel for every wave count; FP32: both within 1e-6 of the oracle."""
import ctypes

import numpy as np
import pytest

import gems_amd as G
import support as S

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def branches(p):
    out = []
    for u in range(2 * p.ntaxa - 2):
        for k in range(3):
            v = int(p.nei[u, k])
            if v > u:
                out.append((u, v))
    return out


def rel(a, b):
    return abs(a - b) / max(abs(b), 1e-300)


@pytest.mark.parametrize("safe", [0, 1])
@pytest.mark.parametrize("T,sites,ml", [(32, 10000, 0.1), (120, 700, 0.6), (300, 200, 0.3)])
def test_fp64_identical_to_category_kernel(T, sites, ml, safe):
    p = G.Problem(G.SEQ_DNA, T, sites, mean_len=ml, alpha=0.5, seed=T + safe)
    say("level fp64", T, sites, safe, "nptn", p.nptn)
    ref = G.Tree(p, scaling=safe)
    ref.set_level_waves(0)
    br = branches(p)
    want = [ref.loglik(u, v, with_patterns=True) for (u, v) in br[:: max(1, len(br) // 6)]]
    inner = [v for v in range(T, 2 * T - 2, max(1, T // 5))]
    want_part = [ref.partial(int(p.nei[v, 0]), v, safe=bool(safe)) for v in inner]
    for waves in (1, 2, 3, 4):
        g = G.Tree(p, scaling=safe)
        g.set_level_waves(waves)
        got = [g.loglik(u, v, with_patterns=True) for (u, v) in br[:: max(1, len(br) // 6)]]
        for (a, pa), (b, pb) in zip(got, want):
            assert a == b and np.array_equal(pa, pb), waves
        for v, (pw, sw) in zip(inner, want_part):
            pg, sg = g.partial(int(p.nei[v, 0]), v, safe=bool(safe))
            assert np.array_equal(pg, pw) and np.array_equal(sg, sw), (waves, v)
    o = S.Oracle(p, safe=bool(safe))
    u, v = br[0]
    assert rel(got[0][0], o.loglik(u, v)) <= 1e-10


def test_fp32_level_kernel_vs_oracle():
    """The FP32 level kernel forms P(t) and the tip rows with the P(t) kernel's own
    arithmetic, so it matches the category-split FP32 kernel bit for bit."""
    p = G.Problem(G.SEQ_DNA, 64, 5000, mean_len=0.15, alpha=0.5, seed=8)
    o = S.Oracle(p, nthreads=4)
    got = {}
    for waves in (0, 2, 4):
        g = G.Tree(p, precision=G.FP32)
        g.set_level_waves(waves)
        got[waves] = [g.loglik(u, v, with_patterns=True) for (u, v) in branches(p)[:4]]
        for (u, v), (l, _) in zip(branches(p)[:4], got[waves]):
            assert rel(l, o.loglik(u, v)) <= 1e-6, (waves, u, v)
    for waves in (2, 4):
        for (a, pa), (b, pb) in zip(got[waves], got[0]):
            assert a == b and np.array_equal(pa, pb), waves


def test_deep_caterpillar_forced_scaling_and_exponents():
    T = 300
    rng = np.random.default_rng(5)
    pp = S.make_problem(4, T, 150, rng, tree=S.caterpillar_tree(T, 1.5), ambiguous=0.05)
    o = S.Oracle(pp)
    a, b = 0, int(pp.nei[0, 0])
    lo = o.loglik(a, b)
    for prec, tol in ((G.FP64, 1e-10), (G.FP32, 1e-6)):
        g = G.Tree(pp, precision=prec)
        g.set_level_waves(4)
        assert rel(g.loglik(a, b), lo) <= tol, prec


def test_op_list_with_slot_reuse():
    """Raw op lists may rewrite a slot after it was read (write-after-read) or rewrite it
    twice (write-after-write); the level sort keeps the list's sequential meaning."""
    p = G.Problem(G.SEQ_DNA, 6, 500, seed=4)
    vals = {}
    for waves in (0, 4):
        g = G.Tree(p)
        g.set_level_waves(waves)
        ctx = g.context()
        ops = (G.plk_op * 4)()

        def setop(i, out, c0, c1, l0, l1):
            ops[i].out_slot, ops[i].child[0], ops[i].child[1] = out, c0, c1
            ops[i].reserved = 0
            ops[i].len[0], ops[i].len[1] = l0, l1

        setop(0, 0, -1, -2, 0.1, 0.2)  # slot0 <- tips 0, 1
        setop(1, 1, 0, -3, 0.3, 0.1)   # slot1 <- slot0, tip 2
        setop(2, 0, -4, -5, 0.2, 0.2)  # slot0 rewritten after it was read
        setop(3, 2, 1, 0, 0.15, 0.25)  # reads slot1 and the new slot0
        assert G.lib().plk_update_partials(ctx, ops, 4) == 0
        v = ctypes.c_double()
        assert G.lib().plk_branch_loglik(ctx, 2, -6, 0.05, ctypes.byref(v), None) == 0
        vals[waves] = v.value
    assert vals[0] == vals[4]


def test_all_directions_and_multi_derivs():
    p = G.Problem(G.SEQ_DNA, 40, 3000, seed=17)
    res = {}
    for waves in (0, 4):
        g = G.Tree(p, all_directions=True)
        g.set_level_waves(waves)
        g.full_traversal()
        res[waves] = g.all_branch_derivs()
    assert res[0][0] == res[4][0]
    assert np.array_equal(res[0][1], res[4][1]) and np.array_equal(res[0][2], res[4][2])
