"""Op order independence: a traversal queued as one op list (whose tiles follow two
different valid orders of the list) leaves partials, scale counts and exponents
bit-identical to the same ops run one launch at a time in list order."""
import ctypes

import numpy as np
import pytest

import gems_amd as G

pytestmark = pytest.mark.gpu


def post_order_ops(p, root):
    """Ops for every internal node but the root, children before parents; node v's
    partial (towards its parent) lives in slot v - ntaxa."""
    T = p.ntaxa
    ops = []

    def ref(u):
        return -(u + 1) if u < T else u - T

    def visit(v, parent):
        kids = [(int(u), float(p.blen[v][k])) for k, u in enumerate(p.nei[v]) if int(u) != parent]
        for u, _ in kids:
            if u >= T:
                visit(u, v)
        if parent >= 0:
            ops.append((v - T, ref(kids[0][0]), ref(kids[1][0]), kids[0][1], kids[1][1]))

    visit(root, -1)
    return ops


def run(p, prec, safe, ops, one_by_one):
    g = G.Tree(p, precision=prec, scaling=safe)
    ctx = g.context()
    arr = (G.plk_op * len(ops))()
    for i, (out, c0, c1, l0, l1) in enumerate(ops):
        arr[i].out_slot, arr[i].child[0], arr[i].child[1], arr[i].reserved = out, c0, c1, 0
        arr[i].len[0], arr[i].len[1] = l0, l1
    if one_by_one:
        for i in range(len(ops)):
            assert G.lib().plk_update_partials(ctx, ctypes.byref(arr[i]), 1) == 0
            assert G.lib().plk_flush(ctx) == 0
    else:
        assert G.lib().plk_update_partials(ctx, arr, len(ops)) == 0
        assert G.lib().plk_flush(ctx) == 0
    n, ncat, nptn = p.n, len(p.rates), p.nptn
    res = []
    for (out, *_rest) in ops:
        part = np.zeros(nptn * n * ncat)
        sc = np.zeros(nptn * ncat, dtype=np.uint16)
        scp = sc.ctypes.data_as(ctypes.POINTER(ctypes.c_uint16)) if prec == G.FP64 else None
        assert G.lib().plk_get_partial(ctx, out, part.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), scp) == 0
        ex = np.zeros(nptn, dtype=np.int32)
        if prec == G.FP32:
            assert G.lib().plk_get_exponents(ctx, out, ex.ctypes.data_as(ctypes.POINTER(ctypes.c_int))) == 0
        res.append((part, sc, ex))
    g.close()
    return res


@pytest.mark.parametrize("prec,safe", [(G.FP64, 0), (G.FP64, 1), (G.FP32, 0)], ids=["f64", "f64-safe", "f32"])
def test_two_orders_bit_identical(prec, safe):
    p = G.Problem(G.SEQ_PROTEIN, 40, 1500, mean_len=0.3, seed=43)
    print("order", prec, safe, "nptn", p.nptn, flush=True)
    ops = post_order_ops(p, p.ntaxa)
    assert len(ops) == p.ntaxa - 3
    a = run(p, prec, safe, ops, one_by_one=False)
    b = run(p, prec, safe, ops, one_by_one=True)
    for (pa, sa, ea), (pb, sb, eb) in zip(a, b):
        assert np.array_equal(pa, pb) and np.array_equal(sa, sb) and np.array_equal(ea, eb)
