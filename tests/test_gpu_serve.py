"""Serving loop: requests of mixed taxa/sites/state counts admitted by HBM footprint,
their traversals batched into shared launches (continuous batching) and their Newton
phases overlapping on worker streams; every result must equal the same request run
alone."""
import numpy as np
import pytest

import gems_amd as G

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def make_requests(n, seed, precision=G.FP64):
    rng = np.random.default_rng(seed)
    reqs = []
    for i in range(n):
        seq = G.SEQ_DNA if rng.random() < 0.6 else G.SEQ_PROTEIN
        T = int(np.exp(rng.uniform(np.log(8), np.log(48))))
        sites = int(np.exp(rng.uniform(np.log(200), np.log(4000))))
        p = G.Problem(seq, T, sites, mean_len=0.1 if seq == G.SEQ_DNA else 0.2, seed=seed * 100 + i)
        kind = [G.REQ_LOGL, G.REQ_OPT_BRANCH, G.REQ_OPT_ALL][i % 3]
        u = T + int(rng.integers(T - 2))
        reqs.append(G.Request(p, kind=kind, precision=precision, opt_branch=(u, int(p.nei[u, 0])), iterations=2))
    return reqs


@pytest.mark.parametrize("prec", [G.FP64, G.FP32], ids=["fp64", "fp32"])
def test_alone_equals_mixed(prec):
    alone = make_requests(12, 5, prec)
    mixed = make_requests(12, 5, prec)
    s1 = G.Server(nworkers=1)
    for r in alone:
        s1.submit(r)
        s1.drain()
    s1.close()
    s = G.Server(nworkers=6, hbm_budget=8 << 30)
    for r in mixed:
        s.submit(r)
    s.drain()
    st = s.stats()
    cn = s.counters()
    say("stats", st, cn)
    s.close()
    assert st["completed"] == 12 and st["failed"] == 0
    assert st["peak_concurrent"] >= 2
    # continuous batching: traversals of several requests shared launches, trees were reused
    assert cn["batched"] > cn["launches"]
    assert cn["pooled_trees"] < 12
    for a, m in zip(alone, mixed):
        ra, rm = a.result(), m.result()
        assert ra["status"] == 0 and rm["status"] == 0
        assert ra["logl_before"] == rm["logl_before"]
        assert ra["logl_after"] == rm["logl_after"]
        assert ra["opt_len"] == rm["opt_len"]
        assert rm["logl_after"] >= rm["logl_before"] - 1e-6 * abs(rm["logl_before"])


def test_admission_respects_budget():
    reqs = make_requests(8, 9)
    biggest = max(r.r.nptn for r in reqs)
    import gems_amd as GG
    fp = [GG.lib().plk_memory_required(r.r.nstates, r.r.ncat, r.r.ntaxa, r.r.nptn, r.r.ntaxa - 2, 0, 0) for r in reqs]
    budget = max(fp) + 1
    s = G.Server(nworkers=8, hbm_budget=budget)
    for r in reqs:
        s.submit(r)
    s.drain()
    st = s.stats()
    s.close()
    assert st["failed"] == 0
    assert st["peak_hbm_bytes"] <= budget
    # an impossible request fails cleanly instead of blocking the queue
    big = make_requests(1, 11)[0]
    s = G.Server(nworkers=2, hbm_budget=1 << 20)
    s.submit(big)
    s.drain()
    assert big.result()["status"] == -5
    s.close()


@pytest.mark.parametrize("prec", [G.FP64, G.FP32], ids=["fp64", "fp32"])
def test_tree_reshape_equals_fresh_tree(prec):
    """A pooled tree reshaped to another problem (larger, then smaller, then a different
    taxon count) gives exactly the logL and Newton result of a fresh tree."""
    import ctypes
    probs = [G.Problem(G.SEQ_PROTEIN, 12, 400, mean_len=0.2, seed=1), G.Problem(G.SEQ_PROTEIN, 30, 3000, seed=2),
             G.Problem(G.SEQ_PROTEIN, 9, 150, seed=3), G.Problem(G.SEQ_PROTEIN, 40, 900, seed=4)]
    pooled = G.Tree(probs[0], precision=prec)
    first = True
    for p in probs:
        if not first:
            G._check(G.lib().plk_tree_reshape(pooled.h, p.ntaxa, p.nptn))
            pooled.p = p
            pooled.set_model(p)
            st = np.ascontiguousarray(p.states, dtype=np.uint8)
            fr = np.ascontiguousarray(p.freq, dtype=np.float64)
            G._check(G.lib().plk_tree_set_patterns(pooled.h, st.ctypes.data_as(G._u8p), G._d(fr), None))
            nei = np.ascontiguousarray(p.nei, dtype=np.int32)
            bl = np.ascontiguousarray(p.blen, dtype=np.float64)
            G._check(G.lib().plk_tree_set_topology(pooled.h, nei.ctypes.data_as(G._ip), G._d(bl)))
        first = False
        fresh = G.Tree(p, precision=prec)
        a, b = 0, int(p.nei[0, 0])
        assert pooled.loglik(a, b) == fresh.loglik(a, b)
        u = p.ntaxa
        assert pooled.optimize_one(u, int(p.nei[u, 0])) == fresh.optimize_one(u, int(p.nei[u, 0]))
        fresh.close()
    pooled.close()
