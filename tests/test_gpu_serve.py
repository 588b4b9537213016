"""Serving loop: requests of mixed taxa/sites/state counts admitted by HBM footprint
onto concurrent workers; every result must equal the same request run alone."""
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
    say("stats", st)
    s.close()
    assert st["completed"] == 12 and st["failed"] == 0
    assert st["peak_concurrent"] >= 2
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
