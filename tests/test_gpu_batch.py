"""Batched unlinked partitions: one launch per traversal for many partitions (mixed DNA
and protein, ragged lengths) must give, per partition, exactly the value of evaluating
it alone, and match the oracle; graph replays must repeat the eager result."""
import numpy as np
import pytest

import gems_amd as G
import support as S

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def make_partitions(nparts, ntaxa=24, seed=0, sizes=None):
    base = G.Problem(G.SEQ_DNA, ntaxa, 50, seed=seed + 1000)
    rng = np.random.default_rng(seed)
    probs = []
    for i in range(nparts):
        seq = G.SEQ_DNA if rng.random() < 0.7 else G.SEQ_PROTEIN
        nsites = int(sizes[i]) if sizes is not None else int(np.clip(np.round(rng.lognormal(np.log(300), 1.0)), 30, 3000))
        probs.append(G.Problem(seq, ntaxa, nsites, mean_len=0.1 if seq == G.SEQ_DNA else 0.2,
                               alpha=0.5 if seq == G.SEQ_DNA else 0.8, seed=seed * 1000 + i, topology=base.nei))
    return probs


@pytest.mark.parametrize("prec", [G.FP64, G.FP32], ids=["fp64", "fp32"])
def test_batch_equals_alone_and_oracle(prec):
    probs = make_partitions(14, sizes=[1, 63, 64, 65, 200, 1000, 3000, 17, 500, 129, 640, 2000, 7, 300])
    say("batch", prec, [p.nptn for p in probs])
    trees = [G.Tree(p, precision=prec) for p in probs]
    alone = [t.full_traversal() for t in trees]
    batch = G.Batch(trees)
    br = [(0, int(p.nei[0, 0])) for p in probs]
    vals, tot = batch.loglik(br, clear=True)
    assert np.array_equal(vals, np.array(alone)), (vals, alone)
    assert tot == pytest.approx(sum(alone), rel=1e-14)
    tol = 1e-10 if prec == G.FP64 else 1e-6
    for p, v in zip(probs, vals):
        ref = S.Oracle(p, nthreads=2).loglik(0, int(p.nei[0, 0]))
        assert abs(v - ref) <= tol * abs(ref)


def test_batch_graph_replay_and_autotune():
    probs = make_partitions(20, seed=3)
    trees = [G.Tree(p, precision=G.FP64) for p in probs]
    br = [(0, int(p.nei[0, 0])) for p in probs]
    batch = G.Batch(trees)
    batch.set_launch(grid=0, use_graph=False)
    eager, _ = batch.loglik(br, clear=True)
    batch.set_launch(grid=0, use_graph=True)
    for rep in range(4):
        v, _ = batch.loglik(br, clear=True)
        assert np.array_equal(v, eager), rep
        # eager work of the program between replays
        assert trees[rep].full_traversal() == eager[rep]
    g, ms = batch.autotune(br, reps=2)
    say("autotune grid", g, "ms", ms)
    v, _ = batch.loglik(br, clear=True)
    assert np.array_equal(v, eager)
    # a different branch per partition after lengths changed (partial recomputation)
    for t, p in zip(trees, probs):
        u = p.ntaxa
        t.set_length(u, int(p.nei[u, 0]), 0.05)
    v2, _ = batch.loglik(br)
    alone2 = [t.loglik(*b) for t, b in zip(trees, br)]
    assert np.array_equal(v2, np.array(alone2))
