"""Per-configuration measurements of the engine (the five workloads of the measurement
plan), printed as one JSON document.

  1  DNA GTR+G4, 32 taxa, 10,000 sites: full traversal + logL, all-branch df/ddf
  2  20-state protein +G4, 100 taxa, 50,000 sites (also bench.py's headline)
  3  300 unlinked partitions on 64 taxa (70% DNA, 30% protein, lognormal lengths)
  4  61-state codon +G4, 48 taxa, 5,000 codons
  5  serving: Poisson stream of logL / Newton requests, 16-512 taxa, 1k-100k sites

Usage: python -m gems_amd.bench_configs [--configs 1,2,3,4,5] [--quick]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

import gems_amd as G  # noqa: E402
from gems_amd import workloads as W  # noqa: E402


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_oracle(p, reps=5, seconds=8.0, threads=None):
    import support as S
    threads = threads or max(1, min(16, os.cpu_count() or 1))
    o = S.Oracle(p, safe=(p.n == 61), nthreads=threads)
    a, b = 0, int(p.nei[0, 0])
    o.lib.orc_clear_all(o.h)
    ref = o.loglik(a, b)
    ts = []
    t0 = time.perf_counter()
    while len(ts) < 2 or (len(ts) < reps and time.perf_counter() - t0 < seconds):
        o.lib.orc_clear_all(o.h)
        t1 = time.perf_counter()
        o.loglik(a, b)
        ts.append(time.perf_counter() - t1)
    med = float(np.median(ts))
    o.close()
    return ref, (p.ntaxa - 2) * p.nptn / med, threads


def traversal_rate(p, precision, steps, scaling=G.NORM):
    g = G.Tree(p, precision=precision, scaling=scaling)
    g.full_traversal()
    g.set_timing(True)
    t0 = time.perf_counter()
    for _ in range(steps):
        v = g.full_traversal()
    el = time.perf_counter() - t0
    k_ms, k_n = g.timing(0)
    g.close()
    upd = (p.ntaxa - 2) * p.nptn
    return {"updates_per_s": upd * steps / el, "ms_per_step": el / steps * 1e3,
            "kernel_ms": k_ms / max(k_n, 1), "logl": v}


def all_branch_derivs(p, precision, scaling=G.NORM, reps=20):
    """df/ddf on every branch from scratch: all 3(T-2) directional partials (one traversal
    launch, all-directions slots) + every branch's derivatives (one launch)."""
    g = G.Tree(p, precision=precision, scaling=scaling, all_directions=True)
    lib = G.lib()
    g.all_branch_derivs()
    t0 = time.perf_counter()
    for _ in range(reps):
        lib.plk_tree_clear_all(g.h)
        br, df, ddf = g.all_branch_derivs()
    el = (time.perf_counter() - t0) / reps
    g.close()
    return br, np.stack([df, ddf], axis=1), el


def nni_pass(p, precision, reps=3):
    """Both NNI5 moves of every internal branch (getBestNNIForBran on all of them), batched."""
    g = G.Tree(p, precision=precision, all_directions=2)
    T = p.ntaxa
    ib = [(u, int(v)) for u in range(T, 2 * T - 2) for v in p.nei[u] if v > u]
    g.nni_evaluate(ib)  # warm-up (all directional partials)
    t0 = time.perf_counter()
    for _ in range(reps):
        sc, _ = g.nni_evaluate(ib)
    el = (time.perf_counter() - t0) / reps
    g.close()
    return ib, sc, el


def nni_cpu(p, ib, threads):
    import support as S
    o = S.Oracle(p, nthreads=threads)
    t0 = time.perf_counter()
    sc = np.array([[o.nni_eval(u, v, m)[0] for m in (0, 1)] for (u, v) in ib])
    el = time.perf_counter() - t0
    o.close()
    return sc, el


def param_loop(p, precision, threads=None, gtr=False):
    """ModelFactory::optimizeParameters (branch lengths + Gamma shape from 1.0, and with gtr
    the exchangeabilities from all ones) on the device, the oracle's restatement timed
    beside it on `threads` host cores."""
    import support as S
    g = G.Tree(p, precision=precision)
    g.set_gamma_shape(1.0)
    if gtr:
        g.set_gtr(np.ones(6), p.pi)
    c0 = g.counters()
    t0 = time.perf_counter()
    ag, lg = g.optimize_parameters(logl_epsilon=0.01, gradient_epsilon=1e-4)
    el = time.perf_counter() - t0
    c1 = g.counters()
    g.close()
    res = {"alpha": ag, "logl": lg, "ms": el * 1e3, "partial_updates": int(c1[0] - c0[0]),
           "derivative_evals": int(c1[1] - c0[1])}
    if threads:
        o = S.Oracle(p, nthreads=threads)
        o.set_gamma_shape(1.0)
        t0 = time.perf_counter()
        if gtr:
            o.set_gtr(np.ones(6), p.pi)
            o.optimize_parameters_model(logl_epsilon=0.01, gradient_epsilon=1e-4)
            ao, lo = o.gamma_shape(), o.compute_likelihood()
        else:
            ao, lo = o.optimize_parameters(logl_epsilon=0.01, gradient_epsilon=1e-4)
        res.update({"cpu_oracle_ms": (time.perf_counter() - t0) * 1e3, "cpu_threads": threads, "cpu_alpha": ao,
                    "cpu_logl": lo, "logl_rel_err": abs(lg - lo) / abs(lo), "alpha_rel_err": abs(ag - ao) / ao})
        o.close()
    return res


def config1(quick):
    p = W.config1()
    log(f"config1: {p.nptn} patterns")
    steps = 50 if quick else 400
    r64 = traversal_rate(p, G.FP64, steps)
    r32 = traversal_rate(p, G.FP32, steps)
    ref, cpu, thr = cpu_oracle(p, reps=20)
    import support as S
    o = S.Oracle(p, nthreads=thr)
    br, d64, t64 = all_branch_derivs(p, G.FP64)
    _, d32, t32 = all_branch_derivs(p, G.FP32)
    do = np.array([o.derivs(u, v)[:2] for (u, v) in br])
    e64 = float(np.max(np.abs(d64 - do) / np.maximum(1.0, np.abs(do))))
    e32 = float(np.max(np.abs(d32 - do) / np.maximum(1.0, np.abs(do))))
    ib, n64, tn64 = nni_pass(p, G.FP64)
    _, n32, tn32 = nni_pass(p, G.FP32)
    nco, tnc = nni_cpu(p, ib, thr)
    nni = {"branches": len(ib), "moves": 2 * len(ib), "fp64_ms": tn64 * 1e3, "fp32_ms": tn32 * 1e3,
           "cpu_oracle_ms": tnc * 1e3, "cpu_threads": thr,
           "score_max_rel_err_fp64": float(np.max(np.abs(n64 - nco) / np.abs(nco))),
           "score_max_rel_err_fp32": float(np.max(np.abs(n32 - nco) / np.abs(nco))),
           "best_move_agrees_fp64": bool(np.array_equal(np.argmax(n64, axis=1), np.argmax(nco, axis=1)))}
    params = {"fp64": param_loop(p, G.FP64, thr), "fp32": param_loop(p, G.FP32),
              "fp64_gtr": param_loop(p, G.FP64, thr, gtr=True)}
    return {"patterns": p.nptn, "fp64": r64, "fp32": r32, "cpu_updates_per_s": cpu, "cpu_threads": thr,
            "optimize_parameters": params,
            "logl_rel_err_fp64": abs(r64["logl"] - ref) / abs(ref), "logl_rel_err_fp32": abs(r32["logl"] - ref) / abs(ref),
            "nni5_all_branches": nni,
            "all_branch_derivs_ms": {"fp64": t64 * 1e3, "fp32": t32 * 1e3, "branches": len(br),
                                     "updates_per_pass": 3 * (p.ntaxa - 2) * p.nptn,
                                     "fp64_updates_per_s": 3 * (p.ntaxa - 2) * p.nptn / t64,
                                     "fp32_updates_per_s": 3 * (p.ntaxa - 2) * p.nptn / t32},
            "deriv_max_rel_err": {"fp64": e64, "fp32": e32},
            "fp32_over_fp64": r32["updates_per_s"] / r64["updates_per_s"],
            "fp32_over_cpu": r32["updates_per_s"] / cpu}


def _topology(ntaxa, shape):
    """Unrooted binary topology as plk_sim_create_on_tree takes it ((2T-2) x 3, leaves
    0..T-1 with their neighbour in slot 0): 'caterpillar' (T-2 dependent levels from the
    first leaf's branch) or 'balanced' (about log2 T levels)."""
    nn = 2 * ntaxa - 2
    nei = -np.ones((nn, 3), dtype=np.int32)

    def link(a, b):
        ka = list(nei[a]).index(-1)
        kb = list(nei[b]).index(-1)
        nei[a, ka], nei[b, kb] = b, a

    internals = list(range(ntaxa, nn))
    if shape == "caterpillar":
        link(internals[0], 0)
        link(internals[0], 1)
        for i in range(len(internals) - 1):
            link(internals[i], internals[i + 1])
            link(internals[i + 1], i + 2)
        link(internals[-1], ntaxa - 1)
    else:
        # pair leaves, then pair subtrees, until three remain for the last internal node
        nodes, nxt = list(range(ntaxa)), 0
        while len(nodes) > 3:
            a, b = nodes.pop(0), nodes.pop(0)
            v = internals[nxt]
            nxt += 1
            link(v, a)
            link(v, b)
            nodes.append(v)
        v = internals[nxt]
        for x in nodes:
            link(v, x)
    return nei


def config1_levels(quick):
    """Where a DNA traversal's time goes: the same 30 ops x 6,002 patterns (32 taxa, 10k
    sites) on trees whose post-order has 5, 13 (the Yule tree of config 1) or 30 dependent
    levels, and the Yule tree at 2k / 20k sites; kernel time per traversal (HIP events)."""
    out = {}
    base = G.Problem(G.SEQ_DNA, 32, 10000, mean_len=0.1, alpha=0.5, seed=1)
    cases = [("yule", base.nei, 10000), ("balanced", _topology(32, "balanced"), 10000),
             ("caterpillar", _topology(32, "caterpillar"), 10000), ("yule", base.nei, 2000),
             ("yule", base.nei, 20000)]
    for name, nei, sites in cases:
        p = G.Problem(G.SEQ_DNA, 32, sites, mean_len=0.1, alpha=0.5, seed=1, topology=nei)
        row = {"patterns": p.nptn}
        for prec, fname in ((G.FP64, "fp64"), (G.FP32, "fp32")):
            r = traversal_rate(p, prec, 50 if quick else 300)
            row[fname] = {"kernel_us": r["kernel_ms"] * 1e3, "step_us": r["ms_per_step"] * 1e3}
        out[f"{name}_{sites}"] = row
        log(f"  {name} {sites} sites ({p.nptn} patterns): fp64 kernel {row['fp64']['kernel_us']:.1f} us, "
            f"fp32 {row['fp32']['kernel_us']:.1f} us")
    return out


def rell_and_newton(p, o, thr, quick):
    """UFBoot RELL (1000 replicates) and branch-length Newton on the config-2 problem: the
    device against the oracle's restatements on `thr` host cores (RELL: one core)."""
    import support as S
    rng = np.random.default_rng(7)
    ns = 1000
    w = p.freq / p.freq.sum()
    counts = rng.multinomial(int(p.freq.sum()), w, size=ns).astype(np.float32)
    uni = rng.random(ns)
    a, b = 0, int(p.nei[0, 0])
    g = G.Tree(p, precision=G.FP64)
    lnl = g.loglik(a, b)
    r = G.Rell(counts)
    r.save_tree(g.context(), lnl, uni)
    reps = 20
    t0 = time.perf_counter()
    for _ in range(reps):
        r.save_tree(g.context(), lnl, uni)
    rell_gpu = (time.perf_counter() - t0) / reps
    r.close()
    lo, plo = o.loglik(a, b, with_patterns=True)
    orl = S.OracleRell(counts)
    t0 = time.perf_counter()
    for _ in range(3):
        orl.save_tree(plo, lo, uniform=uni)
    rell_cpu = (time.perf_counter() - t0) / 3
    res = {"rell": {"replicates": ns, "patterns": p.nptn, "gpu_ms_per_tree": rell_gpu * 1e3,
                    "cpu_oracle_ms_per_tree": rell_cpu * 1e3, "cpu_threads": 1,
                    "gpu_bytes_per_s_counts": ns * p.nptn * 4 / rell_gpu}}
    # Newton: all branches once (optimizeAllBranches with one iteration), device loop
    iters = 1
    g.full_traversal()
    t0 = time.perf_counter()
    lg = g.optimize_all(iterations=iters, tol=0.01)
    newton_gpu = time.perf_counter() - t0
    c = g.counters()
    g.close()
    res["optimize_all_branches_1_round"] = {"branches": 2 * p.ntaxa - 3, "gpu_ms": newton_gpu * 1e3,
                                            "derivative_evals": int(c[1]), "logl_gpu": lg}
    if not quick:
        t0 = time.perf_counter()
        lc = o.optimize_all(iterations=iters, tol=0.01)
        res["optimize_all_branches_1_round"].update({"cpu_oracle_ms": (time.perf_counter() - t0) * 1e3,
                                                     "cpu_threads": thr, "logl_cpu": lc,
                                                     "logl_rel_err": abs(lg - lc) / abs(lc)})
    return res


def config2(quick):
    p = W.config2()
    log(f"config2: {p.nptn} patterns")
    steps = 5 if quick else 20
    r64 = traversal_rate(p, G.FP64, steps)
    r32 = traversal_rate(p, G.FP32, steps)
    ref, cpu, thr = cpu_oracle(p, reps=5)
    import support as S
    o = S.Oracle(p, nthreads=thr)
    br = [(0, int(p.nei[0, 0])), (100, int(p.nei[100, 0]))]
    errs = {}
    for prec, name in ((G.FP64, "fp64"), (G.FP32, "fp32")):
        g = G.Tree(p, precision=prec)
        e = 0.0
        for (u, v) in br:
            dg = g.derivs(u, v)[:2]
            dd = o.derivs(u, v)[:2]
            e = max(e, max(abs(dg[i] - dd[i]) / max(1.0, abs(dd[i])) for i in range(2)))
        errs[name] = e
        g.close()
    extras = rell_and_newton(p, o, thr, quick)
    o.close()
    return {"patterns": p.nptn, "fp64": r64, "fp32": r32, "cpu_updates_per_s": cpu, "cpu_threads": thr,
            "logl_rel_err_fp64": abs(r64["logl"] - ref) / abs(ref), "logl_rel_err_fp32": abs(r32["logl"] - ref) / abs(ref),
            "deriv_max_rel_err": errs, "fp32_over_fp64": r32["updates_per_s"] / r64["updates_per_s"],
            "fp32_over_cpu": r32["updates_per_s"] / cpu, **extras}


def config3(quick, nparts=300):
    probs = W.config3(nparts)
    upd = sum((p.ntaxa - 2) * p.nptn for p in probs)
    log(f"config3: {nparts} partitions, {sum(p.nptn for p in probs)} patterns, {upd} updates per traversal")
    br = [(0, int(p.nei[0, 0])) for p in probs]
    out = {"partitions": nparts, "patterns": int(sum(p.nptn for p in probs)), "updates_per_traversal": upd,
           "dna_partitions": int(sum(p.n == 4 for p in probs))}
    steps = 5 if quick else 20
    for prec, name in ((G.FP64, "fp64"), (G.FP32, "fp32")):
        trees = [G.Tree(p, precision=prec) for p in probs]
        # per-partition launches (the reference's loop over partition trees, one after another)
        t0 = time.perf_counter()
        for _ in range(max(1, steps // 5)):
            seq_vals = [t.full_traversal() for t in trees]
        t_seq = (time.perf_counter() - t0) / max(1, steps // 5)
        b = G.Batch(trees)
        grid, tune_ms = b.autotune(br, reps=2)
        vals, _ = b.loglik(br, clear=True)
        same = bool(np.array_equal(vals, np.array(seq_vals)))
        t0 = time.perf_counter()
        for _ in range(steps):
            vals, tot = b.loglik(br, clear=True)
        el = (time.perf_counter() - t0) / steps
        ms, n, _ = b.timing()
        out[name] = {"batched_ms": el * 1e3, "batched_updates_per_s": upd / el, "sequential_ms": t_seq * 1e3,
                     "sequential_updates_per_s": upd / t_seq, "speedup_vs_sequential": t_seq / el,
                     "autotuned_grid": grid, "work_items": b.items(), "batched_equals_alone": same,
                     "total_logl": float(tot)}
        b.close()
        for t in trees:
            t.close()
    return out


def config4(quick):
    p = W.config4()
    log(f"config4: {p.nptn} patterns")
    steps = 3 if quick else 10
    r64 = traversal_rate(p, G.FP64, steps, scaling=G.SAFE)
    r32 = traversal_rate(p, G.FP32, steps)
    ref, cpu, thr = cpu_oracle(p, reps=3)
    import support as S
    o = S.Oracle(p, safe=True, nthreads=thr)
    u, v = 0, int(p.nei[0, 0])
    dd = o.derivs(u, v)[:2]
    errs = {}
    for prec, name, sc in ((G.FP64, "fp64", G.SAFE), (G.FP32, "fp32", G.NORM)):
        g = G.Tree(p, precision=prec, scaling=sc)
        dg = g.derivs(u, v)[:2]
        errs[name] = max(abs(dg[i] - dd[i]) / max(1.0, abs(dd[i])) for i in range(2))
        g.close()
    flops = 4 * (4 * 61 * 61 + 61) * (p.ntaxa - 2) * p.nptn
    return {"patterns": p.nptn, "fp64": r64, "fp32": r32, "cpu_updates_per_s": cpu, "cpu_threads": thr,
            "logl_rel_err_fp64": abs(r64["logl"] - ref) / abs(ref), "logl_rel_err_fp32": abs(r32["logl"] - ref) / abs(ref),
            "deriv_max_rel_err": errs,
            "fp32_alg_tflops": flops / (r32["kernel_ms"] * 1e-3) / 1e12,
            "fp32_over_fp64": r32["updates_per_s"] / r64["updates_per_s"]}


def config5(quick, nreq=None):
    nreq = nreq or (16 if quick else 40)
    specs, rng = W.config5_specs(nreq)
    t0 = time.perf_counter()
    probs = [W.config5_problem(sp) for sp in specs]
    log(f"config5: generated {nreq} requests in {time.perf_counter() - t0:.1f}s")
    updates = sum((p.ntaxa - 2) * p.nptn for p in probs)
    out = {"requests": nreq, "updates_per_pass": updates, "sweep": {}}
    rates = [5.0, 20.0, 80.0, 1e9] if not quick else [20.0, 1e9]
    for prec, name in ((G.FP64, "fp64"), (G.FP32, "fp32")):
        res = []
        for rate in rates:
            reqs = [G.Request(p, kind=G.REQ_LOGL if k == 0 else G.REQ_OPT_BRANCH, precision=prec,
                              opt_branch=(T, int(p.nei[T, 0])))
                    for p, (s, T, n, k, sd) in zip(probs, specs)]
            gaps = rng.exponential(1.0 / rate, size=nreq)
            arrivals = np.cumsum(gaps) - gaps[0]
            srv = G.Server(nworkers=8, hbm_budget=200 << 30)
            t0 = srv.now()
            for r, at in zip(reqs, arrivals):
                delay = t0 + at - srv.now()
                if delay > 0:
                    time.sleep(delay)
                srv.submit(r)
            srv.drain()
            st = srv.stats()
            srv.close()
            res.append({"offered_rate": rate if rate < 1e8 else "saturation", "p50_s": st["p50_s"],
                        "p99_s": st["p99_s"], "mean_s": st["mean_s"], "makespan_s": st["makespan_s"],
                        "throughput_req_s": nreq / st["makespan_s"], "updates_per_s": updates / st["makespan_s"],
                        "failed": st["failed"], "peak_concurrent": st["peak_concurrent"],
                        "peak_hbm_gb": st["peak_hbm_bytes"] / 1e9})
            log(f"  {name} rate {rate}: p50 {st['p50_s'] * 1e3:.1f} ms p99 {st['p99_s'] * 1e3:.1f} ms "
                f"thr {nreq / st['makespan_s']:.1f} req/s")
        out["sweep"][name] = res
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--configs", default="1,2,3,4,5")
    ap.add_argument("--quick", action="store_true")
    ap.add_argument("--out", default=None)
    args = ap.parse_args()
    fns = {"1": config1, "2": config2, "3": config3, "4": config4, "5": config5, "1L": config1_levels}
    res = {}
    for c in args.configs.split(","):
        t0 = time.perf_counter()
        res["config" + c] = fns[c](args.quick)
        res["config" + c]["wall_s"] = time.perf_counter() - t0
        log(json.dumps({"config" + c: res["config" + c]}))
    txt = json.dumps(res, indent=1)
    if args.out:
        with open(args.out, "w") as f:
            f.write(txt)
    print(txt)


if __name__ == "__main__":
    main()
