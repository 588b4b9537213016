"""Benchmark: site-node partial-likelihood updates/s of the MI355X engine.

Workload (one step): a full post-order traversal of all T-2 internal partials plus the
branch logL, on the 20-state protein +G4 problem with 100 taxa and 50,000 simulated
sites (Yule tree, Exp(mean 0.2) branch lengths, Gamma(1,1) exchangeabilities,
Dirichlet(1^20) frequencies, alpha 0.8), seeded synthetic data resident in HBM.

Prints one JSON line.  The roofline block is measured live with HIP events on the
engine's stream around every traversal launch; the CPU baseline is the oracle (the
FP64 restatement of the reference kernel) on this host's cores over a bounded sample.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "site·node partial-lh updates/s FP32 vs FP64 at stated logL err; p50/p99 latency"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(p, seconds=10.0, min_reps=2):
    """Time the oracle's full traversal + branch logL on the host cores."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import support as S

    threads = max(1, min(16, os.cpu_count() or 1))
    o = S.Oracle(p, safe=False, nthreads=threads)
    a, b = 0, int(p.nei[0, 0])
    o.lib.orc_clear_all(o.h)
    ref = o.loglik(a, b)  # warm-up
    times = []
    t_start = time.perf_counter()
    while len(times) < min_reps or (time.perf_counter() - t_start < seconds and len(times) < 50):
        o.lib.orc_clear_all(o.h)
        t0 = time.perf_counter()
        o.loglik(a, b)
        times.append(time.perf_counter() - t0)
    med = float(np.median(times))
    upd = (p.ntaxa - 2) * p.nptn / med
    o.close()
    return ref, {
        "value": upd,
        "unit": "updates/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{len(times)} full traversals + branch logL of the bench problem (median), "
                  f"{S.oracle_lib().orc_lanes()}-lane SIMD oracle, OpenMP packets",
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--taxa", type=int, default=100)
    ap.add_argument("--sites", type=int, default=50000)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); "
              "the pattern-sharded multi-GPU path is not part of this build", file=sys.stderr)
        sys.exit(2)

    import gems_amd as G

    if G.device_count() < 1:
        raise SystemExit("bench.py: no GPU visible")
    t0 = time.perf_counter()
    p = G.Problem(G.SEQ_PROTEIN, args.taxa, args.sites, mean_len=0.2, alpha=0.8, ncat=4, seed=1)
    log(f"problem: {p.ntaxa} taxa, {p.nsites} sites, {p.nptn} patterns ({time.perf_counter() - t0:.1f}s)")
    n_ops = p.ntaxa - 2
    upd_per_step = n_ops * p.nptn

    g = G.Tree(p, precision=G.FP64, scaling=G.NORM)
    for _ in range(args.warmup):
        g.full_traversal()
    g.set_timing(True)
    G.lib().plk_synchronize(g.context())
    t0 = time.perf_counter()
    for _ in range(args.steps):
        logl = g.full_traversal()
    G.lib().plk_synchronize(g.context())
    el = time.perf_counter() - t0
    k_ms, k_n = g.timing(0)
    g.set_timing(False)
    ups = upd_per_step * args.steps / el
    k_avg_s = k_ms / max(k_n, 1) * 1e-3
    bytes_per_update = 2 * p.ncat * p.n * 8
    achieved = upd_per_step * bytes_per_update / k_avg_s / 1e9
    log(f"fp64: {el / args.steps * 1e3:.3f} ms/step, kernel {k_avg_s * 1e3:.3f} ms, {ups / 1e9:.3f} G upd/s, "
        f"{achieved:.0f} GB/s algorithmic")

    cpu = None
    rel_err = None
    if not args.no_cpu:
        ref, cpu = cpu_baseline(p)
        rel_err = abs(logl - ref) / abs(ref)
        log(f"cpu: {cpu['value'] / 1e9:.4f} G upd/s on {cpu['cores']} threads; |dlogL|/|logL| = {rel_err:.2e}")

    out = {
        "metric": METRIC,
        "value": ups,
        "unit": "updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": el / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded simulation, resident in HBM)",
        "config": {
            "workload": "protein20_G4_100taxa_50000sites_full_traversal",
            "taxa": p.ntaxa, "sites": p.nsites, "patterns": p.nptn, "states": p.n, "ncat": p.ncat,
            "form": "fp64", "updates_per_step": upd_per_step, "logl": logl, "logl_rel_err_vs_oracle": rel_err,
        },
        "roofline": {
            "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": None,
            "kernel": "k_traverse_f64<20,MODE_LOGL>", "kernel_avg_ms": k_avg_s * 1e3,
            "alg_bytes_per_update": bytes_per_update,
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
