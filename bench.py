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


PROFILE = os.path.join(ROOT, "profiles", "k_traverse_f32_mfma_protein20_100taxa_50k.md")


def profiled_traffic():
    """HBM bytes per launch of the benchmark kernel, from the committed rocprofv3 summary
    (FETCH_SIZE and WRITE_SIZE counter passes, each scaled by what the counter reads for a
    1 GiB device copy in the same run)."""
    try:
        with open(PROFILE) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return int(line.split(":", 1)[1])
    except (OSError, ValueError):
        pass
    return None


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

    def run_form(precision, label):
        g = G.Tree(p, precision=precision, scaling=G.NORM)
        tile = 1
        if precision == G.FP32:
            tile, _ = g.autotune(reps=2)  # matrix-core tile shape (waves per category)
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
        g.close()
        ups = upd_per_step * args.steps / el
        k_avg_s = k_ms / max(k_n, 1) * 1e-3
        w = 8 if precision == G.FP64 else 4
        bpu = 2 * p.ncat * p.n * w
        achieved = upd_per_step * bpu / k_avg_s / 1e9
        log(f"{label}: {el / args.steps * 1e3:.3f} ms/step, kernel {k_avg_s * 1e3:.3f} ms, {ups / 1e9:.3f} G upd/s, "
            f"{achieved:.0f} GB/s algorithmic, logL {logl:.6f}")
        return {"ups": ups, "ms_per_step": el / args.steps * 1e3, "kernel_avg_ms": k_avg_s * 1e3,
                "achieved": achieved, "bytes_per_update": bpu, "logl": logl, "tile_waves_per_category": tile}

    def f32_kernel_name(tile):
        """The kernel the FP32 launch of this problem runs (the rule of traverse_f32_n)."""
        if tile == 0:
            return "k_traverse_f32_pbw<20,MODE_LOGL>"
        ntiles, ncu = (p.nptn + 63) // 64, G.device_units()
        if tile == 1 and ncu > 0:
            per, extra = ntiles // ncu, ntiles % ncu
            if 1 <= per <= 3 and 0 < extra and 4 * extra <= ncu:
                return (f"k_traverse_f32_mfma_split<20,MODE_LOGL> ({ntiles - extra} whole tiles + "
                        f"{4 * extra} single-block units on {ncu} compute units)")
        return f"k_traverse_f32_mfma<20,MODE_LOGL,wpc={tile}>"

    traffic = profiled_traffic() if (args.taxa, args.sites) == (100, 50000) else None
    copy_gbps = G.probe_copy(1 << 30, 5)
    log(f"measured device copy bandwidth: {copy_gbps:.0f} GB/s")

    f64 = run_form(G.FP64, "fp64")
    f32 = run_form(G.FP32, "fp32")

    cpu = None
    err64 = err32 = None
    if not args.no_cpu:
        ref, cpu = cpu_baseline(p)
        err64 = abs(f64["logl"] - ref) / abs(ref)
        err32 = abs(f32["logl"] - ref) / abs(ref)
        log(f"cpu: {cpu['value'] / 1e9:.4f} G upd/s on {cpu['cores']} threads; |dlogL|/|logL| fp64 {err64:.2e}, "
            f"fp32 {err32:.2e}")

    out = {
        "metric": METRIC,
        "value": f32["ups"],
        "unit": "updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": f32["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded simulation, resident in HBM)",
        "config": {
            "workload": "protein20_G4_100taxa_50000sites_full_traversal",
            "taxa": p.ntaxa, "sites": p.nsites, "patterns": p.nptn, "states": p.n, "ncat": p.ncat,
            "form": "fp32 partials on matrix cores, fp64 P(t)/exponents/logL reduction",
            "updates_per_step": upd_per_step,
            "logl_fp32": f32["logl"], "logl_rel_err_fp32_vs_oracle": err32,
            "fp64": {"value": f64["ups"], "ms_per_step": f64["ms_per_step"], "logl": f64["logl"],
                     "logl_rel_err_vs_oracle": err64, "kernel": "k_traverse_f64<20,MODE_LOGL>",
                     "kernel_avg_ms": f64["kernel_avg_ms"], "achieved_GBps": f64["achieved"],
                     "roofline_frac": f64["achieved"] / HBM_PEAK_GBS},
            "fp32_over_fp64": f32["ups"] / f64["ups"],
            "gpu_fp32_over_cpu": (f32["ups"] / cpu["value"]) if cpu else None,
        },
        "roofline": {
            "bound": "hbm", "achieved": f32["achieved"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": f32["achieved"] / HBM_PEAK_GBS, "traffic": traffic,
            "traffic_source": os.path.relpath(PROFILE, ROOT) if traffic else None,
            "traffic_over_algorithmic": traffic / (upd_per_step * f32["bytes_per_update"]) if traffic else None,
            "kernel": f32_kernel_name(f32["tile_waves_per_category"]), "kernel_avg_ms": f32["kernel_avg_ms"],
            "alg_bytes_per_update": f32["bytes_per_update"],
            "measured_copy_GBps": copy_gbps, "frac_of_measured_copy": f32["achieved"] / copy_gbps,
            "tile_waves_per_category": f32["tile_waves_per_category"],
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
