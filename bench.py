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
# the oracle's OpenMP team (CPU baseline leg) keeps its threads on neighbouring cores; read
# by the OpenMP runtime when the oracle library is loaded
os.environ.setdefault("OMP_PROC_BIND", "close")

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


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def host_cores():
    """The cores this process may run on (the GPU box's CPU share)."""
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def time_oracle(p, threads, seconds, min_reps=2, max_reps=50, safe=False):
    """Median wall time of the oracle's full traversal + branch logL on `threads` cores."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import support as S

    o = S.Oracle(p, safe=safe, nthreads=threads)
    a, b = 0, int(p.nei[0, 0])
    o.lib.orc_clear_all(o.h)
    ref = o.loglik(a, b)  # warm-up
    times = []
    t_start = time.perf_counter()
    while len(times) < min_reps or (time.perf_counter() - t_start < seconds and len(times) < max_reps):
        o.lib.orc_clear_all(o.h)
        t0 = time.perf_counter()
        o.loglik(a, b)
        times.append(time.perf_counter() - t0)
    o.close()
    return ref, float(np.median(times)), len(times)


def cpu_leg(workload, threads, seconds, max_reps):
    """Time the oracle in a child process that never touches the GPU (so no GPU runtime
    thread shares a core with the oracle's pinned OpenMP team): (reference logL, median s, reps)."""
    import subprocess
    out = subprocess.run([sys.executable, os.path.abspath(__file__), "--cpu-leg", workload, "--threads", str(threads),
                          "--seconds", str(seconds), "--max-reps", str(max_reps)],
                         check=True, capture_output=True, text=True, env=dict(os.environ, OMP_PROC_BIND="close"))
    r = json.loads(out.stdout.strip().splitlines()[-1])
    return r["ref"], r["median_s"], r["reps"]


CPU_LEGS = {}  # workload -> (reference logL, median s, reps), measured before any GPU or oracle use here


def run_cpu_legs(seconds=10.0):
    """All CPU-baseline timings, first thing in the run: each in its own child process, while
    this process has neither a GPU runtime nor an OpenMP team of its own running."""
    cores = host_cores()
    CPU_LEGS["protein"] = cpu_leg("protein", cores, seconds, 50)
    CPU_LEGS["protein_1core"] = cpu_leg("protein", 1, seconds / 2, 3)
    CPU_LEGS["dna"] = cpu_leg("dna", cores, 4.0, 50)
    CPU_LEGS["codon"] = cpu_leg("codon", cores, 4.0, 3)


def cpu_leg_main(args):
    from gems_amd import workloads as W
    p = {"protein": W.config2, "dna": W.config1, "codon": W.config4}[args.cpu_leg]()
    ref, med, n = time_oracle(p, args.threads, args.seconds, max_reps=args.max_reps,
                              min_reps=1 if args.threads == 1 else 2, safe=(p.n == 61))
    print(json.dumps({"ref": ref, "median_s": med, "reps": n}), flush=True)


def cpu_baseline(p, seconds=10.0):
    """The oracle (FP64 restatement of the reference kernel, pattern-interleaved SIMD lanes,
    OpenMP packets) on every host core, threads bound close, and on one core."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import support as S

    cores = host_cores()
    ref, med, n = CPU_LEGS["protein"]
    _, med1, n1 = CPU_LEGS["protein_1core"]
    upd = (p.ntaxa - 2) * p.nptn
    return ref, {
        "value": upd / med,
        "unit": "updates/s",
        "cores": cores,
        "kind": "port",
        "sample": f"{n} full traversals + branch logL of the bench problem (median), "
                  f"{S.oracle_lib().orc_lanes()}-lane SIMD oracle, OpenMP packets, OMP_PROC_BIND=close, "
                  f"all {cores} cores of this host, in a child process that does not use the GPU",
        "value_1core": upd / med1,
        "sample_1core": f"{n1} traversals on one core (median)",
        "cpu_model": cpu_model(),
    }


ALG_FLOPS = {4: 272, 20: 6480, 61: 59780}  # ncat(4n^2 + n) per update, ncat = 4
PEAK = {"fp64_tflops": 78.6, "fp32_tflops": 157.3}


def timed_traversals(G, p, precision, steps, scaling):
    """Wall ms per full traversal + logL step and the traversal kernel's mean duration (HIP
    events on the engine stream)."""
    g = G.Tree(p, precision=precision, scaling=scaling)
    g.full_traversal()
    g.set_timing(True)
    G.lib().plk_synchronize(g.context())
    t0 = time.perf_counter()
    for _ in range(steps):
        v = g.full_traversal()
    G.lib().plk_synchronize(g.context())
    el = (time.perf_counter() - t0) / steps
    k_ms, k_n = g.timing(0)
    g.close()
    return el * 1e3, k_ms / max(k_n, 1), v


def other_configs(G, no_cpu):
    """Compact figures of the DNA (32 taxa x 10k sites), batched-partition (300 x 64 taxa)
    and codon (48 taxa x 5k codons) workloads, both forms, each with its roofline fraction;
    CPU oracle beside the DNA and codon ones on every host core."""
    from gems_amd import workloads as W

    res = {}
    # DNA: HBM-bound unit (256 B FP64, 128 B FP32 per update)
    p = W.config1()
    upd = (p.ntaxa - 2) * p.nptn
    r = {"workload": "dna4_G4_32taxa_10000sites_full_traversal", "patterns": p.nptn}
    for prec, name, w in ((G.FP64, "fp64", 8), (G.FP32, "fp32", 4)):
        ms, kms, v = timed_traversals(G, p, prec, 300, G.NORM)
        bpu = 2 * 4 * 4 * w
        r[name] = {"updates_per_s": upd / (ms * 1e-3), "ms_per_step": ms, "kernel_ms": kms, "logl": v,
                   "roofline": {"bound": "hbm", "unit": "GB/s", "achieved": upd * bpu / (kms * 1e-3) / 1e9,
                                "peak": HBM_PEAK_GBS, "frac": upd * bpu / (kms * 1e-3) / 1e9 / HBM_PEAK_GBS}}
    r["fp32_over_fp64"] = r["fp32"]["updates_per_s"] / r["fp64"]["updates_per_s"]
    if not no_cpu:
        ref, med, n = CPU_LEGS["dna"]
        r["cpu_updates_per_s"] = upd / med
        r["cpu_cores"] = host_cores()
        r["fp32_over_cpu"] = r["fp32"]["updates_per_s"] / r["cpu_updates_per_s"]
        r["logl_rel_err"] = {k: abs(r[k]["logl"] - ref) / abs(ref) for k in ("fp64", "fp32")}
    log(f"dna 32x10k: fp64 {r['fp64']['ms_per_step'] * 1e3:.1f} us, fp32 {r['fp32']['ms_per_step'] * 1e3:.1f} us")
    res["dna_32taxa_10k"] = r

    # batched partitions: bytes per update by state count, summed over the partitions
    probs = W.config3()
    upd = sum((q.ntaxa - 2) * q.nptn for q in probs)
    br = [(0, int(q.nei[0, 0])) for q in probs]
    r = {"workload": "300_unlinked_partitions_64taxa_70pct_dna", "patterns": int(sum(q.nptn for q in probs)),
         "updates_per_traversal": upd}
    for prec, name, w in ((G.FP64, "fp64", 8), (G.FP32, "fp32", 4)):
        trees = [G.Tree(q, precision=prec) for q in probs]
        b = G.Batch(trees)
        b.autotune(br, reps=2)
        b.loglik(br, clear=True)
        kms0, kn0, _ = b.timing()  # cumulative event time of the batch's launches
        steps = 20
        t0 = time.perf_counter()
        for _ in range(steps):
            vals, tot = b.loglik(br, clear=True)
        el = (time.perf_counter() - t0) / steps
        kms, kn, _ = b.timing()
        kms = (kms - kms0) / max(kn - kn0, 1)
        nbytes = sum((q.ntaxa - 2) * q.nptn * 2 * len(q.rates) * q.n * w for q in probs)
        r[name] = {"updates_per_s": upd / el, "ms_per_step": el * 1e3, "kernel_ms": kms, "total_logl": tot,
                   "roofline": {"bound": "hbm", "unit": "GB/s", "achieved": nbytes / (kms * 1e-3) / 1e9,
                                "peak": HBM_PEAK_GBS, "frac": nbytes / (kms * 1e-3) / 1e9 / HBM_PEAK_GBS}}
        b.close()
        for t in trees:
            t.close()
    log(f"batched 300 partitions: fp64 {r['fp64']['ms_per_step']:.2f} ms, fp32 {r['fp32']['ms_per_step']:.2f} ms")
    res["batched_300_partitions"] = r

    # codon: matrix-core (flop) bound, 59,780 algorithmic flops per update
    p = W.config4()
    upd = (p.ntaxa - 2) * p.nptn
    r = {"workload": "codon61_G4_48taxa_5000codons_full_traversal", "patterns": p.nptn}
    for prec, name, sc, pk in ((G.FP64, "fp64", G.SAFE, PEAK["fp64_tflops"]), (G.FP32, "fp32", G.NORM, PEAK["fp32_tflops"])):
        ms, kms, v = timed_traversals(G, p, prec, 10, sc)
        tf = upd * ALG_FLOPS[61] / (kms * 1e-3) / 1e12
        r[name] = {"updates_per_s": upd / (ms * 1e-3), "ms_per_step": ms, "kernel_ms": kms, "logl": v,
                   "roofline": {"bound": "mfma", "unit": "TFLOP/s", "achieved": tf, "peak": pk, "frac": tf / pk}}
    r["fp32_over_fp64"] = r["fp32"]["updates_per_s"] / r["fp64"]["updates_per_s"]
    if not no_cpu:
        ref, med, n = CPU_LEGS["codon"]
        r["cpu_updates_per_s"] = upd / med
        r["cpu_cores"] = host_cores()
        r["logl_rel_err"] = {k: abs(r[k]["logl"] - ref) / abs(ref) for k in ("fp64", "fp32")}
    log(f"codon 48x5k: fp64 {r['fp64']['ms_per_step']:.2f} ms, fp32 {r['fp32']['ms_per_step']:.2f} ms")
    res["codon_48taxa_5k"] = r

    # serving: 16 requests (16-512 taxa, 1k-100k sites, DNA or protein, logL or Newton on
    # one branch) submitted at once (the saturated end of the arrival-rate sweep)
    specs, _ = W.config5_specs(16)
    probs = [W.config5_problem(sp) for sp in specs]
    upd = sum((q.ntaxa - 2) * q.nptn for q in probs)
    r = {"workload": "serving_16_requests_16-512taxa_1k-100k_sites_saturation", "requests": len(probs),
         "first_traversal_updates": upd}
    for prec, name in ((G.FP64, "fp64"), (G.FP32, "fp32")):
        best = None
        for rep in range(2):  # the first run also fills the tree pool
            reqs = [G.Request(q, kind=G.REQ_LOGL if sp[3] == 0 else G.REQ_OPT_BRANCH, precision=prec,
                              opt_branch=(q.ntaxa, int(q.nei[q.ntaxa, 0]))) for q, sp in zip(probs, specs)]
            srv = G.Server(nworkers=8)
            for q in reqs:
                srv.submit(q)
            srv.drain()
            st, cn = srv.stats(), srv.counters()
            srv.close()
            best = (st, cn) if best is None or st["makespan_s"] < best[0]["makespan_s"] else best
        st, cn = best
        r[name] = {"p50_s": st["p50_s"], "p99_s": st["p99_s"], "makespan_s": st["makespan_s"],
                   "throughput_req_s": len(probs) / st["makespan_s"], "failed": st["failed"],
                   "shared_launches": cn["launches"], "request_traversals_in_them": cn["batched"],
                   "peak_concurrent": st["peak_concurrent"]}
    log(f"serving 16 requests: fp64 p50 {r['fp64']['p50_s'] * 1e3:.1f} ms p99 {r['fp64']['p99_s'] * 1e3:.1f} ms, "
        f"fp32 p50 {r['fp32']['p50_s'] * 1e3:.1f} ms p99 {r['fp32']['p99_s'] * 1e3:.1f} ms")
    res["serving"] = r
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--taxa", type=int, default=100)
    ap.add_argument("--sites", type=int, default=50000)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-configs", action="store_true", help="skip the compact figures of the other workloads")
    ap.add_argument("--cpu-leg", default=None, help=argparse.SUPPRESS)  # child process: time the oracle only
    ap.add_argument("--threads", type=int, default=1, help=argparse.SUPPRESS)
    ap.add_argument("--seconds", type=float, default=4.0, help=argparse.SUPPRESS)
    ap.add_argument("--max-reps", type=int, default=50, help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.cpu_leg:
        cpu_leg_main(args)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); "
              "the pattern-sharded multi-GPU path is not part of this build", file=sys.stderr)
        sys.exit(2)

    if not args.no_cpu:
        run_cpu_legs()
        log("cpu legs: " + ", ".join(f"{k} {v[1] * 1e3:.2f} ms" for k, v in CPU_LEGS.items()))

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

    # df/ddf error of both forms against the oracle on two branches of this problem
    deriv = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import support as S
        o = S.Oracle(p, nthreads=host_cores())
        deriv = {}
        brs = [(0, int(p.nei[0, 0])), (p.ntaxa, int(p.nei[p.ntaxa, 0]))]
        dref = [o.derivs(u, v)[:2] for (u, v) in brs]
        o.close()
        for prec, name in ((G.FP64, "fp64"), (G.FP32, "fp32")):
            g = G.Tree(p, precision=prec)
            e = 0.0
            for (u, v), dd in zip(brs, dref):
                dg = g.derivs(u, v)[:2]
                e = max(e, max(abs(dg[i] - dd[i]) / max(1.0, abs(dd[i])) for i in range(2)))
            g.close()
            deriv[name] = e
        log(f"derivatives: max rel err fp64 {deriv['fp64']:.2e}, fp32 {deriv['fp32']:.2e}")

    others = None if args.no_configs else other_configs(G, args.no_cpu)

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
            "latency": "latency_p50_s/latency_p99_s: serving, 16 requests (16-512 taxa, 1k-100k sites) at saturation, "
                       "FP32 form (_fp64: FP64 form); other_workloads.serving has both",
            "deriv_max_rel_err_vs_oracle": deriv,
            "other_workloads": others,
        },
        "latency_p50_s": others["serving"]["fp32"]["p50_s"] if others else None,
        "latency_p99_s": others["serving"]["fp32"]["p99_s"] if others else None,
        "latency_p50_s_fp64": others["serving"]["fp64"]["p50_s"] if others else None,
        "latency_p99_s_fp64": others["serving"]["fp64"]["p99_s"] if others else None,
        "value_fp64": f64["ups"],
        "ms_per_step_fp64": f64["ms_per_step"],
        "logl_rel_err_fp64": err64,
        "logl_rel_err_fp32": err32,
        "roofline_fp64": {
            "bound": "hbm", "achieved": f64["achieved"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": f64["achieved"] / HBM_PEAK_GBS, "kernel": "k_traverse_f64<20,MODE_LOGL>",
            "kernel_avg_ms": f64["kernel_avg_ms"], "alg_bytes_per_update": f64["bytes_per_update"],
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
