"""The five measured workloads as seeded synthetic problems (shared by the benchmarks and
the full-size parity tests).

  1  DNA GTR+G4, 32 taxa, 10,000 sites, Exp(mean 0.1) lengths, alpha 0.5
  2  20-state protein +G4, 100 taxa, 50,000 sites, Exp(mean 0.2) lengths, alpha 0.8
  3  300 unlinked partitions on one 64-taxon topology (70% DNA, 30% protein), lengths
     lognormal(median 500, sigma 1.0) clipped to [100, 5000] sites, alpha U[0.3, 1.5]
  4  61-state codon +G4, 48 taxa, 5,000 codons, alpha 0.5, Exp(mean 0.1) lengths
  5  serving requests: 16-512 taxa and 1k-100k sites (log-uniform), DNA or protein
"""
import numpy as np

from . import SEQ_CODON, SEQ_DNA, SEQ_PROTEIN, Problem


def config1():
    return Problem(SEQ_DNA, 32, 10000, mean_len=0.1, alpha=0.5, seed=1)


def config2():
    return Problem(SEQ_PROTEIN, 100, 50000, mean_len=0.2, alpha=0.8, seed=1)


def config3(nparts=300):
    rng = np.random.default_rng(3)
    base = Problem(SEQ_DNA, 64, 50, seed=3000)
    probs = []
    for i in range(nparts):
        seq = SEQ_DNA if rng.random() < 0.7 else SEQ_PROTEIN
        # lognormal, median 500 sites, sigma 1.0 (sigma is not fixed by the workload description)
        nsites = int(np.clip(np.round(rng.lognormal(np.log(500), 1.0)), 100, 5000))
        probs.append(Problem(seq, 64, nsites, mean_len=0.1 if seq == SEQ_DNA else 0.2,
                             alpha=float(rng.uniform(0.3, 1.5)), seed=30000 + i, topology=base.nei))
    return probs


def config4():
    return Problem(SEQ_CODON, 48, 5000, mean_len=0.1, alpha=0.5, seed=4)


def config5_specs(nreq, seed=5):
    """(seq_type, taxa, sites, request kind 0 logL / 1 Newton on one branch, seed) per request."""
    rng = np.random.default_rng(seed)
    specs = []
    for i in range(nreq):
        seq = SEQ_DNA if rng.random() < 0.5 else SEQ_PROTEIN
        T = int(np.exp(rng.uniform(np.log(16), np.log(512))))
        sites = int(np.exp(rng.uniform(np.log(1000), np.log(100000))))
        kind = 0 if rng.random() < 0.5 else 1
        specs.append((seq, T, sites, kind, 500 + i))
    return specs, rng


def config5_problem(spec):
    seq, T, sites, _kind, seed = spec
    return Problem(seq, T, sites, mean_len=0.1 if seq == SEQ_DNA else 0.2, alpha=0.5 if seq == SEQ_DNA else 0.8,
                   seed=seed)
