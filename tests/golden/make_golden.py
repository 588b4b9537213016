"""Writes tests/golden/model_fixtures.json: discrete-Gamma rates and transition matrices
computed with scipy (exact special functions / expm), used to pin the engine's host
model code.  Run: python tests/golden/make_golden.py"""
import json
import pathlib

import numpy as np
import scipy.linalg
import scipy.special
import scipy.stats


def gamma_rates(alpha, k):
    cuts = scipy.stats.gamma.ppf(np.arange(1, k) / k, a=alpha, scale=1.0 / alpha)
    upper = scipy.special.gammainc(alpha + 1, cuts * alpha)
    edges = np.concatenate([[0.0], upper, [1.0]])
    r = np.diff(edges) * k
    return (r * k / r.sum()).tolist()


def main():
    out = {"gamma": []}
    for alpha in (0.05, 0.1, 0.3, 0.5, 0.8, 1.0, 2.0, 5.0, 20.0, 100.0):
        for k in (2, 4, 6, 8):
            out["gamma"].append({"alpha": alpha, "ncat": k, "rates": gamma_rates(alpha, k)})
    rng = np.random.default_rng(20261202)
    mats = []
    for n in (4, 20):
        pi = rng.dirichlet(np.ones(n))
        exch = rng.uniform(0.5, 2.0, n * (n - 1) // 2)
        R = np.zeros((n, n))
        R[np.triu_indices(n, 1)] = exch
        R = R + R.T
        Q = R * pi[None, :]
        np.fill_diagonal(Q, 0.0)
        rs = Q.sum(1)
        Q -= np.diag(rs)
        Q /= rs @ pi
        for t in (0.001, 0.1, 1.0, 5.0):
            mats.append({"n": n, "pi": pi.tolist(), "exch": exch.tolist(), "t": t,
                         "Q": Q.ravel().tolist(), "P": scipy.linalg.expm(Q * t).ravel().tolist()})
    out["pmatrix"] = mats
    path = pathlib.Path(__file__).with_name("model_fixtures.json")
    path.write_text(json.dumps(out))
    print("wrote", path)


if __name__ == "__main__":
    main()
