"""Test support: ctypes binding of the CPU oracle, independent numpy references
(brute-force enumeration, real-space pruning with scipy expm), model builders and
random trees.  Nothing here is used by the product path."""
import ctypes
import os
import pathlib

import numpy as np
import scipy.linalg
import scipy.special
import scipy.stats

ROOT = pathlib.Path(__file__).resolve().parent.parent
ORACLE_DIR = ROOT / "oracle"

SEQ_DNA, SEQ_PROTEIN, SEQ_CODON = 0, 1, 2
LOG_SCALING_THRESHOLD = -177.4456782233459932741


def _cpu_has_avx512():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("flags"):
                    return "avx512f" in line.split()
    except OSError:
        pass
    return False


_oracle = None
ORC_FN1 = ctypes.CFUNCTYPE(ctypes.c_double, ctypes.c_void_p, ctypes.c_double)


def oracle_lib(lanes=None):
    """Load the oracle shared library (built by oracle/Makefile)."""
    global _oracle
    if lanes is None and _oracle is not None:
        return _oracle
    if lanes is None:
        lanes = 8 if _cpu_has_avx512() else 4
    path = ORACLE_DIR / f"libplk_oracle_v{lanes}.so"
    if not path.exists():
        raise RuntimeError(f"oracle not built: {path} (run make -C oracle)")
    lib = _bind(ctypes.CDLL(str(path)))
    if _oracle is None or lanes is None:
        _oracle = lib
    return lib


def load_oracle(path):
    """An oracle library built elsewhere (the planted-mistake build of a test)."""
    return _bind(ctypes.CDLL(str(path)))


def _bind(lib):
    P = ctypes.c_void_p
    dp = ctypes.POINTER(ctypes.c_double)
    ip = ctypes.POINTER(ctypes.c_int)
    u8p = ctypes.POINTER(ctypes.c_uint8)
    u16p = ctypes.POINTER(ctypes.c_uint16)
    sig = {
        "orc_lanes": (ctypes.c_int, []),
        "orc_updates": (ctypes.c_longlong, [P]),
        "orc_create": (P, [ctypes.c_int] * 6),
        "orc_destroy": (None, [P]),
        "orc_set_model": (None, [P, dp, dp, dp, dp, dp]),
        "orc_set_tip_lh": (None, [P, ctypes.c_int, dp]),
        "orc_set_patterns": (None, [P, u8p, dp, dp]),
        "orc_set_tree": (None, [P, ip, dp]),
        "orc_set_length": (None, [P, ctypes.c_int, ctypes.c_int, ctypes.c_double]),
        "orc_get_length": (ctypes.c_double, [P, ctypes.c_int, ctypes.c_int]),
        "orc_clear_all": (None, [P]),
        "orc_branch_loglik": (ctypes.c_double, [P, ctypes.c_int, ctypes.c_int, dp]),
        "orc_branch_derivs": (ctypes.c_int, [P, ctypes.c_int, ctypes.c_int, ctypes.c_int, dp, dp]),
        "orc_loglik_from_theta": (ctypes.c_double, [P, ctypes.c_int, ctypes.c_int]),
        "orc_get_partial": (ctypes.c_int, [P, ctypes.c_int, ctypes.c_int, dp, u16p]),
        "orc_optimize_one_branch": (ctypes.c_double, [P, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                                      ctypes.c_double, ctypes.c_int]),
        "orc_optimize_all_branches": (ctypes.c_double, [P, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                                        ctypes.c_double, ctypes.c_int]),
        "orc_tip_lh": (None, [ctypes.c_int, ctypes.c_int, dp, dp]),
        "orc_state_unknown": (ctypes.c_int, [ctypes.c_int, ctypes.c_int]),
        "orc_nni_eval": (ctypes.c_double, [P, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                           ctypes.c_double, ctypes.c_int, dp]),
        "orc_rell_save_tree": (None, [ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_float), dp, ctypes.c_double,
                                      ctypes.c_double, dp, dp, dp, ctypes.POINTER(ctypes.c_int32), dp,
                                      ctypes.POINTER(ctypes.c_int32)]),
        "orc_gamma_rates": (None, [ctypes.c_double, ctypes.c_int, dp]),
        "orc_gamma_rates_median": (None, [ctypes.c_double, ctypes.c_int, dp]),
        "orc_set_gamma_median": (None, [P, ctypes.c_int]),
        "orc_minimize_one_dimen": (ctypes.c_double, [ORC_FN1, P, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                                     ctypes.c_double, dp, dp]),
        "orc_set_gamma_shape": (None, [P, ctypes.c_double]),
        "orc_set_eigen": (None, [P, dp, dp, dp]),
        "orc_get_gamma_shape": (ctypes.c_double, [P]),
        "orc_optimize_gamma": (ctypes.c_double, [P, ctypes.c_double, ctypes.c_double, ctypes.c_double]),
        "orc_optimize_parameters": (ctypes.c_double, [P, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                                      ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                                      ctypes.c_double, ctypes.c_int]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class Oracle:
    """Thin object wrapper over one orc_tree."""

    def __init__(self, prob, safe=False, nthreads=2, lanes=None, lib=None):
        self.lib = lib if lib is not None else oracle_lib(lanes)
        self.p = prob
        n, ncat = prob.n, len(prob.rates)
        self.h = self.lib.orc_create(n, ncat, prob.ntaxa, prob.nptn, int(safe), nthreads)
        self._keep = []
        m = prob
        self.lib.orc_set_model(self.h, _dp(m.eval), _dp(m.evec), _dp(m.inv_evec), _dp(m.rates), _dp(m.props))
        su = self.lib.orc_state_unknown(n, m.seq_type)
        tip = np.zeros((su + 1) * n)
        self.lib.orc_tip_lh(n, m.seq_type, _dp(m.inv_evec), _dp(tip))
        self.tip_lh = tip.reshape(su + 1, n)
        self.lib.orc_set_tip_lh(self.h, su, _dp(tip))
        states = np.ascontiguousarray(m.states, dtype=np.uint8)
        freq = np.ascontiguousarray(m.freq, dtype=np.float64)
        invar = np.ascontiguousarray(m.invar if m.invar is not None else np.zeros(m.nptn))
        self.lib.orc_set_patterns(self.h, states.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)),
                                  _dp(freq), _dp(invar))
        nei = np.ascontiguousarray(m.nei, dtype=np.int32)
        blen = np.ascontiguousarray(m.blen, dtype=np.float64)
        self.lib.orc_set_tree(self.h, nei.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), _dp(blen))

    def close(self):
        if self.h:
            self.lib.orc_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def loglik(self, a, b, with_patterns=False):
        pl = np.zeros(self.p.nptn)
        v = self.lib.orc_branch_loglik(self.h, a, b, _dp(pl))
        return (v, pl) if with_patterns else v

    def derivs(self, a, b, reuse=False):
        df = np.zeros(1)
        ddf = np.zeros(1)
        st = self.lib.orc_branch_derivs(self.h, a, b, int(reuse), _dp(df), _dp(ddf))
        return float(df[0]), float(ddf[0]), st

    def loglik_from_theta(self, a, b):
        return self.lib.orc_loglik_from_theta(self.h, a, b)

    def set_length(self, a, b, t):
        self.lib.orc_set_length(self.h, a, b, t)

    def get_length(self, a, b):
        return self.lib.orc_get_length(self.h, a, b)

    def partial(self, dad, node, safe=False):
        n, ncat = self.p.n, len(self.p.rates)
        out = np.zeros(self.p.nptn * n * ncat)
        sc = np.zeros(self.p.nptn * (ncat if safe else 1), dtype=np.uint16)
        st = self.lib.orc_get_partial(self.h, dad, node, _dp(out), sc.ctypes.data_as(ctypes.POINTER(ctypes.c_uint16)))
        assert st == 0
        return out.reshape(self.p.nptn, ncat, n), sc

    def optimize_all(self, iterations=100, tol=1e-3, min_bl=1e-6, max_bl=10.0, max_nr=100):
        return self.lib.orc_optimize_all_branches(self.h, iterations, tol, min_bl, max_bl, max_nr)

    def nni_eval(self, n1, n2, move, min_bl=1e-6, max_bl=10.0, max_nr=10):
        lens = np.zeros(5)
        sc = self.lib.orc_nni_eval(self.h, n1, n2, move, min_bl, max_bl, max_nr, _dp(lens))
        return sc, lens

    def optimize_one(self, a, b, min_bl=1e-6, max_bl=10.0, max_nr=100):
        return self.lib.orc_optimize_one_branch(self.h, a, b, min_bl, max_bl, max_nr)

    def set_gamma_shape(self, alpha):
        self.lib.orc_set_gamma_shape(self.h, alpha)

    def gamma_shape(self):
        return self.lib.orc_get_gamma_shape(self.h)

    def set_gamma_median(self, median):
        self.lib.orc_set_gamma_median(self.h, int(median))

    # ---- reversible model from exchangeabilities (ModelMarkov), for the model search
    def set_gtr(self, exch, pi):
        self.gtr_exch = np.array(exch, dtype=np.float64)
        self.gtr_pi = np.array(pi, dtype=np.float64)
        self._apply_gtr()

    def _apply_gtr(self):
        n = self.p.n
        Q = rate_matrix(self.gtr_exch, self.gtr_pi)
        w, U, Ui = eigen_system(Q, self.gtr_pi)
        w, U, Ui = (np.ascontiguousarray(a, dtype=np.float64) for a in (w, U, Ui))
        self.lib.orc_set_eigen(self.h, _dp(w), _dp(U), _dp(Ui))
        su = self.lib.orc_state_unknown(n, self.p.seq_type)
        tip = np.zeros((su + 1) * n)
        self.lib.orc_tip_lh(n, self.p.seq_type, _dp(Ui), _dp(tip))
        self.lib.orc_set_tip_lh(self.h, su, _dp(tip))

    def gtr_objective(self, x):
        """ModelMarkov::targetFunk: the free exchangeabilities from x, -logL."""
        nd = len(x)
        if any(self.gtr_exch[i] != x[i] for i in range(nd)):
            self.gtr_exch[:nd] = x
            self._apply_gtr()
        return -self.compute_likelihood()

    def optimize_model(self, gradient_epsilon=1e-4):
        """ModelMarkov::optimizeParameters with fixed frequencies (BFGS of optim_oracle)."""
        nd = len(self.gtr_exch) - 1
        x = [float(v) for v in self.gtr_exch[:nd]]
        optim_oracle().dfpmin(self.gtr_objective, x, [1e-4] * nd, [100.0] * nd, max(gradient_epsilon, 1e-4))
        f = self.gtr_objective(x)
        return self.gtr_exch.copy(), -f

    def optimize_parameters_model(self, num_iterations=100, logl_epsilon=0.01, gradient_epsilon=1e-4,
                                  min_bl=1e-6, max_bl=10.0, max_nr=100, gamma=True):
        """ModelFactory::optimizeParameters with the exchangeabilities free (and the Gamma
        shape when gamma), in Python over the oracle's pieces."""
        cur = self.compute_likelihood()
        for i in range(2, num_iterations):
            self.optimize_all(iterations=min(i, 3), tol=logl_epsilon, min_bl=min_bl, max_bl=max_bl, max_nr=max_nr)
            _e, new = self.optimize_model(gradient_epsilon)
            if gamma:
                _a, new = self.optimize_gamma(gradient_epsilon)
            if new > cur + logl_epsilon:
                cur = new
            else:
                cur = self.optimize_all(iterations=100, tol=logl_epsilon, min_bl=min_bl, max_bl=max_bl,
                                        max_nr=max_nr)
                break
        return self.gtr_exch.copy(), cur

    def compute_likelihood(self):
        return self.loglik(0, int(self.p.nei[0][0]))

    def optimize_gamma(self, gradient_epsilon=1e-4, min_gamma=0.02, max_gamma=1000.0):
        v = self.lib.orc_optimize_gamma(self.h, gradient_epsilon, min_gamma, max_gamma)
        return self.gamma_shape(), v

    def optimize_parameters(self, num_iterations=100, logl_epsilon=0.01, gradient_epsilon=1e-4, min_gamma=0.02,
                            max_gamma=1000.0, min_bl=1e-6, max_bl=10.0, max_nr=100):
        v = self.lib.orc_optimize_parameters(self.h, num_iterations, logl_epsilon, gradient_epsilon, min_gamma,
                                             max_gamma, min_bl, max_bl, max_nr)
        return self.gamma_shape(), v


_optim = None


def optim_oracle():
    """oracle/optim_oracle.py (the BFGS restatement)."""
    global _optim
    if _optim is None:
        import importlib.util
        spec = importlib.util.spec_from_file_location("optim_oracle", ORACLE_DIR / "optim_oracle.py")
        _optim = importlib.util.module_from_spec(spec)
        spec.loader.exec_module(_optim)
    return _optim


def minimize_one_dimen(fn, xmin, xguess, xmax, tol):
    """The oracle's minimizeOneDimen on a Python function; returns (x, f(x), f2, calls)."""
    calls = []

    def cb(_ctx, x):
        calls.append(x)
        return float(fn(x))

    c = ORC_FN1(cb)
    fx, f2 = ctypes.c_double(), ctypes.c_double()
    x = oracle_lib().orc_minimize_one_dimen(c, None, xmin, xguess, xmax, tol, ctypes.byref(fx), ctypes.byref(f2))
    return x, fx.value, f2.value, calls


# ---------------------------------------------------------------------------------------
# numpy model builders (independent of the product's C++ host code)


def rate_matrix(exch, pi):
    """Q = R diag(pi), zero row sums, normalised to one substitution per unit time
    (model/modelmarkov.cpp:1472-1514).  exch is the upper-triangle row-major list."""
    n = len(pi)
    R = np.zeros((n, n))
    iu = np.triu_indices(n, 1)
    R[iu] = exch
    R = R + R.T
    Q = R * pi[None, :]
    np.fill_diagonal(Q, 0.0)
    rs = Q.sum(axis=1)
    Q -= np.diag(rs)
    Q /= np.dot(rs, pi)
    return Q


def eigen_system(Q, pi):
    """Symmetrised eigendecomposition (model/modelmarkov.cpp:1519-1562)."""
    s = np.sqrt(pi)
    S = (s[:, None] * Q) / s[None, :]
    S = 0.5 * (S + S.T)
    w, Vm = np.linalg.eigh(S)
    evec = Vm / s[:, None]
    inv_evec = Vm.T * s[None, :]
    return w, np.ascontiguousarray(evec), np.ascontiguousarray(inv_evec)


def gamma_rates_exact(alpha, k):
    """Yang (1994) mean-of-category discrete gamma rates, exact via scipy."""
    cuts = scipy.stats.gamma.ppf(np.arange(1, k) / k, a=alpha, scale=1.0 / alpha)
    upper = scipy.special.gammainc(alpha + 1, cuts * alpha)
    edges = np.concatenate([[0.0], upper, [1.0]])
    rates = np.diff(edges) * k
    return rates * k / rates.sum()


def codon_exch(kappa, omega):
    """GY-style codon exchangeabilities over the 61 sense codons of the standard code,
    codons in AAA..TTT order (model/modelcodon.cpp:366-387, 574-651, 877-906)."""
    code = "KNKNTTTTRSRSIIMIQHQHPPPPRRRRLLLLEDEDAAAAGGGGVVVV*Y*YSSSS*CWCLFLF"
    sense = [c for c in range(64) if code[c] != "*"]
    n = len(sense)
    R = np.zeros((n, n))
    for a in range(n):
        for b in range(a + 1, n):
            ci, cj = sense[a], sense[b]
            d = [(ci >> s) & 3 != (cj >> s) & 3 for s in (4, 2, 0)]
            if sum(d) != 1:
                continue
            pos = [4, 2, 0][d.index(True)]
            n1, n2 = (ci >> pos) & 3, (cj >> pos) & 3
            ts = abs(n1 - n2) == 2
            syn = code[ci] == code[cj]
            r = 1.0
            if syn:
                if ts:
                    r *= kappa
            else:
                r *= omega * kappa if ts else omega
            R[a, b] = r
    return R[np.triu_indices(n, 1)], sense


class Problem:
    """A complete likelihood problem: model, tree, patterns."""

    def __init__(self, **kw):
        self.__dict__.update(kw)


def random_tree(ntaxa, rng, mean_len=0.1):
    """Random unrooted binary tree by repeatedly splitting a leaf edge; nodes 0..T-1 leaves."""
    nn = 2 * ntaxa - 2
    nei = -np.ones((nn, 3), dtype=np.int32)
    blen = np.zeros((nn, 3))
    edges = []

    def link(a, b, t):
        ka = list(nei[a]).index(-1)
        kb = list(nei[b]).index(-1)
        nei[a, ka] = b
        nei[b, kb] = a
        blen[a, ka] = t
        blen[b, kb] = t

    c = ntaxa
    nxt_int = ntaxa + 1
    for leaf in range(3):
        link(c, leaf, rng.exponential(mean_len))
        edges.append((c, leaf))
    for leaf in range(3, ntaxa):
        a, b = edges[rng.integers(len(edges))]
        m = nxt_int
        nxt_int += 1
        # split edge (a,b) with new internal node m, attach leaf
        ka = list(nei[a]).index(b)
        kb = list(nei[b]).index(a)
        nei[a, ka] = -1
        nei[b, kb] = -1
        edges.remove((a, b))
        link(a, m, rng.exponential(mean_len))
        link(m, b, rng.exponential(mean_len))
        link(m, leaf, rng.exponential(mean_len))
        edges += [(a, m), (m, b), (m, leaf)]
    # leaves keep their single neighbour in slot 0
    for leaf in range(ntaxa):
        k = [i for i in range(3) if nei[leaf, i] >= 0][0]
        if k != 0:
            nei[leaf, 0], nei[leaf, k] = nei[leaf, k], -1
            blen[leaf, 0], blen[leaf, k] = blen[leaf, k], 0.0
    return nei, np.maximum(blen, 1e-6) * (nei >= 0)


def caterpillar_tree(ntaxa, length):
    nn = 2 * ntaxa - 2
    nei = -np.ones((nn, 3), dtype=np.int32)
    blen = np.zeros((nn, 3))

    def link(a, b):
        ka = list(nei[a]).index(-1)
        kb = list(nei[b]).index(-1)
        nei[a, ka], nei[b, kb] = b, a
        blen[a, ka] = blen[b, kb] = length

    internals = list(range(ntaxa, nn))
    link(internals[0], 0)
    link(internals[0], 1)
    for i in range(len(internals) - 1):
        link(internals[i], internals[i + 1])
        link(internals[i + 1], i + 2)
    link(internals[-1], ntaxa - 1)
    return nei, blen


def make_problem(n, ntaxa, nsites, rng, seq_type=None, alpha=0.5, ncat=4, mean_len=0.1,
                 ambiguous=0.0, tree=None, compress=True):
    if seq_type is None:
        seq_type = {4: SEQ_DNA, 20: SEQ_PROTEIN}.get(n, SEQ_CODON)
    pi = rng.dirichlet(np.ones(n))
    if n == 4:
        exch = rng.uniform(0.5, 2.0, 6)
    elif n == 20:
        exch = rng.gamma(1.0, 1.0, 190)
    else:
        exch, _ = codon_exch(rng.uniform(1, 4), rng.uniform(0.1, 1))
    Q = rate_matrix(exch, pi)
    w, U, Ui = eigen_system(Q, pi)
    rates = gamma_rates_exact(alpha, ncat) if ncat > 1 else np.ones(1)
    props = np.full(ncat, 1.0 / ncat)
    if tree is None:
        nei, blen = random_tree(ntaxa, rng, mean_len)
    else:
        nei, blen = tree
    # simulate i.i.d. random states (not down the tree: tests want arbitrary patterns)
    states = rng.integers(0, n, size=(ntaxa, nsites)).astype(np.uint8)
    su = {SEQ_DNA: 18, SEQ_PROTEIN: 23}.get(seq_type, n)
    if ambiguous > 0:
        mask = rng.random(states.shape) < ambiguous
        if seq_type == SEQ_DNA:
            amb = rng.integers(n, su + 1, size=states.shape)
        elif seq_type == SEQ_PROTEIN:
            amb = rng.integers(n, su + 1, size=states.shape)
        else:
            amb = np.full(states.shape, su)
        states = np.where(mask, amb, states).astype(np.uint8)
    if compress:
        cols, inv, cnt = np.unique(states.T, axis=0, return_inverse=True, return_counts=True)
        states = np.ascontiguousarray(cols.T)
        freq = cnt.astype(np.float64)
    else:
        freq = np.ones(nsites)
    prob = Problem(n=n, ntaxa=ntaxa, nptn=states.shape[1], seq_type=seq_type, pi=pi, Q=Q,
                   eval=np.ascontiguousarray(w), evec=U, inv_evec=Ui, rates=np.ascontiguousarray(rates),
                   props=props, nei=nei, blen=blen, states=states, freq=freq, invar=None, su=su)
    prob.tip_lh = np.stack([Ui @ tip_likelihood(prob, s) for s in range(su + 1)])
    return prob


def tip_likelihood(prob, s):
    """Real-space tip vector of a state (independent of the oracle)."""
    n = prob.n
    v = np.zeros(n)
    if s < n:
        v[s] = 1.0
    elif prob.seq_type == SEQ_DNA and s < 18:
        for j in range(n):
            if (s - n + 1) & (1 << j):
                v[j] = 1.0
    elif prob.seq_type == SEQ_PROTEIN and s < n + 3:
        amb = [(2, 3), (5, 6), (9, 10)][s - n]
        v[list(amb)] = 1.0
    else:
        v[:] = 1.0
    return v


def pruning_loglik(prob, a, b, dtype=np.float64):
    """Independent real-space Felsenstein pruning with scipy expm and per-node
    renormalisation (log accumulation); returns per-pattern logL and total."""
    nei, blen = prob.nei, prob.blen
    ncat = len(prob.rates)
    P = {}

    def pmat(t, c):
        key = (round(t, 17), c)
        if key not in P:
            P[key] = scipy.linalg.expm(prob.Q * prob.rates[c] * t)
        return P[key]

    def length(u, v):
        k = list(nei[u]).index(v)
        return blen[u, k]

    def down(node, dad):
        # returns (L[c][ptn][x], logscale[ptn])
        if node < prob.ntaxa:
            T = np.stack([tip_likelihood(prob, s) for s in prob.states[node]])  # ptn x n
            return np.repeat(T[None], ncat, axis=0), np.zeros(prob.nptn)
        acc = np.ones((ncat, prob.nptn, prob.n))
        lsc = np.zeros(prob.nptn)
        for k in range(3):
            ch = nei[node, k]
            if ch < 0 or ch == dad:
                continue
            Lc, sc = down(ch, node)
            t = blen[node, k]
            for c in range(ncat):
                acc[c] *= Lc[c] @ pmat(t, c).T
            lsc += sc
        m = acc.max(axis=(0, 2))
        m[m == 0] = 1.0
        acc /= m[None, :, None]
        return acc, lsc + np.log(m)

    La, sa = down(a, b)
    Lb, sb = down(b, a)
    t = length(a, b)
    lh = np.zeros(prob.nptn)
    for c in range(ncat):
        lh += prob.props[c] * np.einsum("px,px->p", La[c] * prob.pi[None, :], Lb[c] @ pmat(t, c).T)
    lp = np.log(lh) + sa + sb
    inv = getattr(prob, "invar", None)
    if inv is not None:
        # + invariant-site mass, a true (unscaled) value: log(lh e^(sa+sb) + invar)
        with np.errstate(divide="ignore"):
            lp = np.logaddexp(lp, np.log(np.asarray(inv)))
    return lp, float(np.dot(lp, prob.freq))


def pruning_lh_cat(prob, a, b):
    """True per-category pattern likelihoods w_c L_c(p) (ncat x nptn) at branch (a, b) by
    the same independent real-space pruning as pruning_loglik (without ptn_invar)."""
    ncat = len(prob.rates)
    P = {}

    def pmat(t, c):
        key = (round(t, 17), c)
        if key not in P:
            P[key] = scipy.linalg.expm(prob.Q * prob.rates[c] * t)
        return P[key]

    def down(node, dad):
        if node < prob.ntaxa:
            T = np.stack([tip_likelihood(prob, s) for s in prob.states[node]])
            return np.repeat(T[None], ncat, axis=0), np.zeros(prob.nptn)
        acc = np.ones((ncat, prob.nptn, prob.n))
        lsc = np.zeros(prob.nptn)
        for k in range(3):
            ch = prob.nei[node, k]
            if ch < 0 or ch == dad:
                continue
            Lc, sc = down(ch, node)
            for c in range(ncat):
                acc[c] *= Lc[c] @ pmat(prob.blen[node, k], c).T
            lsc += sc
        m = acc.max(axis=(0, 2))
        m[m == 0] = 1.0
        return acc / m[None, :, None], lsc + np.log(m)

    La, sa = down(a, b)
    Lb, sb = down(b, a)
    t = prob.blen[a, list(prob.nei[a]).index(b)]
    cat = np.stack([prob.props[c] * np.einsum("px,px->p", La[c] * prob.pi[None, :], Lb[c] @ pmat(t, c).T)
                    for c in range(ncat)])
    return cat * np.exp(sa + sb)[None, :]


def ptn_invar_identity(prob, p_inv):
    """Invariant-site mass of each pattern from its definition: p_inv times the probability
    that every taxon shows one common unchanged state, i.e. p_inv * sum_x pi_x * prod_t
    [state of taxon t allows x].  Independent of the product's const_char coding."""
    A = np.stack([tip_likelihood(prob, s) for s in range(prob.su + 1)])  # allowed-state indicators
    ind = np.prod(A[np.asarray(prob.states)], axis=0)  # nptn x n
    return p_inv * (ind @ np.asarray(prob.pi))


def apply_pinvar(prob, p_inv):
    """+I+G on a problem: props (1 - p)/ncat, Gamma rates / (1 - p)
    (model/rategammainvar.h:69, model/rategamma.cpp:129-140) and ptn_invar."""
    ncat = len(prob.rates)
    prob.props = np.full(ncat, (1.0 - p_inv) / ncat)
    prob.rates = np.ascontiguousarray(np.asarray(prob.rates) / (1.0 - p_inv))
    prob.invar = ptn_invar_identity(prob, p_inv)
    return prob


def plant_constant_columns(prob, frac, rng):
    """Overwrite a fraction of the pattern columns with constant ones (one state per column,
    some taxa then shown as an ambiguity code allowing it, or unknown); keeps the pattern
    count and frequencies (duplicated columns stay separate patterns, which the likelihood
    does not care about)."""
    st = np.array(prob.states, dtype=np.uint8)
    T, nptn = st.shape
    cols = rng.choice(nptn, size=max(1, int(frac * nptn)), replace=False)
    for j in cols:
        x = int(rng.integers(prob.n))
        col = np.full(T, x, dtype=np.uint8)
        k = int(rng.integers(0, 3))
        who = rng.choice(T, size=k, replace=False)
        for t in who:
            if prob.seq_type == SEQ_DNA:
                # an ambiguity code whose mask includes x, or unknown
                masks = [m for m in range(1, 16) if m & (1 << x) and bin(m).count("1") >= 2]
                m = int(rng.choice(masks))
                col[t] = m + 3
            elif prob.seq_type == SEQ_PROTEIN and x in (2, 3, 5, 6, 9, 10):
                col[t] = 20 + [(2, 3), (5, 6), (9, 10)].index(next(pr for pr in [(2, 3), (5, 6), (9, 10)] if x in pr))
            else:
                col[t] = prob.su
        st[:, j] = col
    # one all-unknown column
    st[:, cols[0]] = prob.su
    prob.states = np.ascontiguousarray(st)
    return prob


def brute_force_loglik(prob, a, b):
    """Enumerate all internal-state assignments (tiny trees only).  Edges are oriented
    away from an internal root: L = sum pi_root * prod P[parent][child]."""
    import itertools
    ntaxa, n = prob.ntaxa, prob.n
    nn = 2 * ntaxa - 2
    internals = list(range(ntaxa, nn))
    root = internals[0]
    directed = []
    stack = [(root, -1)]
    while stack:
        u, par = stack.pop()
        for k in range(3):
            v = prob.nei[u, k]
            if v < 0 or v == par:
                continue
            directed.append((u, v, prob.blen[u, k]))
            if v >= ntaxa:
                stack.append((v, u))
    ncat = len(prob.rates)
    tips = np.stack([[tip_likelihood(prob, s) for s in prob.states[i]] for i in range(ntaxa)])  # T x ptn x n
    total = np.zeros(prob.nptn)
    for c in range(ncat):
        Pm = {(u, v): scipy.linalg.expm(prob.Q * prob.rates[c] * t) for (u, v, t) in directed}
        lc = np.zeros(prob.nptn)
        for assign in itertools.product(range(n), repeat=len(internals)):
            st = dict(zip(internals, assign))
            term = np.full(prob.nptn, prob.pi[st[root]])
            for (u, v, t) in directed:
                M = Pm[(u, v)]
                if v >= ntaxa:
                    term = term * M[st[u], st[v]]
                else:
                    term = term * (tips[v] @ M[st[u]])
            lc += term
        total += prob.props[c] * lc
    if getattr(prob, "invar", None) is not None:
        total = total + np.asarray(prob.invar)
    lp = np.log(total)
    return lp, float(np.dot(lp, prob.freq))


class OracleRell:
    """UFBoot replicate state scored by the oracle's RELL loop (orc_rell_save_tree)."""

    def __init__(self, counts):
        self.counts = np.ascontiguousarray(counts, dtype=np.float32)
        self.ns, self.nptn = self.counts.shape
        self.logl = np.full(self.ns, -np.finfo(np.float64).max)
        self.orig = np.full(self.ns, -np.finfo(np.float64).max)
        self.best_counts = np.zeros(self.ns, dtype=np.int32)

    def save_tree(self, ptn_lh, cur_logl, eps=0.5, uniform=None):
        lib = oracle_lib()
        ptn_lh = np.ascontiguousarray(ptn_lh, dtype=np.float64)
        rell = np.zeros(self.ns)
        improved = np.zeros(self.ns, dtype=np.int32)
        i32 = ctypes.POINTER(ctypes.c_int32)
        u = None if uniform is None else _dp(np.ascontiguousarray(uniform, dtype=np.float64))
        lib.orc_rell_save_tree(self.ns, self.nptn, self.counts.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                               _dp(ptn_lh), cur_logl, eps, u, _dp(rell), _dp(self.logl),
                               self.best_counts.ctypes.data_as(i32), _dp(self.orig), improved.ctypes.data_as(i32))
        return rell, improved


# ---------------------------------------------------------------------------------------
# scale_num counts derived from the reference's rule, independently of the oracle


def derived_scale_counts(prob, dad, node, safe, rule="reference", memo=None):
    """Scale counts of node's subtree seen from dad, derived from the reference's rule
    (tree/phylokernelnew.h:1643-1644 tip-tip: 0 and no check; 1757-1760 tip-internal: the
    internal child's count; 1929-1936 internal-internal: the sum; then 1997-2034 /
    1893-1907: one increment when the largest |entry| of the eigen-space partial, over
    the whole pattern block (NORM) or over one category (SAFE), is below 2^-256 and
    ptn_invar is 0).

    Nothing here shares the oracle's arithmetic.  Each direction carries the real-space
    conditional likelihood as a mantissa per (category, pattern, state) times 2^E with an
    integer exponent E per (category, pattern), propagated with scipy expm and
    renormalised with frexp after every node, so the true values never underflow.  The
    eigen-space partial the rule inspects is U^-1 applied to that vector
    (tree/phylokernelnew.h:1989-1994), and the value the kernel checks is the true one
    times 2^(256 * children's counts), all scalings being exact powers of two.  So the
    count follows from log2 of the true maximum alone.

    rule selects planted variants used to show the check is sensitive:
      "reference"        the rule above
      "last_category"    NORM check on the last category's maximum only (a per-category
                         maximum left over from SAFE)
      "threshold_2^-255" a threshold one binade too high
      "before_projection" the maximum taken over the real-space product, before U^-1

    Returns (counts, ambiguous): counts [ptn] (NORM) or [ptn, ncat] (SAFE) as int64, and
    a boolean mask of patterns whose checked maximum lies within 1e-6 (log2 units) of the
    threshold, where the rounding of either computation could decide the count.
    """
    if memo is None:
        memo = {}
    nei, blen = prob.nei, prob.blen
    ncat, n, nptn = len(prob.rates), prob.n, prob.nptn
    invar = np.zeros(nptn) if getattr(prob, "invar", None) is None else np.asarray(prob.invar)
    thr = -255.0 if rule == "threshold_2^-255" else -256.0
    pcache = memo.setdefault("_P", {})
    done = memo.setdefault((bool(safe), rule), {})

    def pmat(t, c):
        key = (float(t), c)
        if key not in pcache:
            pcache[key] = scipy.linalg.expm(prob.Q * prob.rates[c] * t)
        return pcache[key]

    def tips(node_):
        T = np.stack([tip_likelihood(prob, s) for s in range(prob.su + 1)])
        return np.repeat(T[prob.states[node_]][None], ncat, axis=0)  # c x ptn x n

    def walk(d, v):
        # returns (mantissa c x ptn x n, exponent c x ptn, counts, ambiguous, is_tip)
        key = (d, v)
        if key in done:
            return done[key]
        if v < prob.ntaxa:
            res = (tips(v), np.zeros((ncat, nptn), dtype=np.int64), None, np.zeros(nptn, bool), True)
            done[key] = res
            return res
        kids = [(int(nei[v, k]), blen[v, k]) for k in range(3) if nei[v, k] >= 0 and nei[v, k] != d]
        assert len(kids) == 2
        sub = [walk(v, ch) for (ch, _t) in kids]
        mant = np.ones((ncat, nptn, n))
        expo = np.zeros((ncat, nptn), dtype=np.int64)
        for (ch, t), (m, e, _s, _a, _tip) in zip(kids, sub):
            for c in range(ncat):
                mant[c] *= m[c] @ pmat(t, c).T
            expo += e
        # renormalise: mantissa maximum in [0.5, 1)
        mx = np.abs(mant).max(axis=2)
        _f, ex = np.frexp(np.where(mx > 0, mx, 1.0))
        mant = mant / np.ldexp(1.0, ex)[..., None]
        expo = expo + ex
        eig = np.einsum("jx,cpx->cpj", prob.inv_evec, mant)
        src = mant if rule == "before_projection" else eig
        with np.errstate(divide="ignore"):
            lmax = np.log2(np.abs(src).max(axis=2)) + expo  # c x ptn, log2 of the true maximum
        tip_flags = [s[4] for s in sub]
        child_counts = [s[2] for s in sub if not s[4]]
        shape = (nptn, ncat) if safe else (nptn,)
        base = np.zeros(shape, dtype=np.int64)
        for cc in child_counts:
            base = base + cc
        amb = np.zeros(nptn, bool)
        for s in sub:
            amb |= s[3]
        if all(tip_flags):
            counts = base  # tip-tip: 0, no check
        elif safe:
            pre = lmax.T + 256.0 * base  # ptn x c
            hit = (pre < thr) & (invar[:, None] == 0.0)
            amb |= np.any(np.abs(pre - thr) < 1e-6, axis=1)
            counts = base + hit
        else:
            m = lmax[-1] if rule == "last_category" else lmax.max(axis=0)
            pre = m + 256.0 * base
            hit = (pre < thr) & (invar == 0.0)
            amb |= np.abs(pre - thr) < 1e-6
            counts = base + hit
        res = (mant, expo, counts, amb, False)
        done[key] = res
        return res

    _m, _e, counts, amb, is_tip = walk(dad, node)
    if is_tip:
        raise ValueError("a tip has no scale counts")
    return counts, amb
