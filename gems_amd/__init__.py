"""Python binding (ctypes) of libplk, the MI355X Felsenstein likelihood engine.

The engine itself is C++ host code plus hand-written gfx950 HIP kernels behind the C
ABI in include/plk.h, include/plk_tree.h and include/plk_model.h.  This module only
loads that library and wraps its entry points for tests and benchmarks; there is no
Python or CPU fallback for any likelihood computation.
"""
import ctypes
import pathlib

import numpy as np

__all__ = ["lib", "LibraryMissing", "PlkError", "Problem", "Tree", "Context", "FP64", "FP32",
           "NORM", "SAFE", "SEQ_DNA", "SEQ_PROTEIN", "SEQ_CODON", "device_count"]

HERE = pathlib.Path(__file__).resolve().parent
LIB_PATH = HERE / "libplk.so"

FP64, FP32 = 0, 1
NORM, SAFE = 0, 1
SEQ_DNA, SEQ_PROTEIN, SEQ_CODON = 0, 1, 2


class LibraryMissing(RuntimeError):
    pass


class PlkError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"plk error {code}: {msg}")
        self.code = code


_lib = None

_P = ctypes.c_void_p
_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int32)
_i64p = ctypes.POINTER(ctypes.c_int64)
_u8p = ctypes.POINTER(ctypes.c_uint8)
_u16p = ctypes.POINTER(ctypes.c_uint16)
_int = ctypes.c_int
_dbl = ctypes.c_double


class plk_request(ctypes.Structure):
    _fields_ = [("nstates", ctypes.c_int), ("ncat", ctypes.c_int), ("ntaxa", ctypes.c_int), ("nptn", ctypes.c_int),
                ("precision", ctypes.c_int), ("scaling", ctypes.c_int), ("state_unknown", ctypes.c_int),
                ("eval", _dp), ("evec", _dp), ("inv_evec", _dp), ("rates", _dp), ("props", _dp), ("tip_lh", _dp),
                ("states", _u8p), ("freq", _dp), ("invar", _dp), ("nei", _ip), ("len", _dp),
                ("kind", ctypes.c_int), ("opt_a", ctypes.c_int), ("opt_b", ctypes.c_int), ("iterations", ctypes.c_int),
                ("tol", ctypes.c_double),
                ("logl_before", ctypes.c_double), ("logl_after", ctypes.c_double), ("opt_len", ctypes.c_double),
                ("status", ctypes.c_int), ("t_submit", ctypes.c_double), ("t_admit", ctypes.c_double),
                ("t_done", ctypes.c_double), ("footprint", ctypes.c_int64)]


class plk_op(ctypes.Structure):
    _fields_ = [("out_slot", ctypes.c_int32), ("child", ctypes.c_int32 * 2), ("reserved", ctypes.c_int32),
                ("len", ctypes.c_double * 2)]


_SIGS = {
    "plk_last_error": (ctypes.c_char_p, []),
    "plk_version": (_int, []),
    "plk_device_count": (_int, []),
    "plk_device_units": (_int, []),
    "plk_create": (_int, [ctypes.POINTER(_P), _int, _int, _int, _int, _int, _int, _int]),
    "plk_destroy": (_int, [_P]),
    "plk_memory_required": (ctypes.c_int64, [_int] * 7),
    "plk_set_patterns": (_int, [_P, _u8p, _dp, _dp]),
    "plk_set_model": (_int, [_P, _dp, _dp, _dp, _dp, _dp, _int, _dp]),
    "plk_update_partials": (_int, [_P, ctypes.POINTER(plk_op), _int]),
    "plk_flush": (_int, [_P]),
    "plk_branch_loglik": (_int, [_P, _int, _int, _dbl, _dp, _dp]),
    "plk_branch_derivs": (_int, [_P, _int, _int, _dbl, _int, _dp, _dp]),
    "plk_loglik_from_theta": (_int, [_P, _dbl, _dp]),
    "plk_newton_branch": (_int, [_P, _int, _int, _dbl, _dbl, _dbl, _dbl, _int, _dp, _ip]),
    "plk_newton_multi": (_int, [_P, _int, _ip, _ip, _dbl, _dp, _dbl, _dbl, _int, _dp, _dp, _ip]),
    # alignment ingest
    "plk_align_read": (_int, [ctypes.POINTER(_P), ctypes.c_char_p, _int, _int]),
    "plk_align_parse": (_int, [ctypes.POINTER(_P), ctypes.c_char_p, ctypes.c_size_t, _int, _int]),
    "plk_align_destroy": (_int, [_P]),
    "plk_align_dims": (_int, [_P, _ip]),
    "plk_align_name": (_int, [_P, _int, ctypes.c_char_p, _int]),
    "plk_align_patterns": (_int, [_P, _u8p, _dp, _ip, _ip]),
    # UFBoot RELL
    "plk_rell_create": (_int, [ctypes.POINTER(_P), _int, _int, ctypes.POINTER(ctypes.c_float)]),
    "plk_rell_destroy": (_int, [_P]),
    "plk_rell_eval": (_int, [_P, _P, _dp]),
    "plk_rell_save_tree": (_int, [_P, _P, _dbl, _dbl, _dp, _ip, ctypes.POINTER(_int)]),
    "plk_rell_get": (_int, [_P, _dp, _ip, _dp]),
    "plk_rell_eval_noalloc": (_int, [_P, _int, _int, ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p]),
    "plk_get_partial": (_int, [_P, _int, _dp, _u16p]),
    "plk_set_timing": (_int, [_P, _int]),
    "plk_get_timing": (_int, [_P, _int, _dp, _i64p]),
    "plk_updates": (ctypes.c_int64, [_P]),
    "plk_synchronize": (_int, [_P]),
    "plk_get_exponents": (_int, [_P, _int, _ip]),
    "plk_set_tile_shape": (_int, [_P, _int]),
    "plk_set_level_waves": (_int, [_P, _int]),
    "plk_set_f64_form": (_int, [_P, _int]),
    "plk_get_pattern_lh_cat": (_int, [_P, _dp]),
    "plk_get_pattern_scaling": (_int, [_P, _dp, _dp]),
    "plk_get_pattern_dlogl": (_int, [_P, _dp]),
    "plk_probe_copy": (_int, [ctypes.c_int64, _int, _dp]),
    # model / synthetic data
    "plk_state_unknown": (_int, [_int, _int]),
    "plk_rate_matrix": (_int, [_int, _dp, _dp, _dp]),
    "plk_eigen": (_int, [_int, _dp, _dp, _dp, _dp, _dp]),
    "plk_gamma_rates": (_int, [_dbl, _int, _dp]),
    "plk_gamma_rates_median": (_int, [_dbl, _int, _dp]),
    "plk_tip_lh": (_int, [_int, _int, _dp, _dp]),
    "plk_ptn_invar": (_int, [_int, _int, _int, _int, _u8p, _dp, _dbl, _dp]),
    "plk_codon_exch": (_int, [_dbl, _dbl, _dp]),
    "plk_sim_create": (_int, [ctypes.POINTER(_P), _int, _int, _int, _dbl, _dbl, _int, ctypes.c_uint64]),
    "plk_sim_destroy": (_int, [_P]),
    "plk_sim_dims": (_int, [_P, _ip]),
    "plk_sim_array": (ctypes.c_void_p, [_P, ctypes.c_char_p, _i64p]),
    # tree layer
    "plk_tree_create": (_int, [ctypes.POINTER(_P), _int, _int, _int, _int, _int, _int]),
    "plk_tree_destroy": (_int, [_P]),
    "plk_tree_context": (_P, [_P]),
    "plk_tree_set_model": (_int, [_P, _dp, _dp, _dp, _dp, _dp, _int, _dp]),
    "plk_tree_set_patterns": (_int, [_P, _u8p, _dp, _dp]),
    "plk_tree_set_topology": (_int, [_P, _ip, _dp]),
    "plk_tree_set_length": (_int, [_P, _int, _int, _dbl]),
    "plk_tree_get_length": (_int, [_P, _int, _int, _dp]),
    "plk_tree_clear_all": (_int, [_P]),
    "plk_tree_loglik": (_int, [_P, _int, _int, _dp, _dp]),
    "plk_tree_compute_likelihood": (_int, [_P, _dp]),
    "plk_tree_derivs": (_int, [_P, _int, _int, _int, _dp, _dp]),
    "plk_tree_loglik_from_theta": (_int, [_P, _int, _int, _dp]),
    "plk_tree_get_partial": (_int, [_P, _int, _int, _dp, _u16p]),
    "plk_tree_optimize_one_branch": (_int, [_P, _int, _int, _dbl, _dbl, _int, _dp]),
    "plk_tree_set_newton": (_int, [_P, _int]),
    "plk_tree_nni_evaluate": (_int, [_P, _int, _ip, _ip, _dbl, _dbl, _int, _dp, _dp]),
    "plk_tree_optimize_all_branches": (_int, [_P, _int, _dbl, _dbl, _dbl, _int, _dp]),
    "plk_tree_full_traversal": (_int, [_P, _dp]),
    "plk_tree_counters": (_int, [_P, _i64p]),
    "plk_tree_create_ex": (_int, [ctypes.POINTER(_P), _int, _int, _int, _int, _int, _int, _int]),
    "plk_tree_all_branch_derivs": (_int, [_P, _ip, _ip, _dp, _dp]),
    "plk_branch_derivs_multi": (_int, [_P, _int, ctypes.POINTER(_int), ctypes.POINTER(_int), _dp, _dp, _dp]),
    "plk_tree_autotune": (_int, [_P, _int, ctypes.POINTER(_int), _dp]),
    "plk_tree_queue_branch": (_int, [_P, _int, _int, _int, ctypes.POINTER(_int), ctypes.POINTER(_int), _dp]),
    "plk_tree_set_gamma_shape": (_int, [_P, _dbl]),
    "plk_tree_get_gamma_shape": (_int, [_P, _dp]),
    "plk_tree_set_gamma_median": (_int, [_P, _int]),
    "plk_tree_set_gtr": (_int, [_P, _int, _dp, _dp]),
    "plk_tree_get_gtr": (_int, [_P, _dp]),
    "plk_tree_optimize_model": (_int, [_P, _dbl, _dp]),
    "plk_tree_optimize_gamma": (_int, [_P, _dbl, _dbl, _dbl, _dp, _dp]),
    "plk_tree_optimize_parameters": (_int, [_P, _int, _dbl, _dbl, _dbl, _dbl, _dbl, _dbl, _int, _dp, _dp]),
    "plk_sim_create_on_tree": (_int, [ctypes.POINTER(_P), _int, _int, _ip, _int, _dbl, _dbl, _int, ctypes.c_uint64]),
    # batched partitions
    "plk_batch_create": (_int, [ctypes.POINTER(_P), ctypes.POINTER(_P), _int]),
    "plk_batch_destroy": (_int, [_P]),
    "plk_batch_set_launch": (_int, [_P, _int, _int]),
    "plk_batch_loglik": (_int, [_P, ctypes.POINTER(_int), ctypes.POINTER(_int), _dp, _dp, _dp]),
    "plk_batch_autotune": (_int, [_P, ctypes.POINTER(_int), ctypes.POINTER(_int), _dp, _int, ctypes.POINTER(_int), _dp]),
    "plk_batch_timing": (_int, [_P, _dp, _i64p, _i64p]),
    "plk_batch_items": (_int, [_P]),
    # serving
    "plk_server_create": (_int, [ctypes.POINTER(_P), _int, ctypes.c_int64]),
    "plk_server_submit": (_int, [_P, ctypes.c_void_p]),
    "plk_server_drain": (_int, [_P]),
    "plk_server_now": (ctypes.c_double, [_P]),
    "plk_server_stats": (_int, [_P, _dp]),
    "plk_server_destroy": (_int, [_P]),
    "plk_server_counters": (_int, [_P, _i64p]),
    "plk_reshape": (_int, [_P, _int, _int, _int]),
    "plk_tree_reshape": (_int, [_P, _int, _int]),
    "plk_batch_rebind": (_int, [_P, ctypes.POINTER(_P), _int]),
}


def lib():
    """Load libplk.so (raises LibraryMissing if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise LibraryMissing(f"{LIB_PATH} is not built; run __graft_entry__.build() or make -C gems_amd")
    L = ctypes.CDLL(str(LIB_PATH))
    for name, (res, args) in _SIGS.items():
        try:
            fn = getattr(L, name)
        except AttributeError:
            continue
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def declared_symbols():
    return list(_SIGS)


def _check(rc):
    if rc != 0:
        msg = lib().plk_last_error()
        raise PlkError(rc, msg.decode() if msg else "")
    return rc


def _d(a):
    return a.ctypes.data_as(_dp)


def device_count():
    return lib().plk_device_count()


def device_units():
    return lib().plk_device_units()


def probe_copy(nbytes=1 << 30, reps=5):
    """Measured device-to-device copy bandwidth in GB/s (read + write bytes)."""
    v = ctypes.c_double()
    _check(lib().plk_probe_copy(nbytes, reps, ctypes.byref(v)))
    return v.value


def ptn_invar(n, seq_type, states, pi, p_invar):
    """plk_ptn_invar: invariant-site mass per pattern for proportion p_invar."""
    st = np.ascontiguousarray(states, dtype=np.uint8)
    pi = np.ascontiguousarray(pi, dtype=np.float64)
    out = np.zeros(st.shape[1])
    _check(lib().plk_ptn_invar(n, seq_type, st.shape[0], st.shape[1], st.ctypes.data_as(_u8p), _d(pi), p_invar,
                               _d(out)))
    return out


class Problem:
    """Seeded synthetic problem generated by the library (plk_sim_create)."""

    def __init__(self, seq_type, ntaxa, nsites, mean_len=0.1, alpha=0.5, ncat=4, seed=1, topology=None):
        L = lib()
        h = _P()
        if topology is None:
            _check(L.plk_sim_create(ctypes.byref(h), seq_type, ntaxa, nsites, mean_len, alpha, ncat, seed))
        else:
            nei = np.ascontiguousarray(topology, dtype=np.int32)
            _check(L.plk_sim_create_on_tree(ctypes.byref(h), seq_type, ntaxa, nei.ctypes.data_as(_ip), nsites,
                                            mean_len, alpha, ncat, seed))
        self._h = h
        dims = np.zeros(7, dtype=np.int32)
        _check(L.plk_sim_dims(h, dims.ctypes.data_as(_ip)))
        (self.n, self.ntaxa, self.nsites, self.nptn, self.ncat, self.su, self.seq_type) = [int(x) for x in dims]
        for name in ("pi", "Q", "eval", "evec", "inv_evec", "rates", "props", "tip_lh", "freq", "blen", "params"):
            setattr(self, name, self._array(name, np.float64))
        self.nei = self._array("nei", np.int32).reshape(-1, 3)
        self.blen = self.blen.reshape(-1, 3)
        self.states = self._array("states", np.uint8).reshape(self.ntaxa, self.nptn)
        self.evec = self.evec.reshape(self.n, self.n)
        self.inv_evec = self.inv_evec.reshape(self.n, self.n)
        self.Q = self.Q.reshape(self.n, self.n)
        self.tip_lh = self.tip_lh.reshape(self.su + 1, self.n)
        self.invar = None
        L.plk_sim_destroy(h)
        self._h = None

    def _array(self, name, dtype):
        cnt = ctypes.c_int64()
        ptr = lib().plk_sim_array(self._h, name.encode(), ctypes.byref(cnt))
        if not ptr:
            raise PlkError(-1, f"no array {name}")
        n = cnt.value
        buf = (ctypes.c_char * (n * np.dtype(dtype).itemsize)).from_address(ptr)
        return np.frombuffer(bytes(buf), dtype=dtype).copy()


class Tree:
    """PhyloTreeGPU handle (plk_tree_*)."""

    def __init__(self, prob, precision=FP64, scaling=NORM, all_directions=False):
        L = lib()
        self.p = prob
        h = _P()
        _check(L.plk_tree_create_ex(ctypes.byref(h), prob.n, len(prob.rates), prob.ntaxa, prob.nptn, precision, scaling,
                                    int(all_directions)))
        self.h = h
        self.set_model(prob)
        st = np.ascontiguousarray(prob.states, dtype=np.uint8)
        fr = np.ascontiguousarray(prob.freq, dtype=np.float64)
        inv = None if getattr(prob, "invar", None) is None else np.ascontiguousarray(prob.invar, dtype=np.float64)
        _check(L.plk_tree_set_patterns(h, st.ctypes.data_as(_u8p), _d(fr), _d(inv) if inv is not None else None))
        nei = np.ascontiguousarray(prob.nei, dtype=np.int32)
        bl = np.ascontiguousarray(prob.blen, dtype=np.float64)
        _check(L.plk_tree_set_topology(h, nei.ctypes.data_as(_ip), _d(bl)))

    def set_model(self, m):
        tip = np.ascontiguousarray(m.tip_lh, dtype=np.float64).ravel()
        su = tip.size // m.n - 1
        arrs = [np.ascontiguousarray(a, dtype=np.float64).ravel() for a in (m.eval, m.evec, m.inv_evec, m.rates, m.props)]
        _check(lib().plk_tree_set_model(self.h, *[_d(a) for a in arrs], su, _d(tip)))

    def close(self):
        if getattr(self, "h", None):
            lib().plk_tree_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def context(self):
        return lib().plk_tree_context(self.h)

    def loglik(self, a, b, with_patterns=False):
        v = ctypes.c_double()
        pl = np.zeros(self.p.nptn) if with_patterns else None
        _check(lib().plk_tree_loglik(self.h, a, b, ctypes.byref(v), _d(pl) if pl is not None else None))
        return (v.value, pl) if with_patterns else v.value

    def compute_likelihood(self):
        """computeLikelihood(): logL at the first leaf's branch, recomputing only dirty partials."""
        v = ctypes.c_double()
        _check(lib().plk_tree_compute_likelihood(self.h, ctypes.byref(v)))
        return v.value

    def full_traversal(self):
        v = ctypes.c_double()
        _check(lib().plk_tree_full_traversal(self.h, ctypes.byref(v)))
        return v.value

    def derivs(self, a, b, reuse=False):
        df, ddf = ctypes.c_double(), ctypes.c_double()
        rc = lib().plk_tree_derivs(self.h, a, b, int(reuse), ctypes.byref(df), ctypes.byref(ddf))
        return df.value, ddf.value, rc

    def loglik_from_theta(self, a, b):
        v = ctypes.c_double()
        _check(lib().plk_tree_loglik_from_theta(self.h, a, b, ctypes.byref(v)))
        return v.value

    def set_length(self, a, b, t):
        _check(lib().plk_tree_set_length(self.h, a, b, t))

    def get_length(self, a, b):
        v = ctypes.c_double()
        _check(lib().plk_tree_get_length(self.h, a, b, ctypes.byref(v)))
        return v.value

    def partial(self, dad, node, safe=False):
        n, ncat = self.p.n, len(self.p.rates)
        out = np.zeros(self.p.nptn * n * ncat)
        sc = np.zeros(self.p.nptn * (ncat if safe else 1), dtype=np.uint16)
        _check(lib().plk_tree_get_partial(self.h, dad, node, _d(out), sc.ctypes.data_as(_u16p)))
        return out.reshape(self.p.nptn, ncat, n), sc

    def optimize_one(self, a, b, min_bl=1e-6, max_bl=10.0, max_nr=100):
        v = ctypes.c_double()
        _check(lib().plk_tree_optimize_one_branch(self.h, a, b, min_bl, max_bl, max_nr, ctypes.byref(v)))
        return v.value

    def nni_evaluate(self, branches, min_bl=1e-6, max_bl=10.0, max_nr=10):
        """Both NNI5 moves of each internal branch (u, v): (scores[nbr, 2], lens[nbr, 2, 5])."""
        n1 = np.array([b[0] for b in branches], dtype=np.int32)
        n2 = np.array([b[1] for b in branches], dtype=np.int32)
        sc = np.zeros(2 * len(branches))
        ln = np.zeros(10 * len(branches))
        _check(lib().plk_tree_nni_evaluate(self.h, len(branches), n1.ctypes.data_as(_ip), n2.ctypes.data_as(_ip),
                                           min_bl, max_bl, max_nr, _d(sc), _d(ln)))
        return sc.reshape(-1, 2), ln.reshape(-1, 2, 5)

    def newton_multi(self, refs, guesses, min_bl=1e-6, max_bl=10.0, max_nr=100):
        """plk_newton_multi on node-reference pairs: (lengths, logls, status)."""
        a = np.array([r[0] for r in refs], dtype=np.int32)
        b = np.array([r[1] for r in refs], dtype=np.int32)
        g = np.ascontiguousarray(guesses, dtype=np.float64)
        out, lg = np.zeros(len(refs)), np.zeros(len(refs))
        st = np.zeros(len(refs), dtype=np.int32)
        _check(lib().plk_newton_multi(self.context(), len(refs), a.ctypes.data_as(_ip), b.ctypes.data_as(_ip), min_bl,
                                      _d(g), max_bl, min_bl, max_nr, _d(out), _d(lg), st.ctypes.data_as(_ip)))
        return out, lg, st

    def set_newton(self, device_resident):
        """Newton solves on the device (True, default) or as a host loop (False)."""
        _check(lib().plk_tree_set_newton(self.h, 1 if device_resident else 0))

    def optimize_all(self, iterations=100, tol=1e-3, min_bl=1e-6, max_bl=10.0, max_nr=100):
        v = ctypes.c_double()
        _check(lib().plk_tree_optimize_all_branches(self.h, iterations, tol, min_bl, max_bl, max_nr, ctypes.byref(v)))
        return v.value

    def set_gamma_shape(self, alpha):
        _check(lib().plk_tree_set_gamma_shape(self.h, alpha))

    def set_gtr(self, exch, pi, seq_type=None):
        exch = np.ascontiguousarray(exch, dtype=np.float64)
        pi = np.ascontiguousarray(pi, dtype=np.float64)
        st = self.p.seq_type if seq_type is None else seq_type
        _check(lib().plk_tree_set_gtr(self.h, st, _d(exch), _d(pi)))

    def gtr(self):
        n = self.p.n
        out = np.zeros(n * (n - 1) // 2)
        _check(lib().plk_tree_get_gtr(self.h, _d(out)))
        return out

    def optimize_model(self, gradient_epsilon=1e-4):
        """ModelMarkov::optimizeParameters; returns (exchangeabilities, logL)."""
        v = ctypes.c_double()
        _check(lib().plk_tree_optimize_model(self.h, gradient_epsilon, ctypes.byref(v)))
        return self.gtr(), v.value

    def set_gamma_median(self, median):
        _check(lib().plk_tree_set_gamma_median(self.h, int(median)))

    def gamma_shape(self):
        v = ctypes.c_double()
        _check(lib().plk_tree_get_gamma_shape(self.h, ctypes.byref(v)))
        return v.value

    def optimize_gamma(self, gradient_epsilon=1e-4, min_gamma=0.02, max_gamma=1000.0):
        """RateGamma::optimizeParameters; returns (alpha, logL)."""
        a, v = ctypes.c_double(), ctypes.c_double()
        _check(lib().plk_tree_optimize_gamma(self.h, gradient_epsilon, min_gamma, max_gamma, ctypes.byref(a),
                                             ctypes.byref(v)))
        return a.value, v.value

    def optimize_parameters(self, num_iterations=100, logl_epsilon=0.01, gradient_epsilon=1e-4, min_gamma=0.02,
                            max_gamma=1000.0, min_bl=1e-6, max_bl=10.0, max_nr=100):
        """ModelFactory::optimizeParameters (branch lengths + Gamma shape); returns (alpha, logL)."""
        a, v = ctypes.c_double(), ctypes.c_double()
        _check(lib().plk_tree_optimize_parameters(self.h, num_iterations, logl_epsilon, gradient_epsilon, min_gamma,
                                                  max_gamma, min_bl, max_bl, max_nr, ctypes.byref(a), ctypes.byref(v)))
        return a.value, v.value

    def counters(self):
        out = np.zeros(3, dtype=np.int64)
        _check(lib().plk_tree_counters(self.h, out.ctypes.data_as(_i64p)))
        return out

    def all_branch_derivs(self):
        nb = 2 * self.p.ntaxa - 3
        a, b = np.zeros(nb, dtype=np.int32), np.zeros(nb, dtype=np.int32)
        df, ddf = np.zeros(nb), np.zeros(nb)
        _check(lib().plk_tree_all_branch_derivs(self.h, a.ctypes.data_as(_ip), b.ctypes.data_as(_ip), _d(df), _d(ddf)))
        return list(zip(a.tolist(), b.tolist())), df, ddf

    def autotune(self, reps=3):
        w, ms = _int(), ctypes.c_double()
        _check(lib().plk_tree_autotune(self.h, reps, ctypes.byref(w), ctypes.byref(ms)))
        return w.value, ms.value

    def set_tile_shape(self, wpc):
        _check(lib().plk_set_tile_shape(self.context(), wpc))

    def pattern_lh_cat(self):
        out = np.zeros(self.p.nptn * len(self.p.rates))
        _check(lib().plk_get_pattern_lh_cat(self.context(), _d(out)))
        return out.reshape(self.p.nptn, len(self.p.rates))

    def pattern_scaling(self):
        lh, sc = np.zeros(self.p.nptn), np.zeros(self.p.nptn)
        _check(lib().plk_get_pattern_scaling(self.context(), _d(lh), _d(sc)))
        return lh, sc

    def pattern_dlogl(self):
        out = np.zeros(self.p.nptn)
        _check(lib().plk_get_pattern_dlogl(self.context(), _d(out)))
        return out

    def set_f64_form(self, form):
        _check(lib().plk_set_f64_form(self.context(), form))

    def set_level_waves(self, waves):
        _check(lib().plk_set_level_waves(self.context(), waves))

    def set_timing(self, on):
        _check(lib().plk_set_timing(self.context(), int(on)))

    def timing(self, kind):
        ms, n = ctypes.c_double(), ctypes.c_int64()
        _check(lib().plk_get_timing(self.context(), kind, ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value

    def queue_branch(self, a, b, clear=False):
        ra, rb, t = ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
        _check(lib().plk_tree_queue_branch(self.h, a, b, int(clear), ctypes.byref(ra), ctypes.byref(rb),
                                           ctypes.byref(t)))
        return ra.value, rb.value, t.value


class Batch:
    """Batched unlinked partitions (plk_batch_*): one launch per traversal for all trees."""

    def __init__(self, trees):
        self.trees = list(trees)
        arr = (_P * len(self.trees))(*[t.context() for t in self.trees])
        h = _P()
        _check(lib().plk_batch_create(ctypes.byref(h), arr, len(self.trees)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().plk_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def set_launch(self, grid=0, use_graph=True):
        _check(lib().plk_batch_set_launch(self.h, grid, int(use_graph)))

    def _queue(self, branches, clear):
        n = len(self.trees)
        a, b, t = (_int * n)(), (_int * n)(), np.zeros(n)
        for i, (tr, (u, v)) in enumerate(zip(self.trees, branches)):
            a[i], b[i], t[i] = tr.queue_branch(u, v, clear)
        return a, b, t

    def loglik(self, branches, clear=False):
        a, b, t = self._queue(branches, clear)
        out = np.zeros(len(self.trees))
        tot = ctypes.c_double()
        _check(lib().plk_batch_loglik(self.h, a, b, _d(t), _d(out), ctypes.byref(tot)))
        return out, tot.value

    def autotune(self, branches, reps=3):
        a, b, t = self._queue(branches, True)
        g, ms = _int(), ctypes.c_double()
        _check(lib().plk_batch_autotune(self.h, a, b, _d(t), reps, ctypes.byref(g), ctypes.byref(ms)))
        # the tuned traversal stays queued: evaluate it once so the trees stay consistent
        out = np.zeros(len(self.trees))
        tot = ctypes.c_double()
        _check(lib().plk_batch_loglik(self.h, a, b, _d(t), _d(out), ctypes.byref(tot)))
        return g.value, ms.value

    def timing(self):
        ms, n, u = ctypes.c_double(), ctypes.c_int64(), ctypes.c_int64()
        _check(lib().plk_batch_timing(self.h, ctypes.byref(ms), ctypes.byref(n), ctypes.byref(u)))
        return ms.value, n.value, u.value

    def items(self):
        return lib().plk_batch_items(self.h)


REQ_LOGL, REQ_OPT_BRANCH, REQ_OPT_ALL = 0, 1, 2


class Request:
    """One serving request built from a Problem; keeps its arrays alive."""

    def __init__(self, prob, kind=REQ_LOGL, precision=FP64, scaling=NORM, opt_branch=None, iterations=1, tol=1e-3):
        self.p = prob
        r = plk_request()
        keep = []

        def arr(a, dt=np.float64):
            a = np.ascontiguousarray(a, dtype=dt).ravel()
            keep.append(a)
            return a

        r.nstates, r.ncat, r.ntaxa, r.nptn = prob.n, len(prob.rates), prob.ntaxa, prob.nptn
        r.precision, r.scaling = precision, scaling
        tip = arr(prob.tip_lh)
        r.state_unknown = tip.size // prob.n - 1
        r.eval, r.evec, r.inv_evec = _d(arr(prob.eval)), _d(arr(prob.evec)), _d(arr(prob.inv_evec))
        r.rates, r.props, r.tip_lh = _d(arr(prob.rates)), _d(arr(prob.props)), _d(tip)
        r.states = arr(prob.states, np.uint8).ctypes.data_as(_u8p)
        r.freq = _d(arr(prob.freq))
        r.invar = None
        r.nei = arr(prob.nei, np.int32).ctypes.data_as(_ip)
        r.len = _d(arr(prob.blen))
        r.kind = kind
        a, b = opt_branch if opt_branch else (0, int(prob.nei[0, 0]))
        r.opt_a, r.opt_b, r.iterations, r.tol = a, b, iterations, tol
        self.r = r
        self._keep = keep

    def result(self):
        r = self.r
        return {"status": r.status, "logl_before": r.logl_before, "logl_after": r.logl_after, "opt_len": r.opt_len,
                "latency": r.t_done - r.t_submit, "queue": r.t_admit - r.t_submit, "footprint": r.footprint}


class Server:
    def __init__(self, nworkers=4, hbm_budget=200 << 30):
        h = _P()
        _check(lib().plk_server_create(ctypes.byref(h), nworkers, hbm_budget))
        self.h = h

    def submit(self, req):
        _check(lib().plk_server_submit(self.h, ctypes.addressof(req.r)))

    def drain(self):
        _check(lib().plk_server_drain(self.h))

    def now(self):
        return lib().plk_server_now(self.h)

    def stats(self):
        st = np.zeros(8)
        _check(lib().plk_server_stats(self.h, _d(st)))
        keys = ["completed", "failed", "p50_s", "p99_s", "mean_s", "makespan_s", "peak_hbm_bytes", "peak_concurrent"]
        return dict(zip(keys, st.tolist()))

    def counters(self):
        """(shared traversal launches, request traversals run in them, pooled trees)"""
        out = np.zeros(3, dtype=np.int64)
        _check(lib().plk_server_counters(self.h, out.ctypes.data_as(_i64p)))
        return {"launches": int(out[0]), "batched": int(out[1]), "pooled_trees": int(out[2])}

    def close(self):
        if getattr(self, "h", None):
            lib().plk_server_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


class Rell:
    """UFBoot replicates resident on the device (plk_rell_*): counts is nsamples x nptn."""

    def __init__(self, counts):
        c = np.ascontiguousarray(counts, dtype=np.float32)
        self.ns, self.nptn = c.shape
        self.h = _P()
        _check(lib().plk_rell_create(ctypes.byref(self.h), self.ns, self.nptn,
                                     c.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))

    def eval(self, ctx):
        out = np.zeros(self.ns)
        _check(lib().plk_rell_eval(self.h, ctx, _d(out)))
        return out

    def save_tree(self, ctx, cur_logl, uniform, eps=0.5):
        u = np.ascontiguousarray(uniform, dtype=np.float64)
        imp = np.zeros(self.ns, dtype=np.int32)
        n = _int()
        _check(lib().plk_rell_save_tree(self.h, ctx, cur_logl, eps, _d(u), imp.ctypes.data_as(_ip), ctypes.byref(n)))
        return imp, n.value

    def state(self):
        lo, orig = np.zeros(self.ns), np.zeros(self.ns)
        cnt = np.zeros(self.ns, dtype=np.int32)
        _check(lib().plk_rell_get(self.h, _d(lo), cnt.ctypes.data_as(_ip), _d(orig)))
        return lo, cnt, orig

    def close(self):
        if self.h:
            lib().plk_rell_destroy(self.h)
            self.h = _P()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Alignment:
    """PHYLIP / FASTA alignment compressed to site patterns (plk_align_*).  fmt: 0 detect,
    1 PHYLIP, 2 PHYLIP sequential, 3 FASTA; seq_type None detects DNA / protein."""

    def __init__(self, path=None, text=None, fmt=0, seq_type=None):
        h = _P()
        st = -1 if seq_type is None else seq_type
        if text is not None:
            raw = text.encode() if isinstance(text, str) else bytes(text)
            _check(lib().plk_align_parse(ctypes.byref(h), raw, len(raw), fmt, st))
        else:
            _check(lib().plk_align_read(ctypes.byref(h), str(path).encode(), fmt, st))
        try:
            d = np.zeros(6, dtype=np.int32)
            _check(lib().plk_align_dims(h, d.ctypes.data_as(_ip)))
            self.ntaxa, self.nsites, self.nptn, self.n, self.seq_type, self.gaps_only = [int(x) for x in d]
            buf = ctypes.create_string_buffer(4096)
            self.names = []
            for i in range(self.ntaxa):
                _check(lib().plk_align_name(h, i, buf, len(buf)))
                self.names.append(buf.value.decode())
            self.states = np.zeros((self.ntaxa, self.nptn), dtype=np.uint8)
            self.freq = np.zeros(self.nptn)
            self.site_pattern = np.zeros(self.nsites, dtype=np.int32)
            self.flags = np.zeros(self.nptn, dtype=np.int32)
            _check(lib().plk_align_patterns(h, self.states.ctypes.data_as(_u8p), _d(self.freq),
                                            self.site_pattern.ctypes.data_as(_ip), self.flags.ctypes.data_as(_ip)))
        finally:
            lib().plk_align_destroy(h)
