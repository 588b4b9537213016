"""plk_rell_eval_noalloc (RELL scoring whose staging buffer is a static device array, for
callers that cannot allocate) against plk_rell_eval through the C ABI, on two calls in a
row, the second with a smaller input.  The test allocates its own device buffers with
hipMalloc; the entry point itself allocates nothing."""
import ctypes
import sys

import numpy as np
import pytest

import gems_amd as G

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def hip():
    lib = ctypes.CDLL("/opt/rocm/lib/libamdhip64.so")
    lib.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    lib.hipFree.argtypes = [ctypes.c_void_p]
    lib.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    lib.hipDeviceSynchronize.argtypes = []
    return lib


H2D, D2H = 1, 2


def device_copy(h, arr):
    p = ctypes.c_void_p()
    assert h.hipMalloc(ctypes.byref(p), arr.nbytes) == 0
    assert h.hipMemcpy(p, arr.ctypes.data, arr.nbytes, H2D) == 0
    return p


def run_case(h, ntaxa, nsites, nsamples, seed):
    p = G.Problem(G.SEQ_DNA, ntaxa, nsites, seed=seed)
    g = G.Tree(p)
    g.loglik(0, int(p.nei[0, 0]))  # leaves the per-pattern logL on the device
    rng = np.random.default_rng(seed)
    counts = rng.multinomial(int(p.freq.sum()), p.freq / p.freq.sum(), size=nsamples).astype(np.float32)
    r = G.Rell(counts)
    want = r.eval(g.context())
    r.close()
    npad = (p.nptn + 63) // 64 * 64
    stride = npad + 4  # a row stride other than the padded pattern count
    host = np.zeros((nsamples, stride), dtype=np.float32)
    host[:, :p.nptn] = counts
    d_counts = device_copy(h, host)
    d_rell = device_copy(h, np.full(nsamples, np.nan))
    say("noalloc case", ntaxa, nsites, "patterns", p.nptn, "samples", nsamples)
    rc = G.lib().plk_rell_eval_noalloc(g.context(), nsamples, p.nptn, d_counts, stride, d_rell)
    assert rc == 0, G.lib().plk_last_error()
    got = np.zeros(nsamples)
    assert h.hipMemcpy(got.ctypes.data, d_rell, got.nbytes, D2H) == 0
    h.hipFree(d_counts)
    h.hipFree(d_rell)
    g.close()
    return got, want


def test_noalloc_rell_equals_rell_eval():
    h = hip()
    got1, want1 = run_case(h, 24, 6000, 300, 1)
    assert np.array_equal(got1, want1)
    got2, want2 = run_case(h, 10, 700, 40, 2)  # smaller, on the same static array
    assert np.array_equal(got2, want2)


def test_noalloc_rejects_sizes_beyond_scratch():
    p = G.Problem(G.SEQ_DNA, 8, 200, seed=3)
    g = G.Tree(p)
    g.loglik(0, int(p.nei[0, 0]))
    big = (1 << 20) + 1
    rc = G.lib().plk_rell_eval_noalloc(g.context(), 4, big, ctypes.c_void_p(16), big + 63, ctypes.c_void_p(16))
    assert rc != 0
    assert b"scratch maximum" in G.lib().plk_last_error()
    g.close()
