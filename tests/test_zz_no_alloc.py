"""titan_chlld_batch_noalloc: the batch Riemann solve on caller-owned device arrays, with no
allocation inside the call.  Its outputs must equal titan_chlld_batch's bit for bit on two calls in
a row (the second with fewer states, so the scratch flags it shares across calls must be cleared),
and its error returns must match: a fatal branch (NaN state), a bad method code, an oversize
batch."""
import ctypes

import numpy as np
import pytest

from helpers import random_normals, random_states

pytestmark = pytest.mark.gpu

_HIP = None


def hip():
    """The HIP runtime libtitan_gpu.so has loaded (device buffers for the test come from the same
    runtime instance)."""
    global _HIP
    if _HIP is None:
        path = None
        with open("/proc/self/maps") as f:
            for line in f:
                if "libamdhip64.so" in line:
                    path = line.split()[-1]
                    break
        assert path, "libamdhip64 is not loaded"
        L = ctypes.CDLL(path, mode=ctypes.RTLD_GLOBAL)
        L.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        L.hipFree.argtypes = [ctypes.c_void_p]
        L.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        L.hipDeviceSynchronize.argtypes = []
        _HIP = L
    return _HIP


H2D, D2H = 1, 2


class Dev:
    """Device copies of host arrays; outputs copied back by name."""

    def __init__(self):
        self.bufs = {}

    def put(self, name, a):
        a = np.ascontiguousarray(a)
        p = ctypes.c_void_p()
        assert hip().hipMalloc(ctypes.byref(p), max(1, a.nbytes)) == 0
        assert hip().hipMemcpy(p, a.ctypes.data, a.nbytes, H2D) == 0
        self.bufs[name] = (p, a.shape, a.dtype)
        return p

    def get(self, name):
        p, shape, dtype = self.bufs[name]
        out = np.empty(shape, dtype)
        assert hip().hipMemcpy(out.ctypes.data, p, out.nbytes, D2H) == 0
        return out

    def free(self):
        for p, _, _ in self.bufs.values():
            hip().hipFree(p)
        self.bufs = {}


def batch(rng, n, K):
    qL, qR = random_states(rng, n), random_states(rng, n)
    return dict(n_state=rng.integers(0, 4, n).astype(np.int8), n_disc=rng.integers(0, 4, n).astype(np.int8),
                flags=rng.integers(0, 3, n).astype(np.uint8), normal=random_normals(rng, n),
                w=rng.normal(0, 1.0, n), qL=qL, qR=qR, kL=rng.uniform(0.1, 2.0, (n, K)),
                kR=rng.uniform(0.1, 2.0, (n, K)))


def run_noalloc(T, b, K, precision):
    T.lib()
    n = b["qL"].shape[0]
    d = Dev()
    try:
        ptr = {k: d.put(k, v) for k, v in b.items()}
        for name, shape, dt in (("q", (n, 8), np.float64), ("k", (n, max(K, 1)), np.float64),
                                ("ds", (n, 3), np.float64), ("branch", (n,), np.uint8)):
            ptr[name] = d.put(name, np.full(shape, 7, dt))
        rc = T.lib().titan_chlld_batch_noalloc(
            n, ptr["n_state"], ptr["n_disc"], ptr["flags"], ptr["normal"], ptr["w"], ptr["qL"], ptr["qR"], K,
            ptr["kL"] if K else None, ptr["kR"] if K else None, ptr["q"], ptr["k"] if K else None, ptr["ds"],
            ptr["branch"], precision, None)
        out = {k: d.get(k) for k in ("q", "k", "ds", "branch")}
    finally:
        d.free()
    return rc, out


@pytest.mark.parametrize("precision", [0, 1])
def test_noalloc_equals_host_entry_point(T, precision):
    rng = np.random.default_rng(11 + precision)
    K = 2
    for n in (5000, 777):   # the second call is smaller: the shared scratch must not leak state
        b = batch(rng, n, K)
        q, k, ds, br, rc_ref = T.chlld(b["n_state"], b["normal"], b["qL"], b["qR"], w=b["w"], n_disc=b["n_disc"],
                                       flags=b["flags"], kL=b["kL"], kR=b["kR"], precision=precision)
        rc, out = run_noalloc(T, b, K, precision)
        print("noalloc n=%d precision=%d rc=%d (host entry %d)" % (n, precision, rc, rc_ref), flush=True)
        assert rc == rc_ref == T.TITAN_OK
        np.testing.assert_array_equal(out["q"], q)
        np.testing.assert_array_equal(out["k"][:, :K], k)
        np.testing.assert_array_equal(out["ds"], ds)
        np.testing.assert_array_equal(out["branch"], br)


def test_noalloc_errors_and_cleared_flags(T):
    rng = np.random.default_rng(21)
    b = batch(rng, 300, 0)
    bad = dict(b, qL=b["qL"].copy(), n_state=b["n_state"].copy(), flags=b["flags"].copy())
    bad["qL"][17, 4] = np.nan   # HLLD without left_ydar finds no region: the reference's fatal branch
    bad["n_state"][17] = 3
    bad["flags"][17] = 0
    rc_ref = T.chlld(bad["n_state"], bad["normal"], bad["qL"], bad["qR"], w=bad["w"], n_disc=bad["n_disc"],
                     flags=bad["flags"])[4]
    rc, _ = run_noalloc(T, bad, 0, 0)
    assert rc == rc_ref == T.TITAN_E_RIEMANN
    rc, out = run_noalloc(T, b, 0, 0)            # the flag of the call before is cleared
    assert rc == T.TITAN_OK
    codes = dict(b, n_state=b["n_state"].copy())
    codes["n_state"][5] = 7                      # titan_chlld_batch rejects it on the host
    rc, _ = run_noalloc(T, codes, 0, 0)
    assert rc == T.TITAN_E_ARG
    rc, out2 = run_noalloc(T, b, 0, 0)
    assert rc == T.TITAN_OK
    np.testing.assert_array_equal(out["q"], out2["q"])
    # an oversize batch is refused before anything is launched (no array is touched)
    assert T.lib().titan_chlld_batch_noalloc((1 << 31) + 1, 8, None, None, 8, None, 8, 8, 0, None, None, 8, None,
                                             None, None, 0, None) == T.TITAN_E_ARG
