"""The allocation-free gradient entry point (hd_mlp_gradient_noalloc)
against hd_mlp_gradient, through the C ABI, on device buffers: two calls in
a row, the second with fewer rows, each bit-identical to the allocating
entry point on the same rows, and the row limit refused with an error."""
import ctypes as C
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _rows(rng, n, s):
    x = rng.normal(0, s, size=(n, 8)).astype(np.float32)
    x[:, 5:7] = rng.uniform(-0.5, 0.5, size=(n, 2))
    t = (x[:, 0] - x[:, 5] * x[:, 1] - x[:, 6] * x[:, 2]).astype(np.float32)
    return x, t


class _Dev:
    """Device buffers from the HIP runtime libheddle itself links (the
    caller's memory; torch bundles a second HIP runtime, which is not mixed
    into this process's HIP state here)."""

    def __init__(self):
        self.hip = C.CDLL("/opt/rocm/lib/libamdhip64.so.7")
        self.hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.hip.hipFree.argtypes = [C.c_void_p]
        self.bufs = []

    def alloc(self, nbytes):
        p = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(p), nbytes) == 0
        self.bufs.append(p)
        return p

    def put(self, a):
        # a fresh buffer holding the host array a
        a = np.ascontiguousarray(a)
        p = self.alloc(a.nbytes)
        assert self.hip.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0  # host to device
        return p

    def get(self, p, dtype, count):
        out = np.empty(count, dtype)
        assert self.hip.hipMemcpy(out.ctypes.data, p, out.nbytes, 2) == 0  # device to host
        return out

    def free(self):
        for p in self.bufs:
            self.hip.hipFree(p)
        self.bufs = []


def test_gradient_noalloc_matches_the_allocating_entry_point():
    from heddlemd_amd.engine import ShakePredictor

    say("case: hd_mlp_gradient_noalloc, 12000 then 3000 rows, vs hd_mlp_gradient")
    rng = np.random.default_rng(41)
    s = 0.01
    m = ShakePredictor(seed=9)
    p = rng.normal(0, 0.3, size=m.n_params).astype(np.float32)
    p[-1] = s
    m.params = p
    lib = m._lib
    d = _Dev()
    g_dev = d.alloc(4 * (m.n_params - 1))
    l_dev = d.alloc(8)
    for n in (12000, 3000):
        x, t = _rows(rng, n, s)
        loss_ref, g_ref = m.gradient(x, t)
        x_dev, t_dev = d.put(x), d.put(t)
        nan_g = np.full(m.n_params - 1, np.nan, np.float32)
        nan_l = np.full(1, np.nan, np.float64)
        assert d.hip.hipMemcpy(g_dev, nan_g.ctypes.data, nan_g.nbytes, 1) == 0
        assert d.hip.hipMemcpy(l_dev, nan_l.ctypes.data, nan_l.nbytes, 1) == 0
        rc = lib.hd_mlp_gradient_noalloc(m._h, x_dev, t_dev, n, g_dev, l_dev)
        assert rc == 0, lib.hd_mlp_last_error(m._h)
        g = d.get(g_dev, np.float32, m.n_params - 1)
        loss = float(d.get(l_dev, np.float64, 1)[0]) / n
        say(f"  {n} rows: loss {loss:.9e} vs {loss_ref:.9e}, "
            f"gradient equal {np.array_equal(g, g_ref)}")
        assert np.array_equal(g, g_ref)
        assert loss == loss_ref
    # beyond the scratch buffer's row limit: a clear error, nothing launched
    big = d.alloc(16385 * 8 * 4)
    rc = lib.hd_mlp_gradient_noalloc(m._h, big, big, 16385, g_dev, l_dev)
    assert rc != 0
    assert b"16384" in lib.hd_mlp_last_error(m._h)
    d.free()
    m.close()
