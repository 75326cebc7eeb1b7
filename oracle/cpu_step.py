"""ctypes binding of the CPU baseline step (oracle/cpu_step.c): the
reference algorithm for the SPC water MD step on the host's cores, f32 with
FP64 reductions, OpenMP.  Benchmark and test infrastructure only."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_lib = None
_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(_HERE, "_build", "libcpustep.so")
        if not os.path.exists(path):
            import subprocess
            subprocess.run(["make", "-s", "-C", _HERE], check=True)
        L = C.CDLL(path)
        L.cs_create.restype = C.c_void_p
        L.cs_create.argtypes = [C.c_int, _f32p, _f32p, _i32p, _f64p, _f32p, _f32p, C.c_double,
                                C.c_double, C.c_double, C.c_double, C.c_int, C.c_double,
                                C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
                                C.c_double, C.c_double]
        L.cs_step.argtypes = [C.c_void_p, C.c_int]
        L.cs_set_velocities.argtypes = [C.c_void_p, _f32p, _f32p, _f32p]
        L.cs_get.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.cs_free.argtypes = [C.c_void_p]
        _lib = L
    return _lib


class CpuStep:
    """One SPC water system (waters O, H, H consecutive) on the host."""

    def __init__(self, system, dt, r_skin, kt, tau, pressure=0.0, tau_p=0.0, kappa=0.0):
        lj = system.lj
        c6 = np.zeros(4, np.float32)
        c12 = np.zeros(4, np.float32)
        for a in range(2):
            for b in range(2):
                s6 = float(lj.sigma[a, b]) ** 6
                c6[2 * a + b] = 4.0 * float(lj.epsilon[a, b]) * s6
                c12[2 * a + b] = 4.0 * float(lj.epsilon[a, b]) * s6 * s6
        self.n = system.n
        self._h = lib().cs_create(
            system.n, np.ascontiguousarray(system.posq, np.float32).reshape(-1),
            np.ascontiguousarray(system.mass, np.float32),
            np.ascontiguousarray(system.types, np.int32),
            np.ascontiguousarray(system.lattice[:3], np.float64), c6, c12,
            float(system.spme.r_cut), float(r_skin), float(system.spme.alpha),
            float(system.spme.k_c), int(system.spme.grid[0]), float(system.water_geom[0]),
            float(system.water_geom[1]), float(dt), float(kt), float(tau), float(pressure),
            float(tau_p), float(kappa))

    def set_velocities(self, v):
        v = np.ascontiguousarray(v, np.float32)
        lib().cs_set_velocities(self._h, v[0].copy(), v[1].copy(), v[2].copy())

    def step(self, n):
        lib().cs_step(self._h, int(n))

    def state(self):
        posq = np.zeros((self.n, 4), np.float32)
        f = np.zeros((self.n, 3), np.float32)
        out = np.zeros(5)
        lib().cs_get(self._h, posq.ctypes.data, f.ctypes.data, out.ctypes.data)
        return posq, f, dict(zip(["e_pot", "virial", "ke", "rebuilds", "threads"], out.tolist()))

    def close(self):
        if self._h:
            lib().cs_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
