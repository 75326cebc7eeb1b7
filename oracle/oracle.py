"""Python face of the CPU oracle (oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke()``
and by the ``cpu_baseline`` leg of ``bench.py``.  The product path in
``heddlemd_amd`` never imports this module.

Every function mirrors a reference routine; the C file cites file:line for
each.  ``md_step`` at the bottom composes them in the order the reference
runner executes one velocity-Verlet + SETTLE + Nose-Hoover-chain step
(src/runner.rs:2585-2687, src/integrator/mod.rs:344-429).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None

_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
_i64 = C.c_int64


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        L.or_pair_set.restype = _i64
        L.or_pair_set.argtypes = [C.c_int, _f32p, _f64p, C.c_double, C.c_double, _i64,
                                  _u32p, _i64, _u32p, C.POINTER(_i64)]
        L.or_pair_forces.restype = None
        L.or_pair_forces.argtypes = [C.c_int, _f32p, _u32p, C.c_int, _f64p, _f64p, _f64p,
                                     _f64p, C.c_double, C.c_double, C.c_double,
                                     C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, _f64p,
                                     C.c_double, _f64p, _f64p, _f64p, _u8p]
        L.or_bond_forces.restype = None
        L.or_bond_forces.argtypes = [C.c_int, _u32p, C.POINTER(C.c_int), _f64p, _f32p, _f64p,
                                     _f64p, _f64p, _f64p]
        L.or_dihedral_forces.restype = None
        L.or_dihedral_forces.argtypes = [C.c_int, _u32p, _f64p, _f32p, _f64p, _f64p, _f64p,
                                         _f64p]
        L.or_bspline.restype = C.c_double
        L.or_bspline.argtypes = [C.c_int, C.c_double]
        L.or_b_factors.argtypes = [C.c_int, C.c_int, _f64p]
        L.or_fft3d.argtypes = [_f64p, _f64p, C.c_int, C.c_int, C.c_int, C.c_int]
        L.or_spme_spread.argtypes = [C.c_int, _f32p, _f64p, C.c_int, C.c_int, C.c_int,
                                     C.c_int, _f64p]
        L.or_spme_recip.argtypes = [C.c_int, _f32p, _f64p, C.c_int, C.c_int, C.c_int,
                                    C.c_int, C.c_double, C.c_double, _f64p, _f64p, _f64p,
                                    C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_void_p]
        L.or_angle_forces.argtypes = [C.c_int, _u32p, _f64p, _f64p, _f32p, _f64p, _f64p,
                                      _f64p, _f64p]
        L.or_shake_positions_f32.argtypes = [C.c_int, _u32p, _u32p, _u32p, _u8p, _u8p, _f32p,
                                             _f32p, _f32p, _f32p, _f32p, _f32p, _f32p, _f32p,
                                             C.c_float, C.c_float, C.c_int, _f32p, _i32p,
                                             C.c_void_p, C.c_void_p]
        L.or_shake_features_f32.argtypes = [C.c_int, _u32p, _u32p, _u32p, _u8p, _u8p, _f32p,
                                            _f32p, _f32p, _f32p, _f32p, C.c_void_p, _f32p]
        L.or_rattle_velocities_f32.argtypes = [C.c_int, _u32p, _u32p, _u32p, _u8p, _u8p,
                                               _f32p, _f32p, _f32p, _f32p, _f32p, _f32p,
                                               C.c_float, C.c_float, C.c_int, _f32p, _i32p]
        L.or_philox4x32_10.argtypes = [C.c_uint32, C.c_uint32, _u32p, _u32p]
        L.or_lan_alpha.restype = C.c_float
        L.or_lan_alpha.argtypes = [C.c_float, C.c_float]
        L.or_philox_gaussian.restype = C.c_double
        L.or_philox_gaussian.argtypes = [C.c_uint32] * 6
        L.or_lan_kick_drift_half_f32.argtypes = [C.c_int, _f32p, _i32p, _f32p, _f32p, _f32p,
                                                 _f32p, _f32p, _f32p, _f32p, _f32p, C.c_float]
        L.or_lan_ou_drift_half_f32.argtypes = [C.c_int, _f32p, _i32p, _f32p, _f32p, _f32p,
                                               _f32p, _f32p, C.c_float, C.c_float, C.c_float,
                                               C.c_uint64, C.c_uint64]
        L.or_vv_kick_drift_f32.argtypes = [C.c_int, _f32p, _i32p, _f32p, _f32p, _f32p,
                                           _f32p, _f32p, _f32p, _f32p, _f32p, C.c_float]
        L.or_vv_kick_f32.argtypes = [C.c_int, _f32p, _f32p, _f32p, _f32p, _f32p, _f32p,
                                     _f32p, C.c_float]
        L.or_vv_kick_drift_lossless.argtypes = [C.c_int, _f32p, _i32p, _f32p, _f32p, _f32p,
                                                _f64p, _f64p, _f64p, _f64p, _f64p, _f64p,
                                                _f32p, _f32p, _f32p, _f32p, _f32p, C.c_float]
        L.or_vv_kick_lossless.argtypes = [C.c_int, _f32p, _f32p, _f32p, _f64p, _f64p, _f64p,
                                          _f32p, _f32p, _f32p, _f32p, C.c_float]
        L.or_settle_positions.argtypes = [C.c_int, _u32p, _f64p, _f64p, _f64p, C.c_double,
                                          C.c_double, C.c_double, C.c_double, C.c_double,
                                          _f64p, C.c_double, _f64p]
        L.or_settle_velocities.argtypes = [C.c_int, _u32p, _f64p, _f64p, C.c_double,
                                           C.c_double, _f64p, C.c_double, _f64p]
        L.or_nhc_half_step.restype = C.c_double
        L.or_nhc_half_step.argtypes = [C.c_int, _f64p, _f64p, _f64p, C.c_int, _f64p, C.c_int,
                                       C.c_double, C.c_double, C.c_double, C.c_double,
                                       C.POINTER(C.c_double)]
        L.or_csvr_factor.restype = C.c_double
        L.or_csvr_factor.argtypes = [C.c_double, C.c_uint32, C.c_uint64, C.c_uint64, C.c_double,
                                     C.c_double, C.POINTER(C.c_double)]
        L.or_berendsen_factor.restype = C.c_double
        L.or_berendsen_factor.argtypes = [C.c_double, C.c_double, C.c_double,
                                          C.POINTER(C.c_double)]
        L.or_andersen.restype = None
        L.or_andersen.argtypes = [C.c_int, _f32p, _f32p, _f32p, _f32p, C.c_uint64, C.c_uint64,
                                  C.c_float, C.c_float]
        L.or_kinetic_energy.restype = C.c_double
        L.or_kinetic_energy.argtypes = [C.c_int, _f32p, _f32p, _f32p, _f32p]
        L.or_num_threads.restype = C.c_int
        L.or_set_erfc_mode.argtypes = [C.c_int]
        L.or_set_shake_form.argtypes = [C.c_int]
        _lib = L
    return _lib


def _lat(lattice):
    return np.ascontiguousarray(lattice, dtype=np.float64)


def pair_set(posq, lattice, r_search, band=4.0 * 2.0 ** -23):
    """Canonical set of pairs (i<j) with r^2 <= r_search^2, plus the pairs in
    the rounding band around the threshold (neighbor_list.rs:1366-1506)."""
    posq = np.ascontiguousarray(posq, np.float32)
    n = posq.shape[0]
    cap = 1 << 20
    bcap = 1 << 16
    while True:
        out = np.zeros(2 * cap, np.uint32)
        bout = np.zeros(2 * bcap, np.uint32)
        nb = _i64(0)
        cnt = lib().or_pair_set(n, posq.reshape(-1), _lat(lattice), float(r_search) ** 2,
                                band, cap, out, bcap, bout, C.byref(nb))
        if cnt <= cap and nb.value <= bcap:
            return out[:2 * cnt].reshape(-1, 2), bout[:2 * nb.value].reshape(-1, 2)
        cap = max(cap, int(cnt) + 1)
        bcap = max(bcap, int(nb.value) + 1)


def pair_forces(posq, types, lj, lattice, coul=None, excl=None, band=0.0):
    """LJ + real-space Ewald pair forces in double.

    lj: dict with ntypes and (ntypes x ntypes) arrays sigma, epsilon, cutoff,
        r_switch (atomic units).
    coul: None or dict(alpha=, r_cut=, k_c=1.0).
    excl: None or (offsets u32[n+1], partners u32[], scales f64[]) with one
        scale for both terms, or (offsets, partners, scale_lj, scale_coul).
    Returns F (n,3), E (n,), W (n,), band_flag (n,) bool.
    """
    posq = np.ascontiguousarray(posq, np.float32)
    n = posq.shape[0]
    nt = int(lj["ntypes"])
    arr = lambda k: np.ascontiguousarray(np.asarray(lj[k], np.float64).reshape(-1))
    F = np.zeros(3 * n)
    E = np.zeros(n)
    W = np.zeros(n)
    bf = np.zeros(max(n, 1), np.uint8)
    alpha = rc = 0.0
    kc = 1.0
    if coul is not None:
        alpha, rc, kc = float(coul["alpha"]), float(coul["r_cut"]), float(coul.get("k_c", 1.0))
    keep = []
    if excl is not None:
        off = np.ascontiguousarray(excl[0], np.uint32)
        par = np.ascontiguousarray(excl[1], np.uint32)
        sca = np.ascontiguousarray(excl[2], np.float64)
        scc = np.ascontiguousarray(excl[3] if len(excl) > 3 else excl[2], np.float64)
        keep = [off, par, sca, scc]
        ptrs = [a.ctypes.data for a in keep]
    else:
        ptrs = [None, None, None, None]
    lib().or_pair_forces(n, posq.reshape(-1), np.ascontiguousarray(types, np.uint32), nt,
                         arr("sigma"), arr("epsilon"), arr("cutoff"), arr("r_switch"),
                         kc, alpha, rc, ptrs[0], ptrs[1], ptrs[2], ptrs[3], _lat(lattice), band,
                         F, E, W, bf)
    del keep
    return F.reshape(n, 3), E, W, bf[:n].astype(bool)


def set_erfc_mode(hastings: bool):
    """Select the real-space erfc: exact (double erfc, the reference's f64
    build) or the Hastings polynomial of its f32 build."""
    lib().or_set_erfc_mode(1 if hastings else 0)


def bspline(p, x):
    return lib().or_bspline(int(p), float(x))


def b_factors(n, p):
    out = np.zeros(n)
    lib().or_b_factors(n, p, out)
    return out


def fft3d(re, im, sign):
    re = np.ascontiguousarray(re, np.float64).copy()
    im = np.ascontiguousarray(im, np.float64).copy()
    na, nb, nc = re.shape
    lib().or_fft3d(re.reshape(-1), im.reshape(-1), na, nb, nc, sign)
    return re, im


def spme_spread(posq, lattice, grid, order):
    posq = np.ascontiguousarray(posq, np.float32)
    na, nb, nc = grid
    rho = np.zeros(na * nb * nc)
    lib().or_spme_spread(posq.shape[0], posq.reshape(-1), _lat(lattice), na, nb, nc, order, rho)
    return rho.reshape(na, nb, nc)


def spme_recip(posq, lattice, grid, order, alpha, k_c=1.0, want_v=False):
    """Reciprocal SPME in double.  Returns F (n,3), E (n,), W (n,), e_recip,
    w_recip and (optionally) the potential grid V."""
    posq = np.ascontiguousarray(posq, np.float32)
    n = posq.shape[0]
    na, nb, nc = grid
    F = np.zeros(3 * n)
    E = np.zeros(n)
    W = np.zeros(n)
    er = C.c_double(0)
    wr = C.c_double(0)
    V = np.zeros(na * nb * nc) if want_v else None
    lib().or_spme_recip(n, posq.reshape(-1), _lat(lattice), na, nb, nc, order, float(alpha),
                        float(k_c), F, E, W, C.byref(er), C.byref(wr),
                        V.ctypes.data if want_v else None)
    out = (F.reshape(n, 3), E, W, er.value, wr.value)
    if want_v:
        out = out + (V.reshape(na, nb, nc),)
    return out


def vv_kick_drift_f32(posq, img, v, f, mass, lattice, dt):
    """In-place f32 kick+drift (integrate.cu:24-140 with LOSSLESS=false)."""
    n = posq.shape[0]
    lib().or_vv_kick_drift_f32(n, posq.reshape(-1), img.reshape(-1), v[0], v[1], v[2],
                               f[0], f[1], f[2], mass,
                               np.ascontiguousarray(lattice, np.float32), np.float32(dt))


def vv_kick_f32(v, f, mass, dt):
    n = v.shape[1]
    lib().or_vv_kick_f32(n, v[0], v[1], v[2], f[0], f[1], f[2], mass, np.float32(dt))


def angle_forces(posq, lattice, ijk, k_theta, theta0):
    """Harmonic angle forces (n,3), per-atom energy and virial thirds, FP64
    (src/forces/angle.rs:258-313)."""
    n = posq.shape[0]
    ijk = np.ascontiguousarray(ijk, np.uint32).reshape(-1)
    F = np.zeros(3 * n)
    E = np.zeros(n)
    W = np.zeros(n)
    lib().or_angle_forces(ijk.size // 3, ijk, np.ascontiguousarray(k_theta, np.float64),
                          np.ascontiguousarray(theta0, np.float64), np.ascontiguousarray(posq),
                          np.ascontiguousarray(lattice, np.float64), F, E, W)
    return F.reshape(n, 3), E, W


def bond_forces(posq, lattice, ij, kind, prm):
    """Harmonic (kind 0: k, r0) and Morse (kind 1: De, a, re) bond forces
    (n,3), per-atom energy and virial halves, FP64
    (src/forces/harmonic_bond.rs:288-330, src/forces/morse.rs:290-320)."""
    n = posq.shape[0]
    ij = np.ascontiguousarray(ij, np.uint32).reshape(-1)
    kind = np.ascontiguousarray(kind, np.int32).reshape(-1)
    F = np.zeros(3 * n)
    E = np.zeros(n)
    W = np.zeros(n)
    lib().or_bond_forces(kind.size, ij, kind.ctypes.data_as(C.POINTER(C.c_int)),
                         np.ascontiguousarray(prm, np.float64).reshape(-1),
                         np.ascontiguousarray(posq, np.float32).reshape(-1),
                         np.ascontiguousarray(lattice, np.float64), F, E, W)
    return F.reshape(n, 3), E, W


def dihedral_forces(posq, lattice, ijkl, prm):
    """Periodic dihedral forces (n,3), per-atom energy and virial quarters,
    FP64 (src/forces/dihedral.rs:360-460); prm rows (k, n, phi0)."""
    n = posq.shape[0]
    ijkl = np.ascontiguousarray(ijkl, np.uint32).reshape(-1)
    F = np.zeros(3 * n)
    E = np.zeros(n)
    W = np.zeros(n)
    lib().or_dihedral_forces(ijkl.size // 4, ijkl,
                             np.ascontiguousarray(prm, np.float64).reshape(-1),
                             np.ascontiguousarray(posq, np.float32).reshape(-1),
                             np.ascontiguousarray(lattice, np.float64), F, E, W)
    return F.reshape(n, 3), E, W


SHAKE_TOL = 3.57e-6     # |r^2 - d^2| in Bohr^2 (shake.cu:251-256)
RATTLE_TOL = 8.63e-17   # |v_rel . r| (shake.cu:446)
SHAKE_MAX_ITER = 32


class ShakeGroups:
    """Constraint groups in the reference's flattened layout
    (src/integrator/shake.rs:250-345): per group a run of atom ids and a
    run of constraints (group-local i, j, r2 = (f32 d)^2)."""

    def __init__(self, groups):
        # groups: list of (atom_ids, [(local_i, local_j, d), ...])
        goff, gat, coff, ci, cj, r2 = [0], [], [0], [], [], []
        for atoms, cons in groups:
            gat.extend(int(a) for a in atoms)
            goff.append(len(gat))
            for a, b, d in cons:
                ci.append(a)
                cj.append(b)
                df = np.float32(d)
                r2.append(df * df)
            coff.append(len(ci))
        self.goff = np.asarray(goff, np.uint32)
        self.atoms = np.asarray(gat, np.uint32)
        self.coff = np.asarray(coff, np.uint32)
        self.ci = np.asarray(ci, np.uint8)
        self.cj = np.asarray(cj, np.uint8)
        self.r2 = np.asarray(r2, np.float32)
        self.n = len(groups)


def set_shake_form(wave_all: bool):
    """Group solve form: False = the reference's sequential form for groups
    of up to 8 atoms (the device's wave form beyond); True = the device's
    wave form (group frame, colour-ordered sweeps) for every group."""
    lib().or_set_shake_form(1 if wave_all else 0)


def shake_positions(G, posq, v, snap, mass, lattice, dt, tol=SHAKE_TOL, max_iter=SHAKE_MAX_ITER,
                    nu0=None, nu_out=None):
    """In-place SHAKE (shake.cu shake_positions), f32; returns (w_cons
    per atom, sweeps per group).  nu0 (one per constraint) warm-starts the
    solve with multipliers nu = lambda (1/m_i + 1/m_j); nu_out receives
    the total nu of each constraint.  The form follows set_shake_form."""
    w = np.zeros(posq.shape[0], np.float32)
    iters = np.zeros(max(G.n, 1), np.int32)
    p0 = None
    if nu0 is not None:
        nu0 = np.ascontiguousarray(nu0, np.float32)
        p0 = nu0.ctypes.data
    p1 = None
    if nu_out is not None:
        assert nu_out.dtype == np.float32 and nu_out.flags.c_contiguous
        p1 = nu_out.ctypes.data
    lib().or_shake_positions_f32(G.n, G.goff, G.atoms, G.coff, G.ci, G.cj, G.r2,
                                 posq.reshape(-1), v[0], v[1], v[2], snap.reshape(-1), mass,
                                 np.ascontiguousarray(lattice, np.float32), np.float32(dt),
                                 np.float32(tol), int(max_iter), w, iters, p0, p1)
    return w, iters[:G.n]


def shake_features(G, posq, snap, mass, lattice, nu_prev=None):
    """Predictor inputs per constraint (8 floats per row), f32 in the
    device's operation order (no reference counterpart); nu_prev = the
    previous solve's converged multipliers (zeros when None)."""
    out = np.zeros((len(G.r2), 8), np.float32)
    pp = None
    if nu_prev is not None:
        nu_prev = np.ascontiguousarray(nu_prev, np.float32)
        pp = nu_prev.ctypes.data
    lib().or_shake_features_f32(G.n, G.goff, G.atoms, G.coff, G.ci, G.cj, G.r2,
                                np.ascontiguousarray(posq, np.float32).reshape(-1),
                                np.ascontiguousarray(snap, np.float32).reshape(-1),
                                np.ascontiguousarray(mass, np.float32),
                                np.ascontiguousarray(lattice, np.float32), pp, out.reshape(-1))
    return out


def rattle_velocities(G, posq, v, mass, lattice, dt, w_cons=None, tol=RATTLE_TOL,
                      max_iter=SHAKE_MAX_ITER):
    """In-place RATTLE (shake.cu rattle_velocities), f32; adds the
    velocity-level virial to w_cons when dt > 0; returns sweeps per group."""
    if w_cons is None:
        w_cons = np.zeros(posq.shape[0], np.float32)
    iters = np.zeros(max(G.n, 1), np.int32)
    lib().or_rattle_velocities_f32(G.n, G.goff, G.atoms, G.coff, G.ci, G.cj,
                                   posq.reshape(-1), v[0], v[1], v[2], mass,
                                   np.ascontiguousarray(lattice, np.float32), np.float32(dt),
                                   np.float32(tol), int(max_iter), w_cons, iters)
    return iters[:G.n]


def philox4x32_10(key, ctr):
    out = np.zeros(4, np.uint32)
    lib().or_philox4x32_10(int(key[0]), int(key[1]), np.asarray(ctr, np.uint32), out)
    return [int(x) for x in out]


def philox_gaussian(seed, counter, pid, axis):
    """Standard normal draw of (seed, draw counter, particle, axis)
    (philox.cuh philox_gaussian), in double."""
    return lib().or_philox_gaussian(seed & 0xFFFFFFFF, seed >> 32, counter & 0xFFFFFFFF,
                                    counter >> 32, pid, axis)


def lan_kick_drift_half(posq, img, v, f, mass, lattice, dt):
    """Langevin B + A (integrate.cu vv_kick, langevin.cu lan_drift_half), f32."""
    n = posq.shape[0]
    lib().or_lan_kick_drift_half_f32(n, posq.reshape(-1), img.reshape(-1), v[0], v[1], v[2],
                                     f[0], f[1], f[2], mass,
                                     np.ascontiguousarray(lattice, np.float32), np.float32(dt))


def lan_ou_drift_half(posq, img, v, mass, lattice, dt, alpha, kt, seed, counter):
    """Langevin O + A (langevin.cu lan_ou_step, lan_drift_half), f32."""
    n = posq.shape[0]
    lib().or_lan_ou_drift_half_f32(n, posq.reshape(-1), img.reshape(-1), v[0], v[1], v[2],
                                   mass, np.ascontiguousarray(lattice, np.float32),
                                   np.float32(dt), np.float32(alpha), np.float32(kt), seed,
                                   counter)


def lan_alpha(friction, dt):
    """exp(-gamma dt) in f32, as the reference computes it on the host
    (langevin_baoab.rs execute, Custom "O")."""
    return np.float32(lib().or_lan_alpha(np.float32(friction), np.float32(dt)))


def vv_kick_drift_lossless(posq, img, v, xlo, vlo, f, mass, lattice, dt):
    n = posq.shape[0]
    lib().or_vv_kick_drift_lossless(n, posq.reshape(-1), img.reshape(-1), v[0], v[1], v[2],
                                    xlo[0], xlo[1], xlo[2], vlo[0], vlo[1], vlo[2],
                                    f[0], f[1], f[2], mass,
                                    np.ascontiguousarray(lattice, np.float32), np.float32(dt))


def vv_kick_lossless(v, vlo, f, mass, dt):
    n = v.shape[1]
    lib().or_vv_kick_lossless(n, v[0], v[1], v[2], vlo[0], vlo[1], vlo[2], f[0], f[1], f[2],
                              mass, np.float32(dt))


def settle_geometry(d_oh, d_hh, m_o, m_h):
    """Canonical (ra, rb, rc) (src/integrator/settle.rs:83-91)."""
    total = m_o + 2.0 * m_h
    h = np.sqrt(d_oh * d_oh - (d_hh * 0.5) ** 2)
    return (2.0 * m_h / total) * h, (m_o / total) * h, d_hh * 0.5


def settle_positions(groups, pos, vel, snap, geom, m_o, m_h, lattice, dt):
    """In-place double SETTLE position reset; returns the per-slot virial."""
    g = np.ascontiguousarray(groups, np.uint32)
    vir = np.zeros(3 * g.shape[0])
    ra, rb, rc = geom
    lib().or_settle_positions(g.shape[0], g.reshape(-1), pos.reshape(-1), vel.reshape(-1),
                              np.ascontiguousarray(snap, np.float64).reshape(-1),
                              ra, rb, rc, m_o, m_h, _lat(lattice), dt, vir)
    return vir


def settle_velocities(groups, pos, vel, m_o, m_h, lattice, dt, virial=None):
    g = np.ascontiguousarray(groups, np.uint32)
    vir = np.zeros(3 * g.shape[0]) if virial is None else virial
    lib().or_settle_velocities(g.shape[0], g.reshape(-1),
                               np.ascontiguousarray(pos, np.float64).reshape(-1),
                               vel.reshape(-1), m_o, m_h, _lat(lattice), dt, vir)
    return vir


YOSHIDA = {
    1: [1.0],
    3: [1.3512071919596577, -1.7024143839193155, 1.3512071919596577],
    5: [0.41449077179437574, 0.41449077179437574, -0.6579630871775030,
        0.41449077179437574, 0.41449077179437574],
    7: [0.7845136104775573, 0.2355732133593582, -1.1776799841788710, 1.3151863206839023,
        -1.1776799841788710, 0.2355732133593582, 0.7845136104775573],
}


class NHCState:
    """Host Nose-Hoover chain (src/integrator/nose_hoover_chain.rs:182-279)."""

    def __init__(self, n_atoms, n_constraints, kt, tau, chain=3, yoshida=3, n_resp=1):
        self.g_dof = float(max(3 * n_atoms - n_constraints - 3, 0))
        self.kt = float(kt)
        t2 = tau * tau
        self.q = np.array([self.g_dof * kt * t2] + [kt * t2] * (chain - 1), np.float64)
        self.xi = np.zeros(chain)
        self.pxi = np.zeros(chain)
        self.w = np.array(YOSHIDA[yoshida], np.float64)
        self.n_resp = n_resp

    def half_step(self, dt, k):
        kout = C.c_double(0)
        f = lib().or_nhc_half_step(len(self.q), self.xi, self.pxi, self.q, len(self.w), self.w,
                                   self.n_resp, float(dt), float(k), self.g_dof, self.kt,
                                   C.byref(kout))
        return f, kout.value

    def conserved_extra(self):
        """Chain term of the nhc_conserved log column
        (nose_hoover_chain.rs:365-381)."""
        c = float(np.sum(self.pxi ** 2 / (2.0 * self.q)))
        if len(self.xi):
            c += self.g_dof * self.kt * self.xi[0] + self.kt * float(np.sum(self.xi[1:]))
        return c


class MtkNpt:
    """MTK-NPT integrator in FP64 (src/integrator/mtk_npt.rs:124-470): the
    symmetric Trotter splitting cell chain / particle chain / p_eps kick /
    cell-coupled velocity half kick / cell-coupled drift with the box scaled
    by exp(p_eps dt / W), the force evaluation, then the mirror image.  The
    particle chain is the oracle's Nose-Hoover chain (NHCState, pinned
    separately) applied with its cumulative factor; the cell chain is the
    same sub-step on the one cell degree of freedom (K = p_eps^2 / W)."""

    def __init__(self, n_atoms, n_constraints, kt, pressure, tau_t, tau_p, chain=3, yoshida=3,
                 n_resp=1):
        self.part = NHCState(n_atoms, n_constraints, kt, tau_t, chain, yoshida, n_resp)
        self.g = max(float(3 * n_atoms - n_constraints - 3), 1.0)
        self.part.g_dof = self.g
        t2 = tau_t * tau_t
        self.part.q = np.array([self.g * kt * t2] + [kt * t2] * (chain - 1), np.float64)
        self.cell = NHCState(1, 0, kt, tau_t, chain, yoshida, n_resp)
        self.cell.g_dof = 1.0
        self.cell.q = np.full(chain, kt * t2)
        self.kt = float(kt)
        self.p0 = float(pressure)
        self.w = (self.g + 3.0) * kt * tau_p * tau_p
        self.p_eps = 0.0
        self.last_p = 0.0
        self.last_v = 0.0

    @staticmethod
    def _shx(a):
        return 1.0 + a * a / 6.0 if abs(a) < 1e-6 else np.sinh(a) / a

    def _cell_chain(self, dt):
        k2 = self.p_eps * self.p_eps / self.w
        f, _ = self.cell.half_step(dt, 0.5 * k2)
        self.p_eps *= f

    def _kick(self, v, f, m, dt):
        a = (1.0 + 3.0 / self.g) * (self.p_eps / self.w)
        e = np.exp(-a * dt / 2.0)
        phi = 0.5 * dt * self._shx(a * dt / 4.0) * np.exp(-a * dt / 4.0)
        return e * v + phi * (f / m)

    def step(self, x, v, m, L, forces, dt, f, w):
        """One step from positions x (n, 3), velocities v (n, 3), masses m
        (n, 1), box lengths L (3,), the forces f and virial w at x;
        forces(x, L) -> (F, W).  Returns x, v, L, f, w after the step."""
        ke = 0.5 * float(np.sum(m * v * v))
        vol = float(np.prod(L))
        pres = (2.0 * ke + w) / (3.0 * vol)
        self._cell_chain(dt)
        fac, ke = self.part.half_step(dt, ke)
        v = v * fac
        self.p_eps += 0.5 * dt * (3.0 * vol * (pres - self.p0) + 6.0 / self.g * ke)
        v = self._kick(v, f, m, dt)
        b = self.p_eps / self.w
        eb = np.exp(b * dt)
        x = eb * x + dt * self._shx(b * dt / 2.0) * np.exp(b * dt / 2.0) * v
        L = L * eb
        f, w = forces(x, L)
        ke = 0.5 * float(np.sum(m * v * v))
        vol = float(np.prod(L))
        pres = (2.0 * ke + w) / (3.0 * vol)
        v = self._kick(v, f, m, dt)
        ke = 0.5 * float(np.sum(m * v * v))
        self.p_eps += 0.5 * dt * (3.0 * vol * (pres - self.p0) + 6.0 / self.g * ke)
        fac, ke = self.part.half_step(dt, ke)
        v = v * fac
        self.last_p, self.last_v = pres, vol
        self._cell_chain(dt)
        return x, v, L, f, w

    def conserved_extra(self):
        """H - K - U (mtk_npt.rs:246-268)."""
        return (self.p0 * self.last_v + 0.5 * self.p_eps ** 2 / self.w
                + self.part.conserved_extra() + float(np.sum(self.cell.pxi ** 2 / (2 * self.cell.q)))
                + self.kt * float(np.sum(self.cell.xi)))


def csvr_factor(k, g_dof, seed, counter, c, kt_per_dof):
    """CSVR rescale (src/integrator/csvr.rs:188-240): (factor, K_new)."""
    kn = C.c_double(0.0)
    f = lib().or_csvr_factor(float(k), int(g_dof), int(seed), int(counter), float(c),
                             float(kt_per_dof), C.byref(kn))
    return f, kn.value


def berendsen_factor(k, k_target, dt_over_tau):
    """Berendsen thermostat (src/integrator/berendsen.rs:139-170):
    (factor, injected energy)."""
    inj = C.c_double(0.0)
    f = lib().or_berendsen_factor(float(k), float(k_target), float(dt_over_tau), C.byref(inj))
    return f, inj.value


def andersen(v, mass, seed, counter, p, kt):
    """Andersen collisions in place on v (3, n) f32 (kernels/andersen.cu)."""
    lib().or_andersen(v.shape[1], v[0], v[1], v[2], np.ascontiguousarray(mass, np.float32),
                      int(seed), int(counter), float(p), float(kt))


def kinetic_energy(v, mass):
    return lib().or_kinetic_energy(v.shape[1], v[0], v[1], v[2], mass)


def num_threads():
    return lib().or_num_threads()


# ---------------------------------------------------------------------
# SHAKE multiplier predictor (no reference counterpart): the 8 -> 64 ->
# 64 -> 1 tanh network in FP64, forward and analytic backward.
# Parameter pack: W1[64][8], b1[64], W2[64][64], b2[64], w3[64], b3, s.
MLP_IN, MLP_H = 8, 64
MLP_NTRAIN = MLP_H * MLP_IN + MLP_H + MLP_H * MLP_H + MLP_H + MLP_H + 1
MLP_NPARAM = MLP_NTRAIN + 1


def mlp_unpack(p):
    p = np.asarray(p, np.float64)
    o = 0
    W1 = p[o:o + MLP_H * MLP_IN].reshape(MLP_H, MLP_IN); o += MLP_H * MLP_IN
    b1 = p[o:o + MLP_H]; o += MLP_H
    W2 = p[o:o + MLP_H * MLP_H].reshape(MLP_H, MLP_H); o += MLP_H * MLP_H
    b2 = p[o:o + MLP_H]; o += MLP_H
    w3 = p[o:o + MLP_H]; o += MLP_H
    b3 = p[o]
    s = p[o + 1]
    return W1, b1, W2, b2, w3, b3, s


def mlp_inputs(x, s):
    """Scaled inputs: the nu1 columns (0-4) and rho (7) divided by s, the
    couplings (5, 6) as they are."""
    x = np.asarray(x, np.float64).reshape(-1, MLP_IN)
    si = 1.0 / s if s > 0 else 0.0
    u = x * si
    u[:, 5:7] = x[:, 5:7]
    return u


def mlp_forward(p, x):
    """Predicted nu = x0 + s o(x) per row."""
    W1, b1, W2, b2, w3, b3, s = mlp_unpack(p)
    x = np.asarray(x, np.float64).reshape(-1, MLP_IN)
    u = mlp_inputs(x, s)
    h1 = np.tanh(u @ W1.T + b1)
    h2 = np.tanh(h1 @ W2.T + b2)
    o = h2 @ w3 + b3
    return x[:, 0] + s * o


def mlp_loss_grad(p, x, t):
    """Mean of (o - (t - x0)/s)^2 over the rows and its gradient with
    respect to the trained parameters (the pack without s)."""
    W1, b1, W2, b2, w3, b3, s = mlp_unpack(p)
    x = np.asarray(x, np.float64).reshape(-1, MLP_IN)
    t = np.asarray(t, np.float64)
    n = x.shape[0]
    si = 1.0 / s if s > 0 else 0.0
    u = mlp_inputs(x, s)
    h1 = np.tanh(u @ W1.T + b1)
    h2 = np.tanh(h1 @ W2.T + b2)
    o = h2 @ w3 + b3
    r = o - (t - x[:, 0]) * si
    loss = float(np.mean(r * r))
    d3 = 2.0 * r / n
    gw3 = h2.T @ d3
    gb3 = d3.sum()
    D2 = np.outer(d3, w3) * (1.0 - h2 * h2)
    gW2 = D2.T @ h1
    gb2 = D2.sum(axis=0)
    D1 = (D2 @ W2) * (1.0 - h1 * h1)
    gW1 = D1.T @ u
    gb1 = D1.sum(axis=0)
    g = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gw3, [gb3]])
    return loss, g


class SteepestDescent:
    """Steepest descent (src/minimizer/steepest_descent.rs initial_state,
    step, check_convergence): a trial move x += F step / F_max from the
    current forces, kept when the energy drops (step *= step_increase, at
    most max_step) and undone otherwise (step *= step_decrease).
    ``energy_forces(posq) -> (E, F (n,3))``; positions are f32 like the
    device state, the move is formed in f32."""

    def __init__(self, energy_forces, posq, initial_step, max_step, step_increase=1.2,
                 step_decrease=0.2):
        self.ef = energy_forces
        self.posq = np.array(posq, np.float32, copy=True)
        self.step_len = float(initial_step)
        self.max_step = float(max_step)
        self.inc = float(step_increase)
        self.dec = float(step_decrease)
        self.energy, self.F = self.ef(self.posq)

    def max_force(self):
        return float(np.sqrt((self.F.astype(np.float64) ** 2).sum(1).max()))

    def step(self):
        prev, used = self.energy, self.step_len
        fmax = self.max_force()
        if not fmax > 0.0:
            return dict(energy=prev, max_force=0.0, step=used, accepted=False)
        scale = np.float32(used / fmax)
        trial = self.posq.copy()
        trial[:, :3] = trial[:, :3] + self.F.astype(np.float32) * scale
        e, F = self.ef(trial)
        if e < prev:
            self.posq, self.energy, self.F = trial, e, F
            self.step_len = min(self.step_len * self.inc, self.max_step)
            return dict(energy=e, max_force=self.max_force(), step=used, accepted=True)
        self.step_len *= self.dec
        return dict(energy=prev, max_force=self.max_force(), step=used, accepted=False)


def sd_converged(report, prev_energy, force_tolerance, energy_tolerance):
    """check_convergence (src/minimizer/steepest_descent.rs): 'force_zero',
    'force_tolerance', 'energy_tolerance' or None."""
    if report["max_force"] == 0.0:
        return "force_zero"
    if force_tolerance > 0.0 and report["max_force"] <= force_tolerance:
        return "force_tolerance"
    e = report["energy"]
    if report["accepted"] and energy_tolerance > 0.0 and \
            abs(e - prev_energy) <= energy_tolerance * max(abs(e), abs(prev_energy), 1e-30):
        return "energy_tolerance"
    return None


def mc_volume_draw(seed, counter, max_dv):
    """The Monte-Carlo barostat's draw (src/integrator/mc_barostat.rs:263-275):
    u1, u2 from one Philox4x32-10 block with counter (lo, hi, 0, 0); returns
    (dV = max_dv (2 u1 - 1), u2)."""
    o = np.zeros(4, np.uint32)
    lib().or_philox4x32_10(seed & 0xFFFFFFFF, seed >> 32,
                           np.array([counter & 0xFFFFFFFF, counter >> 32, 0, 0], np.uint32), o)
    u1 = (float(o[0]) + 0.5) / 4294967296.0
    u2 = (float(o[1]) + 0.5) / 4294967296.0
    return max_dv * (2.0 * u1 - 1.0), u2


def mc_scale_com(posq, mass, molecules, lattice, scale):
    """The Monte-Carlo barostat's trial positions (kernels/mc_barostat.cu:
    19-74): every molecule (a list of atom indices) shifted by (scale - 1)
    times its centre of mass, the centre taken relative to its first atom
    through minimum images, in double; returns new f32 positions."""
    L = np.asarray(lattice[:3], np.float64)
    out = np.array(posq, np.float32, copy=True)
    f = float(scale) - 1.0
    for atoms in molecules:
        atoms = np.asarray(atoms)
        ref = posq[atoms[0], :3].astype(np.float64)
        d = posq[atoms, :3].astype(np.float64) - ref
        d -= L * np.floor(d / L + 0.5)
        m = np.asarray(mass, np.float64)[atoms]
        com = ref + (m[:, None] * d).sum(0) / m.sum()
        out[atoms, :3] = out[atoms, :3] + (f * com).astype(np.float32)
    return out
