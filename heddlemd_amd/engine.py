"""Python face of the MD step: a thin owner of an ``hd_ctx``.

This mirrors the reference's set-up path (src/runner.rs SimulationSetup,
ForceField::new, run_md_phase): upload the particle state, configure the
force slots (Lennard-Jones, SPME), the integrator (velocity Verlet, f32 or
lossless), the constraint (SETTLE) and the thermostat (Nose-Hoover chain),
then advance steps.  Every numeric operation happens inside libheddle.so on
the GPU; this module only moves host arrays across the C ABI.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import HeddleError


def _c(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


class Engine:
    def __init__(self, device: int = 0):
        self._lib = _lib.load()
        self._ctx = C.c_void_p()
        rc = self._lib.hd_ctx_create(device, C.byref(self._ctx))
        if rc != 0:
            raise HeddleError(f"hd_ctx_create({device}) failed with {rc}")
        self.n = 0
        self._n_constraints_shake = 0
        self._keep = []

    def close(self):
        if self._ctx:
            self._lib.hd_ctx_destroy(self._ctx)
            self._ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc, what):
        if rc != 0:
            msg = self._lib.hd_last_error(self._ctx)
            raise HeddleError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")

    # ---- configuration -------------------------------------------------
    def set_system(self, posq, mass, types, excl_key=None):
        posq = _c(posq, np.float32).reshape(-1, 4)
        self.n = posq.shape[0]
        key = None
        if excl_key is not None:
            key = _c(excl_key, np.uint32)
            self._keep.append(key)
        self._chk(self._lib.hd_set_system(self._ctx, self.n, posq.reshape(-1),
                                          _c(mass, np.float32), _c(types, np.uint32),
                                          key.ctypes.data if key is not None else None),
                  "hd_set_system")

    def set_lattice(self, lattice):
        self._chk(self._lib.hd_set_lattice(self._ctx, _c(lattice, np.float64)), "hd_set_lattice")

    def set_velocities(self, v):
        v = _c(v, np.float32).reshape(3, -1)
        self._chk(self._lib.hd_set_velocities(self._ctx, _c(v[0], np.float32),
                                              _c(v[1], np.float32), _c(v[2], np.float32)),
                  "hd_set_velocities")

    def set_lj(self, lj):
        nt = int(lj.ntypes)
        f = lambda a: _c(np.asarray(a, np.float64).reshape(-1), np.float64)
        self._chk(self._lib.hd_set_lj(self._ctx, nt, f(lj.sigma), f(lj.epsilon), f(lj.cutoff),
                                      f(lj.r_switch)), "hd_set_lj")

    def set_spme(self, spme):
        na, nb, nc = spme.grid
        self._chk(self._lib.hd_set_spme(self._ctx, spme.alpha, spme.r_cut, na, nb, nc,
                                        spme.order, spme.k_c), "hd_set_spme")

    def set_neighbor(self, r_skin):
        self._chk(self._lib.hd_set_neighbor(self._ctx, float(r_skin)), "hd_set_neighbor")

    def set_settle(self, groups, d_oh, d_hh, m_o, m_h):
        g = _c(groups, np.uint32).reshape(-1)
        self._chk(self._lib.hd_set_settle(self._ctx, g.shape[0] // 3, g, d_oh, d_hh, m_o, m_h),
                  "hd_set_settle")

    def set_integrator(self, dt, lossless=False):
        self._chk(self._lib.hd_set_integrator(self._ctx, float(dt), int(lossless)),
                  "hd_set_integrator")

    def set_nhc(self, kt, tau, chain=3, yoshida=3, n_resp=1, n_constraints=0):
        self._chk(self._lib.hd_set_nhc(self._ctx, float(kt), float(tau), chain, yoshida, n_resp,
                                       int(n_constraints)), "hd_set_nhc")

    def set_crescale(self, pressure, tau, compressibility, kt, seed):
        self._chk(self._lib.hd_set_crescale(self._ctx, float(pressure), float(tau),
                                            float(compressibility), float(kt), int(seed)),
                  "hd_set_crescale")

    def set_csvr(self, kt, tau, seed, n_constraints=0):
        self._chk(self._lib.hd_set_csvr(self._ctx, float(kt), float(tau), int(seed),
                                        int(n_constraints)), "hd_set_csvr")

    def set_andersen(self, kt, collision_rate, seed):
        self._chk(self._lib.hd_set_andersen(self._ctx, float(kt), float(collision_rate),
                                            int(seed)), "hd_set_andersen")

    def set_mtk(self, kt, pressure, tau_t, tau_p, chain_length=3, yoshida_order=3, n_resp=1,
                n_constraints=0):
        """MTK-NPT integrator (owns the thermostat and the barostat)."""
        self._chk(self._lib.hd_set_mtk(self._ctx, float(kt), float(pressure), float(tau_t),
                                       float(tau_p), int(chain_length), int(yoshida_order),
                                       int(n_resp), int(n_constraints)), "hd_set_mtk")

    def mtk(self):
        """dict(pressure, volume, extra, p_eps, eps) of the last MTK step;
        extra = the conserved quantity less K + U."""
        out = np.zeros(5, np.float64)
        self._chk(self._lib.hd_get_mtk(self._ctx, out), "hd_get_mtk")
        return dict(zip(["pressure", "volume", "extra", "p_eps", "eps"], (float(x) for x in out)))

    def set_berendsen(self, kt, tau, n_constraints=0):
        self._chk(self._lib.hd_set_berendsen(self._ctx, float(kt), float(tau),
                                             int(n_constraints)), "hd_set_berendsen")

    def set_berendsen_barostat(self, pressure, tau, compressibility):
        self._chk(self._lib.hd_set_berendsen_barostat(self._ctx, float(pressure), float(tau),
                                                      float(compressibility)),
                  "hd_set_berendsen_barostat")

    def thermostat(self):
        """CSVR / Berendsen: injected energy and draw counter."""
        out = np.zeros(2)
        self._chk(self._lib.hd_get_thermostat(self._ctx, out), "hd_get_thermostat")
        return dict(injection=float(out[0]), draws=int(out[1]))

    def minimize_begin(self, initial_step):
        """Steepest descent start: dict(energy, max_force, step, accepted)."""
        out = np.zeros(4)
        self._chk(self._lib.hd_minimize_begin(self._ctx, float(initial_step), out),
                  "hd_minimize_begin")
        return dict(energy=out[0], max_force=out[1], step=out[2], accepted=bool(out[3]))

    def minimize_step(self, max_step, step_increase=1.2, step_decrease=0.2):
        out = np.zeros(4)
        self._chk(self._lib.hd_minimize_step(self._ctx, float(max_step), float(step_increase),
                                             float(step_decrease), out), "hd_minimize_step")
        return dict(energy=out[0], max_force=out[1], step=out[2], accepted=bool(out[3]))

    def set_molecules(self, offsets, atoms):
        off = np.ascontiguousarray(offsets, np.uint32)
        at = np.ascontiguousarray(atoms, np.uint32)
        self._chk(self._lib.hd_set_molecules(self._ctx, len(off) - 1, off, at),
                  "hd_set_molecules")

    def mc_barostat_move(self, pressure, kt, max_dv, seed, counter):
        """One Monte-Carlo volume move: dict(accepted, v_old, v_new, du)."""
        out = np.zeros(4)
        self._chk(self._lib.hd_mc_barostat_move(self._ctx, float(pressure), float(kt),
                                                float(max_dv), int(seed), int(counter), out),
                  "hd_mc_barostat_move")
        return dict(accepted=bool(out[0]), v_old=out[1], v_new=out[2], du=out[3])

    def set_spme_dft(self, mode):
        """0: FP32 radix FFT; 1: split-precision FP16 matrix-core DFT; 2: the
        FFT with every pass a launch of its own (equal to 0 to f32 rounding)."""
        self._chk(self._lib.hd_set_spme_dft(self._ctx, int(mode)), "hd_set_spme_dft")

    def set_angles(self, ijk, k_theta, theta0):
        ijk = np.ascontiguousarray(ijk, np.uint32).reshape(-1)
        n = ijk.size // 3
        self._chk(self._lib.hd_set_angles(self._ctx, n, ijk,
                                          np.ascontiguousarray(k_theta, np.float64),
                                          np.ascontiguousarray(theta0, np.float64)),
                  "hd_set_angles")

    def set_bonds(self, ij, kind, prm):
        """kind 0 harmonic (k, r0, -), 1 Morse (De, a, re) per bond."""
        ij = np.ascontiguousarray(ij, np.uint32).reshape(-1)
        kind = np.ascontiguousarray(kind, np.int32).reshape(-1)
        prm = np.ascontiguousarray(prm, np.float64).reshape(-1)
        if ij.size != 2 * kind.size or prm.size != 3 * kind.size:
            raise ValueError("bonds: ij needs 2 and prm 3 entries per bond")
        self._chk(self._lib.hd_set_bonds(self._ctx, kind.size, ij,
                                         kind.ctypes.data_as(C.POINTER(C.c_int)), prm),
                  "hd_set_bonds")

    def set_dihedrals(self, ijkl, prm):
        """Periodic dihedrals; prm rows (k, n, phi0)."""
        ijkl = np.ascontiguousarray(ijkl, np.uint32).reshape(-1)
        prm = np.asarray(prm, np.float64).reshape(-1, 3)
        if ijkl.size != 4 * len(prm):
            raise ValueError("dihedrals: ijkl needs 4 entries per dihedral")
        self._chk(self._lib.hd_set_dihedrals(self._ctx, len(prm), ijkl,
                                             np.ascontiguousarray(prm[:, 0]),
                                             np.ascontiguousarray(prm[:, 1]),
                                             np.ascontiguousarray(prm[:, 2])),
                  "hd_set_dihedrals")

    def set_exclusions(self, ij, scale_lj, scale_coul):
        """The effective exclusion list: pairs (n, 2) with per-term scales."""
        ij = np.ascontiguousarray(ij, np.uint32).reshape(-1)
        slj = np.ascontiguousarray(scale_lj, np.float64).reshape(-1)
        sc = np.ascontiguousarray(scale_coul, np.float64).reshape(-1)
        if ij.size != 2 * slj.size or sc.size != slj.size:
            raise ValueError("exclusions: 2 atoms and 2 scales per pair")
        self._chk(self._lib.hd_set_exclusions(self._ctx, slj.size, ij, slj, sc),
                  "hd_set_exclusions")

    def set_exclusion_span(self, span):
        self._chk(self._lib.hd_set_exclusion_span(self._ctx, int(span)), "hd_set_exclusion_span")

    def set_pair_list(self, mode):
        """0: packed cluster list (half, j-side atomics); 1: per-atom full
        lists."""
        self._chk(self._lib.hd_set_pair_list(self._ctx, int(mode)), "hd_set_pair_list")

    def set_shake(self, groups):
        """groups: list of (atom_ids, [(local_i, local_j, d), ...])."""
        goff, gat, coff, ci, cj, d = [0], [], [0], [], [], []
        for atoms, cons in groups:
            gat.extend(int(a) for a in atoms)
            goff.append(len(gat))
            for a, b, dist in cons:
                ci.append(a)
                cj.append(b)
                d.append(float(dist))
            coff.append(len(ci))
        self._chk(self._lib.hd_set_shake(self._ctx, len(groups), np.asarray(goff, np.uint32),
                                         np.asarray(gat, np.uint32), np.asarray(coff, np.uint32),
                                         np.asarray(ci, np.uint8), np.asarray(cj, np.uint8),
                                         np.asarray(d, np.float64)), "hd_set_shake")
        self._n_constraints_shake = len(ci)

    def add_settle(self, groups, d_oh, d_hh, m_o, m_h):
        """Further SETTLE groups with their own geometry (hd_add_settle)."""
        g = _c(groups, np.uint32).reshape(-1)
        self._chk(self._lib.hd_add_settle(self._ctx, g.shape[0] // 3, g, d_oh, d_hh, m_o, m_h),
                  "hd_add_settle")

    def set_shake_tolerance(self, tol, rattle_tol, max_iter):
        self._chk(self._lib.hd_set_shake_tolerance(self._ctx, float(tol), float(rattle_tol),
                                                   int(max_iter)), "hd_set_shake_tolerance")

    def constraint_stats(self):
        out = np.zeros(6, np.uint64)
        self._chk(self._lib.hd_get_constraint_stats(self._ctx, out), "hd_get_constraint_stats")
        keys = ["shake_sweeps", "shake_max", "shake_capped", "rattle_sweeps", "rattle_max",
                "rattle_capped"]
        return dict(zip(keys, (int(x) for x in out)))

    def set_shake_predictor(self, params):
        """Warm-start every SHAKE solve from the predictor with these
        parameters (a ShakePredictor or its parameter array); None turns
        prediction off."""
        if params is None:
            ptr = None
        else:
            if isinstance(params, ShakePredictor):
                params = params.params
            params = np.ascontiguousarray(params, np.float32)
            if params.size != self._lib.hd_mlp_param_count():
                raise ValueError("predictor parameter count mismatch")
            ptr = params.ctypes.data
        self._chk(self._lib.hd_set_shake_predictor(self._ctx, ptr), "hd_set_shake_predictor")

    def shake_multipliers(self):
        """(predicted nu0, converged nu) of the last SHAKE solve."""
        n = int(self._n_constraints_shake)
        nu0 = np.zeros(n, np.float32)
        nu = np.zeros(n, np.float32)
        self._chk(self._lib.hd_get_shake_multipliers(self._ctx, nu0.ctypes.data, nu.ctypes.data),
                  "hd_get_shake_multipliers")
        return nu0, nu

    def shake_log_start(self, max_steps):
        self._chk(self._lib.hd_shake_log_start(self._ctx, int(max_steps)), "hd_shake_log_start")

    def shake_log_stop(self):
        self._chk(self._lib.hd_shake_log_stop(self._ctx), "hd_shake_log_stop")

    def shake_log_size(self):
        steps, rows = C.c_uint64(0), C.c_uint64(0)
        self._chk(self._lib.hd_shake_log_size(self._ctx, C.byref(steps), C.byref(rows)),
                  "hd_shake_log_size")
        return int(steps.value), int(rows.value)

    def shake_log(self, first_step=0, n_steps=None):
        """Logged (features [rows, 8], converged nu [rows])."""
        steps, rows = self.shake_log_size()
        if n_steps is None:
            n_steps = steps - first_step
        x = np.zeros((n_steps * rows, 8), np.float32)
        t = np.zeros(n_steps * rows, np.float32)
        self._chk(self._lib.hd_shake_log_read(self._ctx, int(first_step), int(n_steps), x, t),
                  "hd_shake_log_read")
        return x, t

    def set_langevin(self, friction, kt, seed):
        self._chk(self._lib.hd_set_langevin(self._ctx, float(friction), float(kt), int(seed)),
                  "hd_set_langevin")

    def langevin_counter(self):
        n = C.c_uint64(0)
        self._chk(self._lib.hd_get_langevin_counter(self._ctx, C.byref(n)),
                  "hd_get_langevin_counter")
        return int(n.value)

    def setup(self):
        self._chk(self._lib.hd_setup(self._ctx), "hd_setup")

    def set_positions(self, posq):
        self._chk(self._lib.hd_set_positions(self._ctx, _c(posq, np.float32).reshape(-1)),
                  "hd_set_positions")

    def neighbor_pre_step(self):
        st = C.c_uint32(0)
        self._chk(self._lib.hd_neighbor_pre_step(self._ctx, C.byref(st)), "hd_neighbor_pre_step")
        return int(st.value)

    def set_fused_settle(self, on):
        self._chk(self._lib.hd_set_fused_settle(self._ctx, 1 if on else 0), "hd_set_fused_settle")

    def set_small_neighbor(self, on):
        """Small systems: neighbour check, binning and gather in one
        workgroup (default) or the general multi-kernel path (identical
        results)."""
        self._chk(self._lib.hd_set_small_neighbor(self._ctx, 1 if on else 0),
                  "hd_set_small_neighbor")

    def use_graph(self, on):
        self._chk(self._lib.hd_use_graph(self._ctx, int(bool(on))), "hd_use_graph")

    def set_timing(self, on):
        self._chk(self._lib.hd_set_timing(self._ctx, int(bool(on))), "hd_set_timing")

    # ---- work ------------------------------------------------------------
    def force_eval(self, scalars=True):
        self._chk(self._lib.hd_force_eval(self._ctx, int(scalars)), "hd_force_eval")

    def constrain_velocities(self):
        self._chk(self._lib.hd_constrain_velocities(self._ctx), "hd_constrain_velocities")

    def step(self, n_steps, scalars_last=False):
        self._chk(self._lib.hd_step(self._ctx, int(n_steps), int(bool(scalars_last))), "hd_step")

    def negate_velocities(self):
        self._chk(self._lib.hd_negate_velocities(self._ctx), "hd_negate_velocities")

    def sync(self):
        self._chk(self._lib.hd_sync(self._ctx), "hd_sync")

    # ---- downloads ---------------------------------------------------------
    def positions(self, images=False):
        posq = np.zeros((self.n, 4), np.float32)
        if images:
            im = np.zeros((3, self.n), np.int32)
            self._chk(self._lib.hd_get_positions(self._ctx, posq.reshape(-1), im[0].ctypes.data,
                                                 im[1].ctypes.data, im[2].ctypes.data),
                      "hd_get_positions")
            return posq, im
        self._chk(self._lib.hd_get_positions(self._ctx, posq.reshape(-1), None, None, None),
                  "hd_get_positions")
        return posq

    def velocities(self):
        v = np.zeros((3, self.n), np.float32)
        self._chk(self._lib.hd_get_velocities(self._ctx, v[0], v[1], v[2]), "hd_get_velocities")
        return v

    def forces(self):
        f = np.zeros((3, self.n), np.float32)
        self._chk(self._lib.hd_get_forces(self._ctx, f[0], f[1], f[2]), "hd_get_forces")
        return f.T.copy()

    def per_atom(self):
        pe = np.zeros(self.n, np.float32)
        w = np.zeros(self.n, np.float32)
        self._chk(self._lib.hd_get_per_atom(self._ctx, pe, w), "hd_get_per_atom")
        return pe, w

    def lattice(self):
        L = np.zeros(6)
        self._chk(self._lib.hd_get_lattice(self._ctx, L), "hd_get_lattice")
        return L

    def scalars(self):
        out = np.zeros(8)
        self._chk(self._lib.hd_get_scalars(self._ctx, out), "hd_get_scalars")
        keys = ["ke", "pe", "virial", "e_recip", "w_recip", "nhc_chain", "rebuilds", "status"]
        return dict(zip(keys, out.tolist()))

    def pairs(self):
        n = C.c_uint64(0)
        self._chk(self._lib.hd_get_pairs(self._ctx, None, 0, C.byref(n)), "hd_get_pairs")
        out = np.zeros(2 * max(n.value, 1), np.uint32)
        self._chk(self._lib.hd_get_pairs(self._ctx, out.ctypes.data, n.value, C.byref(n)),
                  "hd_get_pairs")
        return out[:2 * n.value].reshape(-1, 2)

    def count_pairs(self):
        n = C.c_uint64(0)
        self._chk(self._lib.hd_count_pairs(self._ctx, C.byref(n)), "hd_count_pairs")
        return int(n.value)

    def barostat(self):
        out = np.zeros(5)
        self._chk(self._lib.hd_get_barostat(self._ctx, out), "hd_get_barostat")
        return dict(zip(["mu", "pressure", "volume", "injection", "draws"], out.tolist()))

    def neighbor_stats(self):
        out = np.zeros(6, np.uint64)
        self._chk(self._lib.hd_get_neighbor_stats(self._ctx, out), "hd_get_neighbor_stats")
        keys = ["blocks", "capacity", "max_entries", "entries", "rebuilds", "status"]
        return dict(zip(keys, [int(x) for x in out]))

    def spme_grids(self, grid):
        na, nb, nc = grid
        rho = np.zeros(na * nb * nc, np.float32)
        v = np.zeros(na * nb * nc, np.float32)
        self._chk(self._lib.hd_get_spme_grids(self._ctx, rho.ctypes.data, v.ctypes.data),
                  "hd_get_spme_grids")
        return rho.reshape(grid), v.reshape(grid)

    def selftest_mfma(self):
        out = np.zeros(1024, np.float32)
        self._chk(self._lib.hd_selftest_mfma(self._ctx, out), "hd_selftest_mfma")
        return out.reshape(64, 16)

    def selftest_mfma_k16(self, bf16=False):
        out = np.zeros(1024, np.float32)
        self._chk(self._lib.hd_selftest_mfma_k16(self._ctx, int(bool(bf16)), out),
                  "hd_selftest_mfma_k16")
        return out.reshape(64, 16)

    def selftest_dpp(self):
        out = np.zeros(64, np.int32)
        self._chk(self._lib.hd_selftest_dpp(self._ctx, out), "hd_selftest_dpp")
        return out

    def stage_times(self):
        names = C.create_string_buffer(64 * 64)
        ms = np.zeros(64)
        k = self._lib.hd_stage_times(self._ctx, names, 64, ms, 64)
        if k < 0:
            self._chk(k, "hd_stage_times")
        raw = names.raw
        return [(raw[i * 64:(i + 1) * 64].split(b"\0")[0].decode(), float(ms[i]))
                for i in range(k)]


def engine_for(system, dt=None, lossless=False, nhc_kt=None, nhc_tau=None, r_skin=None,
               velocities=None, crescale=None, langevin=None, device=0, pair_list=None,
               csvr=None, berendsen=None, berendsen_barostat=None, andersen=None, mtk=None):
    """Configure an Engine for a ``systems.System``; ``langevin`` =
    dict(friction, kt, seed) selects the BAOAB integrator; ``pair_list``
    selects the neighbour-list form (hd_set_pair_list)."""
    e = Engine(device)
    if pair_list is not None:
        e.set_pair_list(pair_list)
    e.set_system(system.posq, system.mass, system.types, system.excl_key)
    e.set_lattice(system.lattice)
    e.set_lj(system.lj)
    if system.spme is not None:
        e.set_spme(system.spme)
    if r_skin is not None:
        e.set_neighbor(r_skin)
    if getattr(system, "shake_groups", None) is not None:
        e.set_shake(system.shake_groups)
    if getattr(system, "angles", None) is not None:
        ijk, kth, th0 = system.angles
        e.set_angles(ijk, kth, th0)
    if getattr(system, "excl_span", 0):
        e.set_exclusion_span(system.excl_span)
    if getattr(system, "bonds", None) is not None:
        e.set_bonds(*system.bonds)
    if getattr(system, "dihedrals", None) is not None:
        e.set_dihedrals(*system.dihedrals)
    if getattr(system, "exclusions", None) is not None:
        e.set_exclusions(*system.exclusions)
    if system.water_groups is not None:
        d_oh, d_hh = system.water_geom
        m_o = float(system.mass[system.water_groups[0, 0]])
        m_h = float(system.mass[system.water_groups[0, 1]])
        e.set_settle(system.water_groups, d_oh, d_hh, m_o, m_h)
    if dt is not None:
        e.set_integrator(dt, lossless)
    if langevin is not None:
        e.set_langevin(**langevin)
    if nhc_kt is not None:
        e.set_nhc(nhc_kt, nhc_tau, n_constraints=system.n_constraints)
    if csvr is not None:
        e.set_csvr(n_constraints=system.n_constraints, **csvr)
    if berendsen is not None:
        e.set_berendsen(n_constraints=system.n_constraints, **berendsen)
    if andersen is not None:
        e.set_andersen(**andersen)
    if mtk is not None:
        e.set_mtk(**mtk)
    if berendsen_barostat is not None:
        e.set_berendsen_barostat(**berendsen_barostat)
    if crescale is not None:
        e.set_crescale(**crescale)
    if velocities is not None:
        e.set_velocities(velocities)
    e.setup()
    return e


class ShakePredictor:
    """The SHAKE multiplier predictor (hd_mlp_*): an 8 -> 64 -> 64 -> 1
    tanh network trained on the device."""

    def __init__(self, device=0, seed=1):
        self._lib = _lib.load()
        h = C.c_void_p()
        rc = self._lib.hd_mlp_create(int(device), int(seed), C.byref(h))
        if rc != 0:
            raise _lib.HeddleError(f"hd_mlp_create failed ({rc})")
        self._h = h
        self.n_params = self._lib.hd_mlp_param_count()

    def _chk(self, rc, what):
        if rc != 0:
            msg = self._lib.hd_mlp_last_error(self._h).decode()
            raise _lib.HeddleError(f"{what} failed ({rc}): {msg}")

    @property
    def params(self):
        p = np.zeros(self.n_params, np.float32)
        self._chk(self._lib.hd_mlp_get_params(self._h, p), "hd_mlp_get_params")
        return p

    @params.setter
    def params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        if p.size != self.n_params:
            raise ValueError("parameter count mismatch")
        self._chk(self._lib.hd_mlp_set_params(self._h, p), "hd_mlp_set_params")

    def predict(self, x):
        x = np.ascontiguousarray(x, np.float32).reshape(-1, 8)
        nu = np.zeros(x.shape[0], np.float32)
        self._chk(self._lib.hd_mlp_predict(self._h, x, x.shape[0], nu), "hd_mlp_predict")
        return nu

    def gradient(self, x, t):
        x = np.ascontiguousarray(x, np.float32).reshape(-1, 8)
        t = np.ascontiguousarray(t, np.float32)
        g = np.zeros(self.n_params - 1, np.float32)
        loss = C.c_double(0)
        self._chk(self._lib.hd_mlp_gradient(self._h, x, t, x.shape[0], g, C.byref(loss)),
                  "hd_mlp_gradient")
        return float(loss.value), g

    def train(self, x, t, epochs=1, batch=1024, lr=1e-3, seed=7):
        x = np.ascontiguousarray(x, np.float32).reshape(-1, 8)
        t = np.ascontiguousarray(t, np.float32)
        loss = np.zeros(epochs, np.float64)
        self._chk(self._lib.hd_mlp_train(self._h, x, t, x.shape[0], int(epochs), int(batch),
                                         float(lr), int(seed), loss.ctypes.data),
                  "hd_mlp_train")
        return loss

    def train_on_log(self, engine, epochs=1, batch=1024, lr=1e-3, seed=7):
        loss = np.zeros(epochs, np.float64)
        self._chk(self._lib.hd_mlp_train_log(self._h, engine._ctx, int(epochs), int(batch),
                                             float(lr), int(seed), loss.ctypes.data),
                  "hd_mlp_train_log")
        return loss

    def close(self):
        if self._h:
            self._lib.hd_mlp_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Many set-up Engines advanced by one graph launch per step
    (hd_batch_*); the engines stay owned by the caller."""

    def __init__(self, engines):
        self._lib = engines[0]._lib
        self._engines = list(engines)
        arr = (C.c_void_p * len(engines))(*[e._ctx.value for e in engines])
        h = C.c_void_p()
        rc = self._lib.hd_batch_create(arr, len(engines), C.byref(h))
        if rc != 0:
            raise RuntimeError(f"hd_batch_create failed ({rc})")
        self._h = h

    def step(self, n_steps, scalars_last=False):
        rc = self._lib.hd_batch_step(self._h, int(n_steps), int(bool(scalars_last)))
        if rc != 0:
            msg = self._lib.hd_batch_last_error(self._h).decode()
            raise RuntimeError(f"hd_batch_step failed ({rc}): {msg}")

    def set_lanes(self, lanes):
        rc = self._lib.hd_batch_set_lanes(self._h, int(lanes))
        if rc != 0:
            raise RuntimeError(f"hd_batch_set_lanes failed ({rc})")

    def autotune(self, steps=10, candidates=(0, 64, 32, 16, 8, 4, 2)):
        """Time `steps` steps per graph shape and keep the fastest; the
        shape does not change any replica's arithmetic (the steps advance
        the replicas, so run this in a warm-up phase)."""
        import time
        best = None
        timings = {}
        for lanes in candidates:
            self.set_lanes(lanes)
            self.step(1)  # capture
            t0 = time.perf_counter()
            self.step(steps)
            dt = time.perf_counter() - t0
            timings[lanes] = dt
            if best is None or dt < timings[best]:
                best = lanes
        self.set_lanes(best)
        return best, timings

    def close(self):
        if self._h:
            self._lib.hd_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

