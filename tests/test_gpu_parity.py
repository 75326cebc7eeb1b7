"""Parity of the HIP path (libheddle.so through the C ABI) with the CPU
oracle on seeded inputs, plus determinism, time reversal and conservation
properties at benchmark sizes.

Tolerances (f32 device arithmetic against the FP64 oracle):
  * forces: max_i |dF_i| / max(|F_i|, F_rms) <= 1e-5 (the FP32-mode bar)
    and RMS |dF| / F_rms <= 5e-6 (tests/forcecheck.py); atoms in a pair
    whose r^2 lies within 1e-6 of a cutoff^2 are excluded (their inclusion
    depends on rounding order);
  * energy and virial totals: relative 1e-5 (argon: the LJ sum cancels ~30x);
  * velocity Verlet f32 update: bit-exact against the scalar restatement;
  * repeated runs, graph vs eager replay: byte-identical state.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from heddlemd_amd import systems, units as U  # noqa: E402
from tests.forcecheck import check_forces, force_errors  # noqa: E402
from heddlemd_amd.engine import engine_for  # noqa: E402


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _jittered_argon(n_atoms=10000, seed=1):
    s = systems.config_argon() if n_atoms == 10000 else systems.argon_lattice(n_atoms=n_atoms)
    rng = np.random.default_rng(seed)
    p = s.posq[:, :3].astype(np.float64) + rng.normal(0, 0.3, size=(s.n, 3))
    s.posq[:, :3] = systems.wrap_into_box(p, s.lattice).astype(np.float32)
    return s




def test_dpp_rotation_lane_layout():
    say("case: dpp selftest")
    from heddlemd_amd.engine import Engine
    e = Engine(0)
    out = e.selftest_dpp()
    e.close()
    assert out.tolist() == [7 * ((l + 1) % 64) + 3 for l in range(64)]


@pytest.mark.parametrize("bf16", [False, True])
def test_mfma_k16_lane_layout(bf16):
    # the split-precision DFT relies on this A/B/D lane layout of the gfx950
    # v_mfma_f32_32x32x16_{f16,bf16} (exact small-integer products)
    say(f"case: mfma 32x32x16 selftest ({'bf16' if bf16 else 'f16'})")
    from heddlemd_amd.engine import Engine
    e = Engine(0)
    out = e.selftest_mfma_k16(bf16)
    e.close()
    A = np.array([[(16 * i + k) % 61 - 30 for k in range(16)] for i in range(32)], np.float64)
    B = np.array([[(32 * k + j) % 23 - 11 for j in range(32)] for k in range(16)], np.float64)
    D = A @ B
    want = np.zeros((64, 16))
    for l in range(64):
        for v in range(16):
            want[l, v] = D[8 * (v // 4) + 4 * (l // 32) + v % 4, l % 32]
    assert np.array_equal(out.astype(np.float64), want)


def test_mfma_f16_lane_layout():
    # the split-precision DFT relies on this A/B/D lane layout of
    # v_mfma_f32_32x32x8_f16 (exact small-integer products)
    say("case: mfma selftest")
    from heddlemd_amd.engine import Engine
    e = Engine(0)
    out = e.selftest_mfma()
    e.close()
    A = np.array([[(8 * i + k) % 61 - 30 for k in range(8)] for i in range(32)], np.float64)
    B = np.array([[(32 * k + j) % 23 - 11 for j in range(32)] for k in range(8)], np.float64)
    D = A @ B
    want = np.zeros((64, 16))
    for l in range(64):
        for v in range(16):
            want[l, v] = D[8 * (v // 4) + 4 * (l // 32) + v % 4, l % 32]
    assert np.array_equal(out.astype(np.float64), want)


def test_argon_forces_energy_virial(oracle):
    say("case: argon 10000 forces")
    s = _jittered_argon()
    e = engine_for(s, dt=U.time(5e-15))
    e.force_eval(scalars=True)
    f = e.forces()
    pe, w = e.per_atom()
    sc = e.scalars()
    e.close()
    F, E, W, band = oracle.pair_forces(s.posq, s.types, s.lj.as_dict(), s.lattice, band=1e-6)
    check_forces(f, F, band, "argon 10000", say)
    assert sc["pe"] == pytest.approx(E.sum(), rel=1e-5)
    assert sc["virial"] == pytest.approx(W.sum(), rel=1e-5)
    assert pe.sum() == pytest.approx(E.sum(), rel=1e-5)


def _argon_parity(s, tag, oracle):
    e = engine_for(s, dt=U.time(5e-15))
    e.force_eval(scalars=True)
    f = e.forces()
    sc = e.scalars()
    e.close()
    F, E, W, band = oracle.pair_forces(s.posq, s.types, s.lj.as_dict(), s.lattice, band=1e-6)
    check_forces(f, F, band, tag, say)
    assert sc["pe"] == pytest.approx(E.sum(), rel=1e-5)
    assert sc["virial"] == pytest.approx(W.sum(), rel=1e-5)


def test_switched_lj_forces_general_loop(oracle):
    # the reference's cubic switch from 0.9 r_c: the general pair loop
    # (shifted images, per-type select, switch polynomial)
    say("case: argon 10000 forces, switched LJ")
    s = _jittered_argon()
    s.lj.r_switch[:] = 0.9 * s.lj.cutoff
    _argon_parity(s, "switched argon", oracle)


def test_narrow_box_forces_minimum_image_loop(oracle):
    # a 2.38 nm box leaves less than the search radius between a block and
    # half the box: the general loop with a minimum image per pair
    say("case: argon 343 in a 2.38 nm box (per-pair minimum image)")
    s = systems.argon_lattice(n_atoms=343, n_side=7, box_m=7 * 3.40e-10)
    rng = np.random.default_rng(3)
    p = s.posq[:, :3].astype(np.float64) + rng.normal(0, 0.3, size=(s.n, 3))
    s.posq[:, :3] = systems.wrap_into_box(p, s.lattice).astype(np.float32)
    _argon_parity(s, "narrow argon", oracle)


def test_packed_list_covers_pair_set(oracle):
    say("case: argon pair set")
    s = _jittered_argon(seed=3)
    e = engine_for(s, dt=U.time(5e-15))
    gp = e.pairs()
    gp2 = e.pairs()
    useful = e.count_pairs()
    e.close()
    rs = float(s.lj.cutoff[0, 0]) * 1.3  # r_search = r_c + 0.3 r_c (default skin)
    want, band = oracle.pair_set(s.posq, s.lattice, rs)
    got = set(zip(gp[:, 0].tolist(), gp[:, 1].tolist()))
    missing = [p for p in zip(want[:, 0].tolist(), want[:, 1].tolist()) if p not in got]
    assert missing == []
    assert np.array_equal(gp, gp2)
    assert np.all(gp[:, 0] < gp[:, 1])
    within_rc, _ = oracle.pair_set(s.posq, s.lattice, float(s.lj.cutoff[0, 0]))
    assert abs(useful - len(within_rc)) <= 4


def _water_engine(s, **kw):
    return engine_for(s, dt=U.time(2e-15), **kw)


def test_water_forces_energy_virial(oracle):
    say("case: water 8192 forces")
    s = systems.config_water_8192()
    e = _water_engine(s)
    e.force_eval(scalars=True)
    f = e.forces()
    sc = e.scalars()
    rho_g, v_g = e.spme_grids(s.spme.grid)
    e.close()
    Fp, Ep, Wp, band = oracle.pair_forces(s.posq, s.types, s.lj.as_dict(), s.lattice,
                                          coul=dict(alpha=s.spme.alpha, r_cut=s.spme.r_cut),
                                          excl=s.exclusion_csr(), band=1e-6)
    Fr, Er, Wr, er, wr, V = oracle.spme_recip(s.posq, s.lattice, s.spme.grid, 4, s.spme.alpha,
                                              want_v=True)
    check_forces(f, Fp + Fr, band, "water 8192", say)
    assert sc["pe"] == pytest.approx((Ep + Er).sum(), rel=2e-6)
    assert sc["virial"] == pytest.approx((Wp + Wr).sum(), rel=2e-6)
    assert sc["e_recip"] == pytest.approx(er, rel=2e-6)
    assert sc["w_recip"] == pytest.approx(wr, rel=2e-6)
    rho_o = oracle.spme_spread(s.posq, s.lattice, s.spme.grid, 4)
    # f32 spline weights rounded into 2^-32 fixed point vs the exact sum
    rho_err = float(np.abs(rho_g - rho_o).max() / np.abs(rho_o).max())
    v_err = float(np.abs(v_g - V).max() / np.abs(V).max())
    say(f"rho err {rho_err:.3e} V err {v_err:.3e}")
    assert rho_err <= 4e-6
    assert v_err <= 2e-5


def test_vv_update_bit_exact(oracle):
    say("case: argon NVE VV bit-exact")
    s = _jittered_argon(seed=5)
    rng = np.random.default_rng(6)
    v0 = rng.normal(0, 2e-4, size=(3, s.n)).astype(np.float32)
    dt = np.float32(U.time(5e-15))
    e = engine_for(s, dt=float(dt), velocities=v0)
    e.use_graph(False)
    e.force_eval(scalars=False)
    f0 = e.forces().T.copy()
    posq0, img0 = e.positions(images=True)
    e.step(1)
    posq1, img1 = e.positions(images=True)
    v1 = e.velocities()
    f1 = e.forces().T.copy()
    e.close()
    p = posq0.copy()
    img = img0.T.copy()
    v = v0.copy()
    L32 = s.lattice.astype(np.float32)
    oracle.vv_kick_drift_f32(p, img, v, f0, s.mass, L32, dt)
    assert np.array_equal(p, posq1)
    assert np.array_equal(img.T, img1)
    oracle.vv_kick_f32(v, f1, s.mass, dt)
    assert np.array_equal(v, v1)


def _run(s, steps, graph, **kw):
    e = _water_engine(s, **kw)
    e.use_graph(graph)
    e.constrain_velocities()
    e.force_eval(scalars=False)
    e.step(steps, scalars_last=True)
    out = (e.positions(images=True), e.velocities(), e.scalars())
    e.close()
    return out


def test_repeat_runs_and_graph_replay_are_byte_identical():
    say("case: water 8192 determinism, 20 steps")
    s = systems.config_water_8192()
    kt = U.temperature(300.0)
    from heddlemd_amd.rng import maxwell_boltzmann
    v = maxwell_boltzmann(s.mass, kt, seed=7, n_constraints=s.n_constraints)
    kw = dict(nhc_kt=kt, nhc_tau=U.time(1e-13), velocities=v)
    (pa, ia), va, sa = _run(s, 20, True, **kw)
    (pb, ib), vb, sb = _run(s, 20, True, **kw)
    (pc, ic), vc, sc = _run(s, 20, False, **kw)
    assert np.isfinite(pa).all()
    for (p, i, vv) in ((pb, ib, vb), (pc, ic, vc)):
        assert pa.tobytes() == p.tobytes()
        assert ia.tobytes() == i.tobytes()
        assert va.tobytes() == vv.tobytes()
    assert sa == sb == sc


def test_settle_water_step_matches_oracle(oracle):
    # one VV step of 8192 SETTLE waters (every atom in a water, so the drift
    # and SETTLE run as one kernel): the f32 kick + drift is bit-exact, and
    # SETTLE on positions and velocities follows the FP64 oracle to f32
    # resolution (2e-5 Bohr on positions ~ 60 Bohr; the velocity resets are
    # position resets / dt, so 4e-5 / dt on velocities)
    say("case: water 8192 SETTLE step against the oracle")
    s = systems.config_water_8192()
    rng = np.random.default_rng(29)
    v0 = rng.normal(0, 2e-4, size=(3, s.n)).astype(np.float32)
    dt = np.float32(U.time(2e-15))
    e = engine_for(s, dt=float(dt), velocities=v0)
    e.use_graph(False)
    e.force_eval(scalars=False)
    f0 = e.forces().T.copy()
    posq0, img0 = e.positions(images=True)
    e.step(1)
    posq1, img1 = e.positions(images=True)
    v1 = e.velocities()
    f1 = e.forces().T.copy()
    e.close()
    p = posq0.copy()
    img = img0.T.copy()
    v = v0.copy()
    L32 = s.lattice.astype(np.float32)
    oracle.vv_kick_drift_f32(p, img, v, f0, s.mass, L32, dt)
    assert np.array_equal(img.T, img1)
    d_oh, d_hh = s.water_geom
    m_o, m_h = float(s.mass[s.water_groups[0, 0]]), float(s.mass[s.water_groups[0, 1]])
    geom = oracle.settle_geometry(d_oh, d_hh, m_o, m_h)
    pos = p[:, :3].astype(np.float64)
    vel = v.T.astype(np.float64).copy()
    L64 = s.lattice.astype(np.float32).astype(np.float64)  # the device's f32 box
    oracle.settle_positions(s.water_groups, pos, vel, posq0[:, :3].astype(np.float64), geom,
                            m_o, m_h, L64, float(dt))
    oracle.settle_velocities(s.water_groups, pos, vel, m_o, m_h, L64, float(dt))
    vel += 0.5 * float(dt) * f1.T.astype(np.float64) / s.mass[:, None].astype(np.float64)
    dp = np.abs(posq1[:, :3] - pos).max()
    dv = np.abs(v1.T - vel).max()
    say(f"  max |dp| {dp:.2e} Bohr, max |dv| {dv:.2e} (|v| ~ {np.abs(vel).max():.1e})")
    # SETTLE moved the atoms by far more than the tolerance
    assert np.abs(p[:, :3] - pos).max() > 1e-3
    assert dp <= 2e-5
    assert dv <= 4e-5 / float(dt)


def test_settle_keeps_water_rigid():
    say("case: water 8192 rigidity, 40 steps")
    s = systems.config_water_8192()
    kt = U.temperature(300.0)
    (p, _), v, sc = _run(s, 40, True, nhc_kt=kt, nhc_tau=U.time(1e-13))
    d_oh, d_hh = s.water_geom
    L = s.lattice[0]
    g = s.water_groups
    def dist(a, b):
        d = p[a, :3].astype(np.float64) - p[b, :3]
        d -= L * np.floor(d / L + 0.5)
        return np.linalg.norm(d, axis=1)
    assert np.abs(dist(g[:, 0], g[:, 1]) / d_oh - 1).max() < 2e-5
    assert np.abs(dist(g[:, 0], g[:, 2]) / d_oh - 1).max() < 2e-5
    assert np.abs(dist(g[:, 1], g[:, 2]) / d_hh - 1).max() < 2e-5


def test_lossless_time_reversal_bit_exact():
    say("case: argon lossless reversal")
    s = _jittered_argon(n_atoms=4000, seed=9)
    rng = np.random.default_rng(10)
    v0 = rng.normal(0, 1e-4, size=(3, s.n)).astype(np.float32)
    e = engine_for(s, dt=U.time(2e-15), lossless=True, velocities=v0, r_skin=U.length(0.5e-9))
    e.force_eval(scalars=False)
    x0 = e.positions()
    reb0 = e.scalars()["rebuilds"]
    e.step(15)
    e.negate_velocities()
    e.step(15)
    x1 = e.positions()
    v1 = e.velocities()
    reb = e.scalars()["rebuilds"] - reb0
    e.close()
    assert reb == 0, "a rebuild between the passes would change the f32 sum order"
    assert np.array_equal(x0, x1)
    assert np.array_equal(-v1, v0)


def test_nve_energy_conservation():
    say("case: argon NVE energy drift, 200 steps, switched LJ")
    s = _jittered_argon(seed=11)
    # a hard cutoff makes the energy jump whenever a pair crosses r_c (the
    # lattice melts here, so the pair count inside r_c changes by ~15%);
    # the reference's cubic switch from 0.9 r_c (its default r_switch,
    # src/io/config.rs:1175) makes energy and force continuous
    s.lj.r_switch[:] = 0.9 * s.lj.cutoff
    from heddlemd_amd.rng import maxwell_boltzmann
    kt = U.temperature(120.0)
    v = maxwell_boltzmann(s.mass, kt, seed=42)
    e = engine_for(s, dt=U.time(2e-15), velocities=v)
    e.force_eval(scalars=True)
    e.step(1, scalars_last=True)
    a = e.scalars()
    e.step(200, scalars_last=True)
    b = e.scalars()
    e.close()
    e0, e1 = a["ke"] + a["pe"], b["ke"] + b["pe"]
    assert abs(e1 - e0) < 2e-3 * abs(a["ke"])


def test_crescale_move_matches_formula():
    say("case: water 8192 C-rescale one step")
    from tests.test_rng_units_systems import philox4x32_10
    s = systems.config_water_8192()
    kt = U.temperature(300.0)
    kappa, p0, tau = U.inv_pressure(4.5e-10), U.pressure(1e5), U.time(1e-12)
    e = _water_engine(s, crescale=dict(pressure=p0, tau=tau, compressibility=kappa, kt=kt,
                                       seed=13))
    L0 = e.lattice()
    e.force_eval(scalars=False)
    e.step(1, scalars_last=True)
    b = e.barostat()
    L1 = e.lattice()
    e.close()
    v_pre = float(np.float32(L0[0])) * float(np.float32(L0[1])) * float(np.float32(L0[2]))
    o = philox4x32_10((13, 0), (0, 0, 0, 0))
    u1 = (o[0] + 0.5) / 2 ** 32
    u2 = (o[1] + 0.5) / 2 ** 32
    r = np.sqrt(-2 * np.log(u1)) * np.cos(2 * np.pi * u2)
    dt = U.time(2e-15)
    mu3 = 1 - kappa * (dt / tau) * (p0 - b["pressure"]) + np.sqrt(2 * kappa * kt * dt / (tau * v_pre)) * r
    assert b["mu"] == pytest.approx(float(np.float32(mu3 ** (1 / 3))), rel=1e-6)
    assert L1[0] == pytest.approx(float(np.float32(L0[0]) * np.float32(b["mu"])), rel=1e-7)
    assert b["draws"] == 1


def test_barostat_virial_only_steps_match_full_scalar_steps():
    # between logged steps a barostat run evaluates only the virial; the
    # virial (and so every box move) is the one a full energy+virial step
    # gives, and the last logged energy survives the virial-only steps
    say("case: water 8192 C-rescale, 6 steps virial-only vs 6 full-scalar steps")
    s = systems.config_water_8192()
    kt = U.temperature(300.0)
    baro = dict(pressure=U.pressure(1e5), tau=U.time(1e-12),
                compressibility=U.inv_pressure(4.5e-10), kt=kt, seed=13)
    from heddlemd_amd.rng import maxwell_boltzmann
    v = maxwell_boltzmann(s.mass, kt, seed=4, n_constraints=s.n_constraints)
    out = []
    for per_step in (False, True):
        e = _water_engine(s, crescale=baro, velocities=v, nhc_kt=kt, nhc_tau=U.time(1e-13))
        e.constrain_velocities()
        e.force_eval(scalars=True)
        pe0 = e.scalars()["pe"]
        if per_step:
            for _ in range(6):
                e.step(1, scalars_last=True)
        else:
            e.step(6, scalars_last=False)
            assert e.scalars()["pe"] == pe0
            e.step(1, scalars_last=True)
        if per_step:
            e.step(1, scalars_last=True)
        p, _ = e.positions(images=True)
        out.append((p, e.velocities(), e.lattice(), e.scalars()))
        e.close()
    (p0, v0, l0, s0), (p1, v1, l1, s1) = out
    assert np.array_equal(l0, l1)
    assert np.array_equal(p0, p1) and np.array_equal(v0, v1)
    assert s0["pe"] == s1["pe"] and s0["virial"] == s1["virial"]


def test_langevin_baoab_step_matches_oracle(oracle):
    # one BAOAB step of jittered argon: B + A with the initial forces, O + A
    # with draw counter 0, then the fused B with the new forces; f32
    # operations in the reference order on both sides (the Box-Muller
    # log/cos are FP64 library calls on each side, so a last-bit difference
    # of a draw is tolerated)
    say("case: argon 10000 Langevin BAOAB one step")
    s = _jittered_argon(seed=9)
    rng = np.random.default_rng(10)
    v0 = rng.normal(0, 2e-4, size=(3, s.n)).astype(np.float32)
    dt = np.float32(U.time(5e-15))
    kt = U.temperature(120.0)
    gamma = 1.0 / U.time(1e-12)
    e = engine_for(s, dt=float(dt), velocities=v0, langevin=dict(friction=gamma, kt=kt, seed=77))
    e.use_graph(False)
    e.force_eval(scalars=False)
    f0 = e.forces().T.copy()
    posq0, img0 = e.positions(images=True)
    e.step(1)
    posq1, img1 = e.positions(images=True)
    v1 = e.velocities()
    f1 = e.forces().T.copy()
    assert e.langevin_counter() == 1
    e.step(3)
    assert e.langevin_counter() == 4
    e.close()
    p = posq0.copy()
    img = img0.T.copy()
    v = v0.copy()
    L32 = s.lattice.astype(np.float32)
    alpha = oracle.lan_alpha(gamma, dt)
    oracle.lan_kick_drift_half(p, img, v, f0, s.mass, L32, dt)
    oracle.lan_ou_drift_half(p, img, v, s.mass, L32, dt, alpha, np.float32(kt), 77, 0)
    oracle.vv_kick_f32(v, f1, s.mass, dt)
    assert np.array_equal(img.T, img1)
    dp = np.abs(p[:, :3] - posq1[:, :3]).max()
    dv = np.abs(v - v1).max() / np.abs(v).max()
    say(f"  max |dx| {dp:.3e} Bohr, max |dv|/max|v| {dv:.3e}")
    assert dp <= 1e-5
    assert dv <= 1e-6


def test_langevin_water_settle_thermalises():
    # rigid SPC water under BAOAB + SETTLE: geometry stays rigid and the
    # kinetic temperature approaches the bath from a cold start
    say("case: water 8192 Langevin + SETTLE, 400 steps")
    s = systems.config_water_8192()
    kt = U.temperature(300.0)
    e = engine_for(s, dt=U.time(1e-15),
                   langevin=dict(friction=1.0 / U.time(5e-14), kt=kt, seed=3))
    e.constrain_velocities()
    e.force_eval(scalars=False)
    e.step(400, scalars_last=True)
    sc = e.scalars()
    p, _ = e.positions(images=True)
    e.close()
    ndof = 3 * s.n - s.n_constraints - 3
    T = 2 * sc["ke"] / ndof * U.TEMPERATURE_AU_TO_K
    say(f"  T = {T:.1f} K")
    # the lattice start releases ~0.6 Eh per molecule over a few ps; without
    # the bath the same run passes 1000 K
    assert 150.0 < T < 600.0
    d_oh, d_hh = s.water_geom
    L = s.lattice[0]
    g = s.water_groups
    for a, b, d in ((0, 1, d_oh), (0, 2, d_oh), (1, 2, d_hh)):
        r = p[g[:, a], :3].astype(np.float64) - p[g[:, b], :3]
        r -= L * np.floor(r / L + 0.5)
        assert np.abs(np.linalg.norm(r, axis=1) / d - 1).max() < 2e-5


def _vv_constrained_step_oracle(oracle, s, G, posq0, img0, v0, f0, f1, dt, w=None):
    """The reference VV step with SHAKE hooks (run_step,
    src/integrator/mod.rs:344-429): kick + drift with a snapshot, SHAKE,
    RATTLE (dt), then the fused kick with the new forces."""
    p = posq0.copy()
    img = img0.T.copy()
    v = v0.copy()
    L32 = s.lattice.astype(np.float32)
    snap = p.copy()
    oracle.vv_kick_drift_f32(p, img, v, f0, s.mass, L32, dt)
    # the device solves every group in the wave form (group frame,
    # colour-ordered sweeps)
    oracle.set_shake_form(True)
    try:
        wc, it_s = oracle.shake_positions(G, p, v, snap, s.mass, L32, dt)
        it_r = oracle.rattle_velocities(G, p, v, s.mass, L32, dt, w_cons=wc)
    finally:
        oracle.set_shake_form(False)
    oracle.vv_kick_f32(v, f1, s.mass, dt)
    return p, img, v, it_s, it_r


def test_shake_water_step_bit_exact(oracle):
    # SPC water with general SHAKE groups in place of SETTLE: one step on
    # the GPU equals the f32 restatement bit for bit (same operation order,
    # correctly rounded division on both sides)
    say("case: water 8192 SHAKE/RATTLE one VV step")
    s = systems.water_as_shake(systems.config_water_8192())
    kt = U.temperature(300.0)
    from heddlemd_amd.rng import maxwell_boltzmann
    v0 = maxwell_boltzmann(s.mass, kt, seed=21, n_constraints=s.n_constraints)
    dt = np.float32(U.time(2e-15))
    e = engine_for(s, dt=float(dt), velocities=v0)
    e.use_graph(False)
    e.constrain_velocities()
    v0 = e.velocities()
    e.force_eval(scalars=False)
    f0 = e.forces().T.copy()
    posq0, img0 = e.positions(images=True)
    e.step(1)
    posq1, img1 = e.positions(images=True)
    v1 = e.velocities()
    f1 = e.forces().T.copy()
    st = e.constraint_stats()
    e.close()
    G = oracle.ShakeGroups(s.shake_groups)
    p, img, v, it_s, it_r = _vv_constrained_step_oracle(oracle, s, G, posq0, img0, v0, f0, f1,
                                                        dt)
    say(f"  sweeps: shake max {st['shake_max']} (oracle {it_s.max()}), rattle max "
        f"{st['rattle_max']} (oracle {it_r.max()})")
    assert st["shake_sweeps"] == int(it_s.sum())
    assert st["rattle_sweeps"] == int(it_r.sum())
    # groups that stop at the sweep cap: the reference's 3.57e-6 Bohr^2 is
    # at the f32 resolution of |r|^2 here, so some do, on both sides alike
    max_iter = oracle.SHAKE_MAX_ITER
    say(f"  capped: shake {st['shake_capped']} (oracle {int((it_s >= max_iter).sum())}), "
        f"rattle {st['rattle_capped']} (oracle {int((it_r >= max_iter).sum())}) of {G.n}")
    assert st["shake_capped"] == int((it_s >= max_iter).sum())
    assert st["rattle_capped"] == int((it_r >= max_iter).sum())
    assert np.array_equal(img.T, img1)
    assert np.array_equal(p, posq1)
    assert np.array_equal(v, v1)


def test_shake_polymer_chains_step_bit_exact(oracle):
    # 50-bead chains as single SHAKE groups (49 coupled constraints each,
    # beyond the reference's per-group caps), force-free so only the
    # integrator and the constraints act
    say("case: 24 chains x 50 beads SHAKE step")
    rng = np.random.default_rng(23)
    n_ch, n_b, d = 24, 50, 2.89
    L = 120.0
    pos = []
    for c in range(n_ch):
        x = rng.uniform(-L / 4, L / 4, size=3)
        chain = [x]
        for k in range(1, n_b):
            u = rng.normal(size=3)
            chain.append(chain[-1] + d * u / np.linalg.norm(u))
        pos.extend(chain)
    pos = np.asarray(pos)
    n = n_ch * n_b
    posq = np.zeros((n, 4), np.float32)
    posq[:, :3] = systems.wrap_into_box(pos, [L, L, L])
    lj = systems.LJTable(ntypes=1, sigma=np.array([[7.46]]), epsilon=np.array([[0.0]]),
                         cutoff=np.array([[18.9]]), r_switch=np.array([[18.9]]))
    s = systems.System(name="chains", posq=posq, mass=np.full(n, 25520.0, np.float32),
                       types=np.zeros(n, np.uint32),
                       lattice=np.array([L, L, L, 0, 0, 0], np.float64), lj=lj,
                       excl_key=np.repeat(np.arange(n_ch, dtype=np.uint32), n_b),
                       shake_groups=[(list(range(c * n_b, (c + 1) * n_b)),
                                      [(k, k + 1, d) for k in range(n_b - 1)])
                                     for c in range(n_ch)])
    v0 = rng.normal(0, 2.2e-4, size=(3, n)).astype(np.float32)
    dt = np.float32(U.time(2e-15))
    e = engine_for(s, dt=float(dt), velocities=v0)
    e.set_shake_tolerance(5e-5, 1e-8, 200)
    e.use_graph(False)
    e.force_eval(scalars=False)
    f0 = e.forces().T.copy()
    posq0, img0 = e.positions(images=True)
    e.step(1)
    posq1, img1 = e.positions(images=True)
    v1 = e.velocities()
    f1 = e.forces().T.copy()
    st = e.constraint_stats()
    e.close()
    assert np.abs(f0).max() == 0.0 and np.abs(f1).max() == 0.0
    G = oracle.ShakeGroups(s.shake_groups)
    p = posq0.copy()
    img = img0.T.copy()
    v = v0.copy()
    L32 = s.lattice.astype(np.float32)
    snap = p.copy()
    oracle.vv_kick_drift_f32(p, img, v, f0, s.mass, L32, dt)
    wc, it_s = oracle.shake_positions(G, p, v, snap, s.mass, L32, dt, tol=5e-5, max_iter=200)
    it_r = oracle.rattle_velocities(G, p, v, s.mass, L32, dt, w_cons=wc, tol=1e-8, max_iter=200)
    oracle.vv_kick_f32(v, f1, s.mass, dt)
    say(f"  shake sweeps max {st['shake_max']} (oracle {it_s.max()}), capped {st['shake_capped']}")
    assert st["shake_capped"] == 0 and it_s.max() < 200
    assert st["shake_sweeps"] == int(it_s.sum())
    assert np.array_equal(p, posq1)
    assert np.array_equal(v, v1)


def test_nose_hoover_chain_matches_oracle(oracle):
    # force-free gas (epsilon 0) started at 1.7x the bath temperature under
    # a 20 fs chain: each step is the pre-kick chain half step (factor on
    # v), the drift, and the post-force half step (factor on v), so the
    # velocities and the chain state follow the oracle's chain alone.  The
    # kinetic energies are FP64 sums in different orders on the two sides,
    # so a factor may round to a neighbouring f32 value: velocities within
    # 1e-6 relative, chain term and KE within 1e-9 relative.
    say("case: argon 4000 force-free NHC 60 steps")
    s = _jittered_argon(n_atoms=4000, seed=3)
    s.lj.epsilon[:] = 0.0
    kt = U.temperature(300.0)
    tau = U.time(2e-14)
    dt = np.float32(U.time(2e-15))
    rng = np.random.default_rng(17)
    v0 = (rng.normal(size=(3, s.n)) * np.sqrt(1.7 * kt / s.mass)).astype(np.float32)
    steps = 60
    e = engine_for(s, dt=float(dt), velocities=v0, nhc_kt=kt, nhc_tau=tau)
    e.force_eval(scalars=False)
    e.step(steps, scalars_last=True)
    v1 = e.velocities()
    sc = e.scalars()
    assert np.abs(e.forces()).max() == 0.0
    e.close()
    st = oracle.NHCState(s.n, 0, kt, tau)
    v = v0.copy()
    for _ in range(steps):
        f, _ = st.half_step(float(dt), oracle.kinetic_energy(v, s.mass))
        v = v * np.float32(f)
        f, _ = st.half_step(float(dt), oracle.kinetic_energy(v, s.mass))
        v = v * np.float32(f)
    k_ref = oracle.kinetic_energy(v, s.mass)
    k0 = oracle.kinetic_energy(v0, s.mass)
    same = np.mean(v1 == v)
    say(f"  KE {k0:.6e} -> {sc['ke']:.6e} (oracle {k_ref:.6e}); chain {sc['nhc_chain']:.9e} "
        f"(oracle {st.conserved_extra():.9e}); identical velocity words {same:.4f}")
    # the chain moved the gas well away from its start
    assert abs(k_ref / k0 - 1.0) > 0.1
    assert np.all(np.abs(v1 - v) <= 1e-6 * np.abs(v))
    assert sc["ke"] == pytest.approx(k_ref, rel=1e-9)
    assert sc["nhc_chain"] == pytest.approx(st.conserved_extra(), rel=1e-9)


def test_split_precision_dft_matches_fp64(oracle):
    # the FP16 split matrix-core DFT path against the FP64 oracle, next to
    # the FP32 radix FFT path on the same state: potential grid, reciprocal
    # energy and virial, and total forces
    say("case: water 8192 split-precision DFT")
    s = systems.config_water_8192()
    out = {}
    for mode in (0, 1):
        e = _water_engine(s)
        e.set_spme_dft(mode)
        e.force_eval(scalars=True)
        out[mode] = (e.forces(), e.scalars(), e.spme_grids(s.spme.grid)[1])
        e.close()
    Fp, Ep, Wp, band = oracle.pair_forces(s.posq, s.types, s.lj.as_dict(), s.lattice,
                                          coul=dict(alpha=s.spme.alpha, r_cut=s.spme.r_cut),
                                          excl=s.exclusion_csr(), band=1e-6)
    Fr, Er, Wr, er, wr, V = oracle.spme_recip(s.posq, s.lattice, s.spme.grid, 4, s.spme.alpha,
                                              want_v=True)
    for mode, name in ((0, "fp32 fft"), (1, "fp16 split dft")):
        f, sc, v_g = out[mode]
        v_err = float(np.abs(v_g - V).max() / np.abs(V).max())
        say(f"  {name}: V err {v_err:.3e}, E_recip rel {abs(sc['e_recip'] / er - 1):.3e}, "
            f"W_recip rel {abs(sc['w_recip'] / wr - 1):.3e}")
        check_forces(f, Fp + Fr, band, f"  {name}", say)
        assert v_err <= 2e-5
        assert sc["e_recip"] == pytest.approx(er, rel=2e-6)
        assert sc["w_recip"] == pytest.approx(wr, rel=2e-6)


def _replica_engine(spec, r):
    from heddlemd_amd.rng import maxwell_boltzmann
    s = systems.replica_system(spec)
    kt = U.temperature(spec["temperature_K"])
    v = maxwell_boltzmann(s.mass, kt, seed=100 + r, n_constraints=s.n_constraints)
    e = engine_for(s, dt=U.time(1e-15), velocities=v,
                   langevin=dict(friction=1.0 / U.time(1e-13), kt=kt, seed=1000 + r))
    e.constrain_velocities()
    return e


def test_batched_replicas_byte_identical_to_solo_runs():
    # uneven argon and water replicas advanced by one graph launch per step
    # end in exactly the state of their own single-replica runs
    say("case: 6 replicas, batched vs solo, 30 steps")
    from heddlemd_amd.engine import Batch
    specs = [s for s in systems.replica_specs(count=256)][::43][:6]
    say("  " + ", ".join(f"{s['kind']}:{s['n_atoms']}" for s in specs))
    solo = []
    for r, sp in enumerate(specs):
        e = _replica_engine(sp, r)
        e.step(30, scalars_last=True)
        solo.append((e.positions(images=True), e.velocities(), e.scalars()))
        e.close()
    engines = [_replica_engine(sp, r) for r, sp in enumerate(specs)]
    b = Batch(engines)
    b.step(10)
    b.step(20, scalars_last=True)
    b.close()
    for r, e in enumerate(engines):
        (p, i), v, sc = e.positions(images=True), e.velocities(), e.scalars()
        (ps, is_), vs, scs = solo[r]
        assert np.isfinite(p).all()
        assert p.tobytes() == ps.tobytes(), f"replica {r} positions"
        assert i.tobytes() == is_.tobytes()
        assert v.tobytes() == vs.tobytes(), f"replica {r} velocities"
        assert sc == scs
        assert e.langevin_counter() == 30
        e.close()


def test_small_neighbor_kernel_matches_the_general_path():
    # systems of at most 8192 atoms run the displacement check, the voxel
    # binning and the slot gather in one workgroup: argon and water replicas
    # (ragged atom counts, SPME with the spline records written by the
    # neighbour kernel, SETTLE) under Langevin, one water replica with the
    # matrix-core DFT, and one argon replica under NHC (the check then runs
    # in the KE pass), end in the same bits, pair list and rebuild count as
    # on the general path
    say("case: one-workgroup neighbour kernel vs general path, 200 steps")
    specs = [s for s in systems.replica_specs(count=256)][::43][:6]
    say("  " + ", ".join(f"{s['kind']}:{s['n_atoms']}" for s in specs))
    runs = [(sp, r, "") for r, sp in enumerate(specs)] + [(specs[0], 0, "nhc")]
    runs += [(sp, r, "dft") for r, sp in enumerate(specs)
             if sp["kind"] == "water" and systems.replica_system(sp).spme.grid[0] == 32][:1]
    for sp, r, mode in runs:
        out = []
        for small in (True, False):
            if mode == "nhc":
                s = systems.replica_system(sp)
                kt = U.temperature(sp["temperature_K"])
                from heddlemd_amd.rng import maxwell_boltzmann
                v = maxwell_boltzmann(s.mass, kt, seed=7)
                e = engine_for(s, dt=U.time(2e-15), velocities=v, nhc_kt=kt,
                               nhc_tau=U.time(1e-13))
            else:
                e = _replica_engine(sp, r)
            if mode == "dft":
                e.set_spme_dft(1)
            e.set_small_neighbor(small)
            e.step(150)
            e.step(50, scalars_last=True)
            (p, i), v, f, sc = e.positions(images=True), e.velocities(), e.forces(), e.scalars()
            out.append((p, i, v, f, sc, e.pairs()))
            e.close()
        (a, b) = out
        say(f"  {sp['kind']}:{sp['n_atoms']} {mode}: rebuilds "
            f"{a[4]['rebuilds']:.0f}, pairs {len(a[5])}")
        assert a[4]["rebuilds"] >= 2
        for k in range(4):
            assert a[k].tobytes() == b[k].tobytes(), f"{sp} output {k}"
        assert a[4] == b[4]
        assert np.array_equal(a[5], b[5])


def test_charged_chains_real_space_key_map(oracle):
    # chains with alternating charges under SPME: one cutoff for LJ and
    # Coulomb sends every block through the regrouped pair loop, and the
    # 1-2 / 1-3 exclusions (key span 2) through its per-block key map, both
    # in entries that hold excludable atoms and in entries that skip the
    # key test
    import dataclasses
    say("case: charged polymer 40 x 50, SPME 32^3, key span 2")
    s = systems.polymer_melt(n_chains=40, n_beads=50)
    posq = s.posq.copy()
    posq[:, 3] = np.where(np.arange(s.n) % 2 == 0, 0.25, -0.25).astype(np.float32)
    rc = float(s.lj.cutoff[0, 0])
    spme = systems.SpmeParams(alpha=systems.ewald_alpha(rc, 1e-5), r_cut=rc, grid=(32, 32, 32),
                              order=4)
    s = dataclasses.replace(s, posq=posq, spme=spme)
    e = engine_for(s)
    e.force_eval(scalars=True)
    f = e.forces()
    sc = e.scalars()
    e.close()
    Fp, Ep, Wp, band = oracle.pair_forces(s.posq, s.types, s.lj.as_dict(), s.lattice,
                                          coul=dict(alpha=spme.alpha, r_cut=spme.r_cut),
                                          excl=s.exclusion_csr(), band=1e-6)
    Fr, Er, Wr, er, wr = oracle.spme_recip(s.posq, s.lattice, spme.grid, 4, spme.alpha)
    Fa, Ea, Wa = oracle.angle_forces(s.posq, s.lattice, *s.angles)
    check_forces(f, Fp + Fr + Fa, band, "  charged polymer", say)
    assert sc["pe"] == pytest.approx((Ep + Er + Ea).sum(), rel=1e-5)
    assert sc["virial"] == pytest.approx((Wp + Wr + Wa).sum(), rel=1e-5,
                                         abs=1e-6 * np.abs(Wp).sum())


def test_polymer_melt_forces_and_shake_dynamics(oracle):
    # united-atom melt: LJ with 1-2 / 1-3 exclusions (key span), harmonic
    # angles folded into the fixed-point accumulators, then NHC dynamics
    # with every bond SHAKE-constrained
    say("case: polymer melt 40 x 50 forces + 50 SHAKE steps")
    s = systems.polymer_melt(n_chains=40, n_beads=50)
    kt = U.temperature(400.0)
    from heddlemd_amd.rng import maxwell_boltzmann
    v = maxwell_boltzmann(s.mass, kt, seed=5, n_constraints=s.n_constraints)
    # the random-walk start packs beads down to 0.28 nm (LJ sigma 0.395 nm),
    # so the first steps run at 1 fs under a 100 fs thermostat
    e = engine_for(s, dt=U.time(1e-15), velocities=v, nhc_kt=kt, nhc_tau=U.time(1e-13))
    e.set_shake_tolerance(5e-5, 1e-8, 500)
    e.constrain_velocities()
    e.force_eval(scalars=True)
    f = e.forces()
    sc = e.scalars()
    Fp, Ep, Wp, band = oracle.pair_forces(s.posq, s.types, s.lj.as_dict(), s.lattice,
                                          excl=s.exclusion_csr(), band=1e-6)
    Fa, Ea, Wa = oracle.angle_forces(s.posq, s.lattice, *s.angles)
    say(f"  pe {sc['pe']:.6e} vs {(Ep + Ea).sum():.6e}")
    check_forces(f, Fp + Fa, band, "  polymer melt", say)
    assert sc["pe"] == pytest.approx((Ep + Ea).sum(), rel=1e-5)
    assert sc["virial"] == pytest.approx((Wp + Wa).sum(), rel=1e-5, abs=1e-6 * np.abs(Wp).sum())
    e.step(50, scalars_last=True)
    st = e.constraint_stats()
    p, _ = e.positions(images=True)
    sc = e.scalars()
    e.close()
    ndof = 3 * s.n - s.n_constraints - 3
    T = 2 * sc["ke"] / ndof * U.TEMPERATURE_AU_TO_K
    say(f"  T {T:.1f} K, shake sweeps max {st['shake_max']} capped {st['shake_capped']}")
    # the packed start releases ~0.005 Eh per bead in the first 100 fs, so
    # the chain heats up well above the bath before the thermostat catches up
    assert np.isfinite(p).all() and 200 < T < 3000
    assert st["shake_capped"] == 0
    L = s.lattice[0]
    pc = p[:, :3].astype(np.float64).reshape(40, 50, 3)
    r = pc[:, 1:] - pc[:, :-1]
    r -= L * np.floor(r / L + 0.5)
    d = np.linalg.norm(r, axis=2)
    d0 = U.length(0.153e-9)
    assert np.abs(d / d0 - 1).max() < 2e-5
