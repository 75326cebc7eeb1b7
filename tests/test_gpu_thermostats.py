"""CSVR and Berendsen thermostats, the Berendsen barostat and the
Monte-Carlo barostat on the GPU, through the C ABI, against the oracle's
restatement of the reference's steps (src/integrator/csvr.rs,
berendsen.rs, berendsen_barostat.rs, mc_barostat.rs)."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from heddlemd_amd import systems, units as U  # noqa: E402
from heddlemd_amd.engine import HeddleError, engine_for  # noqa: E402
from tests.test_gpu_parity import _jittered_argon, _water_engine  # noqa: E402


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _gas(seed):
    s = _jittered_argon(n_atoms=4000, seed=seed)
    s.lj.epsilon[:] = 0.0
    return s


@pytest.mark.parametrize("kind", ["csvr", "berendsen"])
def test_rescale_thermostat_matches_oracle(oracle, kind):
    # force-free gas started at 1.7x the bath temperature: each step is the
    # drift and the post-force rescale (factor on v), so the velocities and
    # the injected energy follow the oracle's factors alone.  The kinetic
    # energies are FP64 sums in different orders on the two sides, so a
    # factor may round to a neighbouring f32 value: velocities within 1e-6
    # relative, KE and the injection within 1e-9 relative.
    say(f"case: argon 4000 force-free {kind} 60 steps")
    s = _gas(3)
    kt = U.temperature(300.0)
    tau = U.time(5e-14)
    dt = np.float32(U.time(2e-15))
    rng = np.random.default_rng(17)
    v0 = (rng.normal(size=(3, s.n)) * np.sqrt(1.7 * kt / s.mass)).astype(np.float32)
    steps = 60
    kw = dict(csvr=dict(kt=kt, tau=tau, seed=11)) if kind == "csvr" else \
        dict(berendsen=dict(kt=kt, tau=tau))
    e = engine_for(s, dt=float(dt), velocities=v0, **kw)
    e.force_eval(scalars=False)
    e.step(steps, scalars_last=True)
    v1 = e.velocities()
    sc = e.scalars()
    th = e.thermostat()
    assert np.abs(e.forces()).max() == 0.0
    e.close()
    g = 3 * s.n - 3
    c = np.exp(-float(dt) / tau)
    v = v0.copy()
    inj = 0.0
    for step in range(steps):
        k = oracle.kinetic_energy(v, s.mass)
        if kind == "csvr":
            f, kn = oracle.csvr_factor(k, g, 11, step, c, kt / 2)
            inj += kn - k
        else:
            f, d = oracle.berendsen_factor(k, 0.5 * g * kt, float(dt) / tau)
            inj += d
        v = v * np.float32(f)
    k_ref = oracle.kinetic_energy(v, s.mass)
    k0 = oracle.kinetic_energy(v0, s.mass)
    say(f"  KE {k0:.6e} -> {sc['ke']:.6e} (oracle {k_ref:.6e}); injected {th['injection']:.9e} "
        f"(oracle {inj:.9e}); draws {th['draws']}; identical words {np.mean(v1 == v):.4f}")
    assert abs(k_ref / k0 - 1.0) > 0.05
    assert np.all(np.abs(v1 - v) <= 1e-6 * np.abs(v))
    assert sc["ke"] == pytest.approx(k_ref, rel=1e-9)
    assert th["injection"] == pytest.approx(inj, rel=1e-9)
    assert sc["nhc_chain"] == pytest.approx(-inj, rel=1e-9)  # conserved-quantity term
    assert th["draws"] == (steps if kind == "csvr" else 0)


def test_csvr_relaxes_a_lj_liquid_to_the_bath():
    # argon at 120 K started from 300 K velocities under a 50 fs CSVR: the
    # temperature over the last 200 of 600 steps is the bath's within 5 %
    say("case: argon 10000 CSVR 600 steps from 300 K to 120 K")
    from heddlemd_amd.rng import maxwell_boltzmann
    s = systems.config_argon()
    kt = U.temperature(120.0)
    v = maxwell_boltzmann(s.mass, U.temperature(300.0), seed=42, n_constraints=0)
    e = engine_for(s, dt=U.time(5e-15), velocities=v,
                   csvr=dict(kt=kt, tau=U.time(5e-14), seed=3))
    e.force_eval(scalars=False)
    e.step(400)
    temps = []
    for _ in range(20):
        e.step(10, scalars_last=True)
        temps.append(2 * e.scalars()["ke"] / (3 * s.n - 3))
    e.close()
    t = np.mean(temps) / kt
    say(f"  <T>/T_bath over the last 200 steps {t:.4f}")
    assert abs(t - 1.0) < 0.05


def test_berendsen_barostat_move_matches_formula():
    say("case: water 8192 Berendsen barostat one step")
    s = systems.config_water_8192()
    kappa, p0, tau = U.inv_pressure(4.5e-10), U.pressure(1e5), U.time(1e-12)
    e = _water_engine(s)
    e.set_berendsen_barostat(p0, tau, kappa)
    L0 = e.lattice()
    e.force_eval(scalars=False)
    e.step(1, scalars_last=True)
    b = e.barostat()
    L1 = e.lattice()
    e.close()
    dt = U.time(2e-15)
    mu3 = 1 - kappa * (dt / tau) * (p0 - b["pressure"])
    say(f"  P {b['pressure']:.6e}, mu {b['mu']:.9f}")
    assert b["mu"] == pytest.approx(float(np.float32(mu3 ** (1 / 3))), rel=1e-6)
    assert L1[0] == pytest.approx(float(np.float32(L0[0]) * np.float32(b["mu"])), rel=1e-7)
    assert b["draws"] == 0 and b["injection"] == 0.0


def test_one_thermostat_per_context():
    s = _gas(5)
    kt = U.temperature(300.0)
    with pytest.raises(HeddleError):
        engine_for(s, dt=U.time(2e-15), nhc_kt=kt, nhc_tau=U.time(1e-13),
                   csvr=dict(kt=kt, tau=U.time(1e-13), seed=1))


def test_mc_barostat_moves_match_the_oracle(oracle):
    # 12 moves on jittered argon (every atom a molecule) with counters 1..12:
    # the volume proposal from the oracle's Philox draw, the trial positions
    # from the oracle's centre scaling, U_new - U_old against the oracle's
    # energies, and the Metropolis decision from those numbers
    say("case: argon 4000, 12 Monte-Carlo volume moves")
    s = _jittered_argon(n_atoms=4000, seed=21)
    kt = U.temperature(120.0)
    p0 = U.pressure(1.0e8)
    e = engine_for(s, dt=U.time(5e-15))
    lj = s.lj.as_dict()

    def energy(p, L):
        return float(oracle.pair_forces(p, s.types, lj, L)[1].sum())

    n_acc = 0
    for ctr in range(1, 13):
        L = e.lattice()
        pos = e.positions()
        v_old = float(np.float32(L[0])) * float(np.float32(L[1])) * float(np.float32(L[2]))
        max_dv = 0.004 * v_old
        r = e.mc_barostat_move(p0, kt, max_dv, 77, ctr)
        dv, u2 = oracle.mc_volume_draw(77, ctr, max_dv)
        assert r["v_old"] == pytest.approx(v_old, rel=1e-12)
        scale = np.cbrt((v_old + dv) / v_old)
        L_new = np.float32(L[:3]) * np.float32(scale)
        trial = oracle.mc_scale_com(pos, s.mass, [[i] for i in range(s.n)], L, scale)
        Lt = np.array([*L_new.astype(np.float64), 0, 0, 0])
        du = energy(trial, Lt) - energy(pos, L)
        # (a pair within rounding of the cutoff may count on one side only)
        assert r["du"] == pytest.approx(du, rel=1e-4, abs=5e-5)
        w = r["du"] + p0 * dv - s.n * kt * np.log((v_old + dv) / v_old)
        assert r["accepted"] == (w <= 0 or u2 < np.exp(-w / kt)), ctr
        after = e.positions()
        if r["accepted"]:
            n_acc += 1
            assert np.abs(after[:, :3].astype(np.float64) - trial[:, :3]).max() < 1e-5
            assert np.allclose(e.lattice()[:3], L_new, rtol=1e-7)
        else:
            assert after.tobytes() == pos.tobytes()
            assert np.array_equal(e.lattice(), L)
    e.step(5)  # the dynamics carry on from the last state
    e.close()
    say(f"  accepted {n_acc} of 12")
    assert n_acc > 0


def test_mc_barostat_keeps_molecules_rigid():
    say("case: water 8192 (SETTLE), 6 Monte-Carlo volume moves on molecule centres")
    s = systems.config_water_8192()
    g = s.water_groups
    off = np.arange(0, s.n + 1, 3, dtype=np.uint32)
    from heddlemd_amd.engine import Engine
    e = Engine(0)
    e.set_system(s.posq, s.mass, s.types, s.excl_key)
    e.set_lattice(s.lattice)
    e.set_lj(s.lj)
    e.set_spme(s.spme)
    d_oh, d_hh = s.water_geom
    e.set_settle(g, d_oh, d_hh, float(s.mass[g[0, 0]]), float(s.mass[g[0, 1]]))
    e.set_molecules(off, g.reshape(-1))
    e.set_integrator(U.time(2e-15))
    e.setup()
    kt = U.temperature(300.0)
    pos0 = e.positions()
    moved = 0
    for ctr in range(1, 7):
        L = e.lattice()
        r = e.mc_barostat_move(U.pressure(1e5), kt, 0.02 * L[0] * L[1] * L[2], 5, ctr)
        moved += r["accepted"]
    pos = e.positions()
    L = e.lattice()
    e.close()

    def oh(p):
        d = p[g[:, 0], :3].astype(np.float64) - p[g[:, 1], :3]
        d -= L[:3] * np.round(d / L[:3])
        return np.linalg.norm(d, axis=1)

    say(f"  accepted {moved} of 6; O-H drift {np.abs(oh(pos) - oh(pos0)).max():.2e} Bohr")
    assert moved > 0
    assert np.abs(oh(pos) - oh(pos0)).max() < 2e-5


def test_andersen_collisions_bit_exact_with_the_oracle(oracle):
    # force-free gas: each step the kick and drift leave v alone and the
    # collisions redraw about rate x dt of the atoms; the device's draws
    # equal the oracle's word for word over 30 steps
    say("case: argon 4000 force-free Andersen 30 steps")
    s = _gas(7)
    kt = U.temperature(150.0)
    dt = np.float32(U.time(2e-15))
    rate = 1.0 / U.time(2e-13)
    rng = np.random.default_rng(3)
    v0 = (rng.normal(size=(3, s.n)) * np.sqrt(2.0 * kt / s.mass)).astype(np.float32)
    e = engine_for(s, dt=float(dt), velocities=v0, andersen=dict(kt=kt, collision_rate=rate,
                                                                 seed=99))
    e.force_eval(scalars=False)
    e.step(30, scalars_last=True)
    v1 = e.velocities()
    th = e.thermostat()
    e.close()
    v = v0.copy()
    p = np.float32(min(max(rate * float(dt), 0.0), 1.0))
    for step in range(30):
        oracle.andersen(v, s.mass, 99, step, p, np.float32(kt))
    changed = np.mean(np.any(v != v0, axis=0))
    say(f"  p = {p:.4f}, atoms redrawn {changed:.3f} (1 - (1-p)^30 = {1 - (1 - p) ** 30:.3f}), "
        f"draws {th['draws']}")
    assert th["draws"] == 30
    assert v1.tobytes() == v.tobytes()
    assert abs(changed - (1 - (1 - p) ** 30)) < 0.05


def _mtk_setup(seed=31):
    s = _jittered_argon(n_atoms=4000, seed=seed)
    s.lj.r_switch[:] = 0.9 * s.lj.cutoff
    from heddlemd_amd.rng import maxwell_boltzmann
    kt = U.temperature(150.0)
    v0 = maxwell_boltzmann(s.mass, kt, seed=seed, n_constraints=0).astype(np.float32)
    prm = dict(kt=kt, pressure=U.pressure(5e7), tau_t=U.time(1e-13), tau_p=U.time(5e-13))
    return s, v0, prm


def test_mtk_npt_steps_match_the_oracle(oracle):
    # 10 MTK-NPT steps of 2 fs on 4,000 LJ atoms (switched LJ, 150 K, 500
    # bar): positions (unwrapped), velocities, box and the cell momentum
    # against the oracle's FP64 restatement of the same splitting
    say("case: argon 4000 MTK-NPT, 10 steps vs the oracle")
    s, v0, prm = _mtk_setup()
    dt = float(np.float32(U.time(2e-15)))
    e = engine_for(s, dt=dt, velocities=v0, mtk=prm)
    e.force_eval(scalars=True)
    e.step(10, scalars_last=True)
    p, im = e.positions(images=True)
    v1 = e.velocities()
    L1 = e.lattice()
    st = e.mtk()
    sc = e.scalars()
    e.close()
    lj = s.lj.as_dict()

    def forces(x, L):
        posq = np.zeros((s.n, 4), np.float32)
        posq[:, :3] = x
        F, E, W, _ = oracle.pair_forces(posq, s.types, lj, np.array([*L, 0, 0, 0]))
        return F, float(W.sum())

    m = s.mass.astype(np.float64)[:, None]
    x = s.posq[:, :3].astype(np.float64)
    v = v0.T.astype(np.float64)
    L = np.array(s.lattice[:3], np.float64)
    mtk = oracle.MtkNpt(s.n, 0, prm["kt"], prm["pressure"], prm["tau_t"], prm["tau_p"])
    f, w = forces(x, L)
    for _ in range(10):
        x, v, L, f, w = mtk.step(x, v, m, L, forces, dt, f, w)
    xd = p[:, :3].astype(np.float64) + im.T * np.asarray(L1[:3])
    dx = np.abs(xd - x).max()
    dv = np.abs(v1.T - v).max() / np.abs(v).max()
    say(f"  box {L1[0]:.6f} (oracle {L[0]:.6f}), p_eps {st['p_eps']:.6e} (oracle "
        f"{mtk.p_eps:.6e}), max |dx| {dx:.2e} Bohr, max |dv|/|v|max {dv:.2e}, "
        f"P {st['pressure']:.4e} (oracle {mtk.last_p:.4e})")
    assert abs(L1[0] / L[0] - 1) < 2e-7
    assert st["p_eps"] == pytest.approx(mtk.p_eps, rel=1e-3, abs=1e-3 * abs(mtk.p_eps) + 1e-9)
    assert dx < 2e-4
    assert dv < 1e-4
    assert st["pressure"] == pytest.approx(mtk.last_p, rel=1e-4)
    assert st["volume"] == pytest.approx(mtk.last_v, rel=1e-6)
    assert sc["ke"] > 0


def test_mtk_npt_conserves_its_extended_energy():
    say("case: argon 4000 MTK-NPT, 400 steps, conserved quantity")
    s, v0, prm = _mtk_setup(seed=33)
    e = engine_for(s, dt=U.time(2e-15), velocities=v0, mtk=prm)
    e.force_eval(scalars=True)
    hs, vols = [], []
    for _ in range(40):
        e.step(10, scalars_last=True)
        sc = e.scalars()
        st = e.mtk()
        hs.append(sc["ke"] + sc["pe"] + st["extra"])
        vols.append(st["volume"])
    e.close()
    k_ref = 1.5 * s.n * prm["kt"]
    drift = (max(hs) - min(hs)) / k_ref
    say(f"  H spread / K {drift:.2e}; volume {vols[0]:.1f} -> {vols[-1]:.1f} Bohr^3")
    assert drift < 5e-3
    assert abs(vols[-1] / vols[0] - 1) > 1e-3


def test_mtk_npt_refuses_other_thermostats_and_constraints():
    s, v0, prm = _mtk_setup()
    with pytest.raises(HeddleError, match="one thermostat|owns"):
        engine_for(s, dt=U.time(2e-15), nhc_kt=prm["kt"], nhc_tau=U.time(1e-13), mtk=prm)
    w = systems.spc_water(n_mol=512, n_side=8, grid=(32, 32, 32), r_cut_m=0.7e-9)
    with pytest.raises(HeddleError, match="mtk-npt does not support constraints"):
        engine_for(w, dt=U.time(2e-15), mtk=prm)
