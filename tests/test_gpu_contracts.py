"""GPU tests of the step's contracts beyond the force parity: SETTLE on the
fused and unfused paths against the FP64 oracle, the neighbour displacement
check against the oracle's rule, the status word polled between graph
launches, graph re-capture after re-parameterisation, and the refusals the
reference makes (lossless mode with constraints or a barostat).

All calls go through the C ABI (libheddle.so)."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from heddlemd_amd import systems, units as U  # noqa: E402
from heddlemd_amd._lib import HeddleError  # noqa: E402
from heddlemd_amd.engine import engine_for  # noqa: E402
from heddlemd_amd.rng import maxwell_boltzmann  # noqa: E402


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _settle_reference_step(oracle, s, posq0, img0, v0, f0, f1, dt, groups):
    """The f32 kick + drift of every atom (bit-exact restatement), then SETTLE
    on positions and velocities of the given waters in FP64 (the oracle's
    Miyamoto-Kollman restatement), then the closing half kick."""
    p = posq0.copy()
    img = img0.T.copy()
    v = v0.copy()
    L32 = s.lattice.astype(np.float32)
    oracle.vv_kick_drift_f32(p, img, v, f0, s.mass, L32, dt)
    d_oh, d_hh = s.water_geom
    m_o, m_h = float(s.mass[groups[0, 0]]), float(s.mass[groups[0, 1]])
    geom = oracle.settle_geometry(d_oh, d_hh, m_o, m_h)
    pos = p[:, :3].astype(np.float64)
    vel = v.T.astype(np.float64).copy()
    # the device holds the box in f32: the oracle takes the same lengths
    L32 = s.lattice.astype(np.float32).astype(np.float64)
    oracle.settle_positions(groups, pos, vel, posq0[:, :3].astype(np.float64), geom, m_o, m_h,
                            L32, float(dt))
    oracle.settle_velocities(groups, pos, vel, m_o, m_h, L32, float(dt))
    vel += 0.5 * float(dt) * f1.T.astype(np.float64) / s.mass[:, None].astype(np.float64)
    return p, img, pos, vel


def _one_step(s, v0, dt, fused=True):
    e = engine_for(s, dt=float(dt), velocities=v0)
    e.use_graph(False)
    e.set_fused_settle(fused)
    e.force_eval(scalars=False)
    f0 = e.forces().T.copy()
    posq0, img0 = e.positions(images=True)
    e.step(1)
    posq1, img1 = e.positions(images=True)
    v1 = e.velocities()
    f1 = e.forces().T.copy()
    e.close()
    return posq0, img0, f0, posq1, img1, v1, f1


def test_settle_fused_and_unfused_paths_bit_identical_and_match_oracle(oracle):
    # one VV step of 8192 SETTLE waters through the fused drift + SETTLE
    # kernel and through the three separate kernels: the same bits, and the
    # FP64 oracle (Miyamoto-Kollman) to f32 resolution
    say("case: water 8192 SETTLE step, fused vs unfused vs oracle")
    s = systems.config_water_8192()
    rng = np.random.default_rng(29)
    v0 = rng.normal(0, 2e-4, size=(3, s.n)).astype(np.float32)
    dt = np.float32(U.time(2e-15))
    out = {f: _one_step(s, v0, dt, fused=f) for f in (True, False)}
    for k in range(len(out[True])):
        assert out[True][k].tobytes() == out[False][k].tobytes(), f"fused/unfused output {k}"
    posq0, img0, f0, posq1, img1, v1, f1 = out[True]
    p, img, pos, vel = _settle_reference_step(oracle, s, posq0, img0, v0, f0, f1, dt,
                                              s.water_groups)
    assert np.array_equal(img.T, img1)
    dp = np.abs(posq1[:, :3] - pos)
    spacing = np.spacing(np.abs(pos).astype(np.float32)).astype(np.float64)
    dv = np.abs(v1.T - vel).max()
    say(f"  max |dp| {dp.max():.2e} Bohr ({(dp / spacing).max():.2f} f32 spacings), "
        f"max |dv| {dv:.2e} (|v| ~ {np.abs(vel).max():.1e})")
    assert np.abs(p[:, :3] - pos).max() > 1e-3  # SETTLE moved the atoms
    assert (dp <= spacing).all()                # within one f32 spacing of FP64
    assert dv <= 1e-5 * np.abs(vel).max()


def test_langevin_settle_fused_and_unfused_paths_bit_identical():
    # BAOAB on 8192 SETTLE waters: the fused kernel (both half steps and the
    # three SETTLE stages) against the five separate kernels, eager and
    # through captured step graphs: the same bits after 12 steps, and the
    # draw counter advanced once per step (the fused kernel's last block
    # advances it)
    say("case: water 8192 Langevin + SETTLE, fused vs unfused, eager and graph, 12 steps")
    s = systems.config_water_8192()
    rng = np.random.default_rng(31)
    v0 = rng.normal(0, 2e-4, size=(3, s.n)).astype(np.float32)
    kt = U.temperature(300.0)
    out = {}
    for fused in (True, False):
        for graph in (False, True):
            e = engine_for(s, dt=U.time(1e-15), velocities=v0,
                           langevin=dict(friction=1.0 / U.time(1e-13), kt=kt, seed=5))
            e.use_graph(graph)
            e.set_fused_settle(fused)
            e.force_eval(scalars=False)
            e.step(6)
            e.step(6, scalars_last=True)
            p, img = e.positions(images=True)
            out[fused, graph] = (p, img, e.velocities(), e.forces(), e.langevin_counter(),
                                 e.scalars())
            e.close()
    ref = out[False, False]
    assert ref[4] == 12
    for key, o in out.items():
        for k in range(4):
            assert o[k].tobytes() == ref[k].tobytes(), f"{key} output {k}"
        assert o[4] == ref[4], key
        assert o[5] == ref[5], key
    say(f"  4 paths byte-identical, counter {ref[4]}, ke {ref[5]['ke']:.6e}")


def test_langevin_free_atoms_fused_and_unfused_paths_bit_identical():
    # BAOAB on argon without constraints: the fused half-step kernel against
    # the separate B + A and O + A kernels, eager and through captured step
    # graphs (a ragged atom count: the last block is partial)
    say("case: argon 10000 Langevin, fused vs unfused, eager and graph, 12 steps")
    s = systems.config_argon()
    kt = U.temperature(120.0)
    v0 = maxwell_boltzmann(s.mass, kt, seed=8)
    out = {}
    for fused in (True, False):
        for graph in (False, True):
            e = engine_for(s, dt=U.time(2e-15), velocities=v0,
                           langevin=dict(friction=1.0 / U.time(1e-13), kt=kt, seed=9))
            e.use_graph(graph)
            e.set_fused_settle(fused)
            e.force_eval(scalars=False)
            e.step(5)
            e.step(7, scalars_last=True)
            p, img = e.positions(images=True)
            out[fused, graph] = (p, img, e.velocities(), e.forces(), e.langevin_counter(),
                                 e.scalars())
            e.close()
    ref = out[False, False]
    assert ref[4] == 12
    for key, o in out.items():
        for k in range(4):
            assert o[k].tobytes() == ref[k].tobytes(), f"{key} output {k}"
        assert o[4] == ref[4], key
        assert o[5] == ref[5], key
    say(f"  4 paths byte-identical, counter {ref[4]}, ke {ref[5]['ke']:.6e}")


def test_settle_waters_beside_free_atoms(oracle):
    # the last two molecules are left out of the SETTLE groups (six free
    # atoms, their intramolecular pairs still excluded), so the step runs
    # the unfused kick/drift, SETTLE and velocity kernels: the waters follow
    # the FP64 oracle, the free atoms plain velocity Verlet
    say("case: water 8190 SETTLE + 6 free atoms")
    s = systems.config_water_8192()
    import dataclasses
    groups = s.water_groups[:-2].copy()
    s2 = dataclasses.replace(s, water_groups=groups)
    rng = np.random.default_rng(31)
    v0 = rng.normal(0, 2e-4, size=(3, s.n)).astype(np.float32)
    dt = np.float32(U.time(2e-15))
    posq0, img0, f0, posq1, img1, v1, f1 = _one_step(s2, v0, dt)
    p, img, pos, vel = _settle_reference_step(oracle, s, posq0, img0, v0, f0, f1, dt, groups)
    free = s.water_groups[-2:].reshape(-1)
    inw = groups.reshape(-1)
    # free atoms: the kick + drift is bit-exact; their closing kick in f32
    v32 = v0.copy()
    pf = posq0.copy()
    imf = img0.T.copy()
    oracle.vv_kick_drift_f32(pf, imf, v32, f0, s.mass, s.lattice.astype(np.float32), dt)
    oracle.vv_kick_f32(v32, f1, s.mass, dt)
    assert np.array_equal(posq1[free], pf[free])
    assert np.array_equal(v1[:, free], v32[:, free])
    dp = np.abs(posq1[inw, :3] - pos[inw])
    spacing = np.spacing(np.abs(pos[inw]).astype(np.float32)).astype(np.float64)
    say(f"  waters: max |dp| {dp.max():.2e} Bohr")
    assert (dp <= spacing).all()
    assert np.abs(v1.T[inw] - vel[inw]).max() <= 1e-5 * np.abs(vel).max()


def test_displacement_check_matches_the_rule(oracle):
    # the pre-step status bit 0 is set exactly when some atom's minimum-image
    # displacement from the last rebuild exceeds r_skin / 2
    # (neighbor.cu:50-75); displacements are kept 1e-3 away from the
    # threshold so f32 and FP64 agree on each side
    say("case: argon 4000 displacement check")
    s = systems.argon_lattice(n_atoms=4000, n_side=16)
    skin = 2.0
    e = engine_for(s, dt=U.time(1e-15), r_skin=skin)
    e.force_eval(scalars=False)
    base = s.posq.copy()
    L = s.lattice[:3]
    rng = np.random.default_rng(5)
    half = 0.5 * skin
    trips = 0
    for trial in range(12):
        e.set_positions(base)
        e.force_eval(scalars=False)  # a tripped previous trial rebuilds here
        assert e.neighbor_pre_step() & 1 == 0
        u = rng.normal(size=(s.n, 3))
        u /= np.linalg.norm(u, axis=1)[:, None]
        mag = rng.uniform(0.0, 0.98 * half, size=s.n)
        movers = rng.choice(s.n, size=3, replace=False)
        mag[movers] = rng.choice([0.97, 0.999, 1.001, 1.03, 1.5], size=3) * half
        p = base.copy()
        p[:, :3] = systems.wrap_into_box(base[:, :3].astype(np.float64) + u * mag[:, None],
                                         s.lattice).astype(np.float32)
        d = p[:, :3].astype(np.float64) - base[:, :3]
        d -= L * np.rint(d / L)
        want = bool((np.linalg.norm(d, axis=1) > half).any())
        e.set_positions(p)
        got = bool(e.neighbor_pre_step() & 1)
        say(f"  trial {trial}: max |d| / (r_skin/2) = {np.linalg.norm(d, axis=1).max() / half:.4f}"
            f" -> rebuild {got} (rule {want})")
        assert got == want
        trips += want
    e.close()
    assert 0 < trips < 12


def _inert_argon():
    s = systems.argon_lattice(n_atoms=4000, n_side=16)
    s.lj.epsilon[:] = 0.0  # no forces: every step is a straight flight
    return s


def test_list_grows_between_graph_launches():
    # an inert gas imploding at constant velocity towards the box centre
    # outgrows its neighbour capacity part-way through one 400-step
    # graph-replayed call; the status word read between graph launches
    # sees the high-water mark and the list grows before it overflows, so
    # every step runs on a complete list and the call succeeds
    say("case: argon 4000 inert, imploding: high water seen between graph launches")
    s = _inert_argon()
    dt = U.time(1e-15)
    x0 = s.posq[:, :3].astype(np.float64)
    v = (-0.45 * x0 / (400 * dt)).T.astype(np.float32)  # the cloud is centred at 0
    e = engine_for(s, dt=dt, velocities=v)
    e.force_eval(scalars=False)
    cap0 = e.neighbor_stats()["capacity"]
    e.step(400)
    st = e.neighbor_stats()
    p = e.positions()[:, :3].astype(np.float64)
    e.close()
    say(f"  capacity {cap0} -> {st['capacity']}, rebuilds {st['rebuilds']}, status {st['status']}")
    assert st["capacity"] > cap0
    assert st["status"] & 8 == 0
    # all 400 steps ran: each atom flew 400 v dt
    assert np.abs(p - (x0 + 400 * dt * v.T.astype(np.float64))).max() < 1e-3


def test_overflow_stops_a_long_call_within_a_few_graph_launches():
    # the same gas squeezed into half the box after set-up overflows at the
    # first rebuild of the next call; the call stops a few launches later
    # instead of running its 400 steps on a truncated list (a uniform drift
    # velocity measures the steps taken)
    say("case: argon 4000 inert, squeezed: overflow seen between graph launches")
    s = _inert_argon()
    dt = U.time(1e-15)
    vd = 1e-4
    v = np.full((3, s.n), vd, np.float32)
    e = engine_for(s, dt=dt, velocities=v)
    e.force_eval(scalars=False)
    p = s.posq.copy()
    p[:, :3] = (0.5 * p[:, :3].astype(np.float64)).astype(np.float32)
    e.set_positions(p)
    with pytest.raises(HeddleError, match="overflow"):
        e.step(400)
    q = e.positions()[:, :3].astype(np.float64)
    e.close()
    taken = np.median((q - p[:, :3]).ravel()) / (vd * dt)
    say(f"  stopped after {taken:.1f} of 400 steps")
    assert round(taken) <= 3 * 8


def test_graph_recapture_after_reparameterisation_matches_eager():
    # ten phases of 20 steps, each ended by a velocity redraw, a velocity
    # projection and a fresh Nose-Hoover chain (every one drops and
    # re-captures the step graphs): graph replay and plain launches end in
    # the same bytes
    say("case: water 8192, 10 re-parameterised phases, graphs vs eager")
    s = systems.config_water_8192()
    kt = U.temperature(300.0)
    tau = U.time(1e-13)
    finals = []
    for graphs in (True, False):
        v = maxwell_boltzmann(s.mass, kt, seed=7, n_constraints=s.n_constraints)
        e = engine_for(s, dt=U.time(1e-15), nhc_kt=kt, nhc_tau=tau, velocities=v)
        e.use_graph(graphs)
        e.constrain_velocities()
        for k in range(10):
            e.step(20, scalars_last=True)
            e.set_velocities(maxwell_boltzmann(s.mass, kt, seed=100 + k,
                                               n_constraints=s.n_constraints))
            e.constrain_velocities()
            e.set_nhc(kt, tau, n_constraints=s.n_constraints)
            if k == 4:
                e.set_integrator(U.time(2e-15))
        e.step(20, scalars_last=True)
        (p, i), v1, sc = e.positions(images=True), e.velocities(), e.scalars()
        e.close()
        finals.append((p, i, v1, sc))
        say(f"  graphs {graphs}: ke {sc['ke']:.9e} pe {sc['pe']:.9e}")
    (pa, ia, va, sa), (pb, ib, vb, sb) = finals
    assert np.isfinite(pa).all()
    assert pa.tobytes() == pb.tobytes()
    assert ia.tobytes() == ib.tobytes()
    assert va.tobytes() == vb.tobytes()
    assert sa == sb


@pytest.mark.parametrize("what", ["settle", "shake", "barostat"])
def test_lossless_mode_refuses_constraints_and_barostat(what):
    # the compensated mode carries FP64 residuals that constraint
    # projections and box scaling would leave stale; the reference refuses
    # constraints in it (src/integrator/velocity_verlet.rs:222-229)
    say(f"case: lossless + {what} refused")
    s = systems.spc_water(n_mol=512, n_side=8, grid=(32, 32, 32), r_cut_m=0.7e-9)
    if what == "shake":
        s = systems.water_as_shake(s)
    with pytest.raises(HeddleError, match="lossless"):
        kw = {}
        if what == "barostat":
            import dataclasses
            s = dataclasses.replace(s, water_groups=None)
            kw["crescale"] = dict(pressure=U.pressure(1e5), tau=U.time(1e-12),
                                  compressibility=U.inv_pressure(4.5e-10),
                                  kt=U.temperature(300.0), seed=1)
        e = engine_for(s, dt=U.time(1e-15), lossless=True, **kw)
        e.close()
