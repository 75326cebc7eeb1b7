"""The benchmark workloads at their full sizes under the GPU, through the C
ABI, against the CPU oracle:

  * 196,608-atom SPC water with a 128^3 SPME grid: forces, energies and
    virials of the FP32 radix FFT and of the split-precision matrix-core
    DFT;
  * the pair sets of the 8,192- and 65,536-molecule water boxes against the
    oracle's brute-force enumeration (every pair within r_search present);
  * the 256 batched replicas, each byte-identical to its own solo run, and
    eight of them, stepped as a batch, against the oracle's forces;
  * the 200 x 50 polymer melt: forces, energy, virial, and one SHAKE step
    bit-exact against the oracle's restatement of the device's sweeps.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from heddlemd_amd import systems, units as U  # noqa: E402
from heddlemd_amd.engine import engine_for  # noqa: E402
from tests.forcecheck import check_forces  # noqa: E402


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _keys(pairs):
    return (pairs[:, 0].astype(np.uint64) << np.uint64(32)) | pairs[:, 1].astype(np.uint64)


def test_water_196k_forces_energy_virial_both_transforms(oracle):
    say("case: water 65536 (196,608 atoms), SPME 128^3, FP32 FFT and FP16 split DFT")
    s = systems.config_water_65536()
    out = {}
    for mode in (0, 1):
        e = engine_for(s, dt=U.time(2e-15))
        e.set_spme_dft(mode)
        e.force_eval(scalars=True)
        out[mode] = (e.forces(), e.scalars())
        e.close()
    say("  oracle ...")
    Fp, Ep, Wp, band = oracle.pair_forces(s.posq, s.types, s.lj.as_dict(), s.lattice,
                                          coul=dict(alpha=s.spme.alpha, r_cut=s.spme.r_cut),
                                          excl=s.exclusion_csr(), band=1e-6)
    Fr, Er, Wr, er, wr = oracle.spme_recip(s.posq, s.lattice, s.spme.grid, 4, s.spme.alpha)
    for mode, name in ((0, "fp32 fft"), (1, "fp16 split dft")):
        f, sc = out[mode]
        say(f"  {name}: E_recip rel {abs(sc['e_recip'] / er - 1):.3e}, "
            f"W_recip rel {abs(sc['w_recip'] / wr - 1):.3e}, "
            f"pe rel {abs(sc['pe'] / (Ep + Er).sum() - 1):.3e}")
        check_forces(f, Fp + Fr, band, f"  196k {name}", say)
        assert sc["e_recip"] == pytest.approx(er, rel=2e-6)
        assert sc["w_recip"] == pytest.approx(wr, rel=2e-6)
        assert sc["pe"] == pytest.approx((Ep + Er).sum(), rel=2e-6)
        assert sc["virial"] == pytest.approx((Wp + Wr).sum(), rel=2e-6)


@pytest.mark.parametrize("which", ["8192", "65536"])
def test_water_pair_set_covers_the_oracle(oracle, which):
    # P_s = {pairs with r^2 <= r_search^2}: every oracle pair is in the
    # packed list; pairs within 4 ulp of r_search^2 are reported apart
    # (their side of the threshold depends on rounding order)
    s = systems.config_water_8192() if which == "8192" else systems.config_water_65536()
    say(f"case: water {which} pair set")
    e = engine_for(s, dt=U.time(2e-15))
    got = _keys(e.pairs())
    useful = e.count_pairs()
    e.close()
    rs = s.spme.r_cut * 1.3  # r_search = r_c + the default skin 0.3 r_c
    want, band = oracle.pair_set(s.posq, s.lattice, rs)
    wk = _keys(want)
    bk = _keys(band) if len(band) else np.zeros(0, np.uint64)
    missing = np.setdiff1d(wk, got, assume_unique=True)
    missing = np.setdiff1d(missing, bk)
    say(f"  list {len(got)} pairs, oracle P_s {len(wk)} (band {len(bk)}), missing {len(missing)}")
    assert len(missing) == 0
    assert np.all(np.diff(got.astype(np.uint64)) > 0)  # canonical, unique
    within_rc, band_rc = oracle.pair_set(s.posq, s.lattice, s.spme.r_cut)
    # useful pairs: within r_c and not excluded (the 3 intramolecular pairs
    # of each water are); pairs within 4 ulp of r_c^2 may fall either side
    n_intra = s.n
    say(f"  useful {useful}, oracle within r_c {len(within_rc)} (minus {n_intra} "
        f"intramolecular), band at r_c {len(band_rc)}")
    assert abs(useful - (len(within_rc) - n_intra)) <= len(band_rc)


def _replica_engine(spec, r):
    from heddlemd_amd.rng import maxwell_boltzmann
    s = systems.replica_system(spec)
    kt = U.temperature(spec["temperature_K"])
    v = maxwell_boltzmann(s.mass, kt, seed=100 + r, n_constraints=s.n_constraints)
    e = engine_for(s, dt=U.time(1e-15), velocities=v,
                   langevin=dict(friction=1.0 / U.time(1e-13), kt=kt, seed=1000 + r))
    e.constrain_velocities()
    return e


def test_all_256_replicas_byte_identical_to_solo_runs():
    say("case: 256 replicas, batched vs solo, 20 steps")
    from heddlemd_amd.engine import Batch
    specs = list(systems.replica_specs(count=256))
    solo = []
    for r, sp in enumerate(specs):
        e = _replica_engine(sp, r)
        e.step(20, scalars_last=True)
        solo.append((e.positions(images=True), e.velocities(), e.scalars()))
        e.close()
    say("  solo runs done")
    engines = [_replica_engine(sp, r) for r, sp in enumerate(specs)]
    b = Batch(engines)
    b.step(8)
    b.step(12, scalars_last=True)
    b.close()
    bad = []
    for r, e in enumerate(engines):
        (p, i), v, sc = e.positions(images=True), e.velocities(), e.scalars()
        (ps, is_), vs, scs = solo[r]
        if not (np.isfinite(p).all() and p.tobytes() == ps.tobytes()
                and i.tobytes() == is_.tobytes() and v.tobytes() == vs.tobytes() and sc == scs
                and e.langevin_counter() == 20):
            bad.append(r)
        e.close()
    say(f"  {len(specs)} replicas, {sum(sp['n_atoms'] for sp in specs)} atoms, mismatches {bad}")
    assert bad == []


def test_replica_forces_in_a_batch_match_the_oracle(oracle):
    # eight of the 256 replicas (the smallest, the largest and two between of
    # each kind; the water grids include a mixed-radix side) stepped 200
    # times (0.2 ps) as one batch; every replica's forces at its own final
    # positions against the FP64 oracle, to the force bar of the single
    # systems.  (The argon replicas start on a perfect lattice, where every
    # atom's force is a near-cancelling sum and F_rms is tiny; 0.2 ps of
    # Langevin dynamics leaves a disordered state.)
    specs = list(systems.replica_specs(count=256))
    pick = []
    for kind in ("argon", "water"):
        idx = sorted((r for r, sp in enumerate(specs) if sp["kind"] == kind),
                     key=lambda r: specs[r]["n_atoms"])
        pick += [idx[0], idx[len(idx) // 3], idx[2 * len(idx) // 3], idx[-1]]
    say("case: replicas " + ", ".join(f"{specs[r]['kind']} {specs[r]['n_atoms']}" for r in pick)
        + " batched 200 steps, forces vs oracle")
    from heddlemd_amd.engine import Batch
    engines = [_replica_engine(specs[r], r) for r in pick]
    b = Batch(engines)
    b.step(200)
    b.close()
    for r, e in zip(pick, engines):
        s = systems.replica_system(specs[r])
        pos = e.positions()
        f = e.forces()
        e.close()
        if specs[r]["kind"] == "argon":
            F, _, _, band = oracle.pair_forces(pos, s.types, s.lj.as_dict(), s.lattice, band=1e-6)
        else:
            F, _, _, band = oracle.pair_forces(pos, s.types, s.lj.as_dict(), s.lattice,
                                               coul=dict(alpha=s.spme.alpha, r_cut=s.spme.r_cut),
                                               excl=s.exclusion_csr(), band=1e-6)
            Fr = oracle.spme_recip(pos, s.lattice, s.spme.grid, 4, s.spme.alpha)[0]
            F = F + Fr
        check_forces(f, F, band, f"replica {r} ({specs[r]['kind']} {specs[r]['n_atoms']}, "
                     f"grid {s.spme.grid[0] if s.spme else '-'})", say)


def test_polymer_melt_200x50_forces_and_shake_step(oracle):
    say("case: polymer melt 200 x 50: forces and one SHAKE step")
    s = systems.config_polymer()
    kt = U.temperature(400.0)
    from heddlemd_amd.rng import maxwell_boltzmann
    v0 = maxwell_boltzmann(s.mass, kt, seed=5, n_constraints=s.n_constraints)
    dt = np.float32(U.time(1e-15))
    e = engine_for(s, dt=float(dt), velocities=v0)
    e.set_shake_tolerance(5e-5, 1e-8, 500)
    e.use_graph(False)
    e.constrain_velocities()
    v0 = e.velocities()
    e.force_eval(scalars=True)
    f0 = e.forces()
    sc = e.scalars()
    posq0, img0 = e.positions(images=True)
    e.step(1)
    posq1, img1 = e.positions(images=True)
    v1 = e.velocities()
    f1 = e.forces().T.copy()
    st = e.constraint_stats()
    e.close()
    Fp, Ep, Wp, band = oracle.pair_forces(s.posq, s.types, s.lj.as_dict(), s.lattice,
                                          excl=s.exclusion_csr(), band=1e-6)
    Fa, Ea, Wa = oracle.angle_forces(s.posq, s.lattice, *s.angles)
    check_forces(f0, Fp + Fa, band, "  melt 200x50", say)
    assert sc["pe"] == pytest.approx((Ep + Ea).sum(), rel=1e-5)
    assert sc["virial"] == pytest.approx((Wp + Wa).sum(), rel=1e-5, abs=1e-6 * np.abs(Wp).sum())
    G = oracle.ShakeGroups(s.shake_groups)
    p = posq0.copy()
    img = img0.T.copy()
    v = v0.copy()
    L32 = s.lattice.astype(np.float32)
    snap = p.copy()
    oracle.vv_kick_drift_f32(p, img, v, f0.T.copy(), s.mass, L32, dt)
    oracle.set_shake_form(True)
    try:
        wc, it_s = oracle.shake_positions(G, p, v, snap, s.mass, L32, dt, tol=5e-5, max_iter=500)
        it_r = oracle.rattle_velocities(G, p, v, s.mass, L32, dt, w_cons=wc, tol=1e-8,
                                        max_iter=500)
    finally:
        oracle.set_shake_form(False)
    oracle.vv_kick_f32(v, f1, s.mass, dt)
    say(f"  sweeps: shake max {st['shake_max']} (oracle {it_s.max()}), capped "
        f"{st['shake_capped']}; rattle max {st['rattle_max']} (oracle {it_r.max()})")
    assert st["shake_capped"] == 0 and st["rattle_capped"] == 0
    assert st["shake_sweeps"] == int(it_s.sum())
    assert st["rattle_sweeps"] == int(it_r.sum())
    assert np.array_equal(img.T, img1)
    assert np.array_equal(p, posq1)
    assert np.array_equal(v, v1)


def test_water_8192_compensated_mode():
    # the 8,192-molecule SPC box in the compensated (f32 value + f32 residual
    # per coordinate) mode.  The reference refuses constraints in that mode
    # (src/integrator/velocity_verlet.rs:222-231), so the waters are held by
    # harmonic O-H bonds and an H-O-H angle, as in its spc-water-256
    # example.  (1) NVE with SPME is time-reversible: 20 steps, negate the
    # velocities, 20 steps, and every velocity word and every position word
    # not within a rounding of zero is back.  (2) Under the Nose-Hoover chain at 300 K the compensated and
    # plain f32 runs agree after 10 steps, and the chain's conserved
    # quantity moves by under 1 % of the kinetic energy over 200 steps in
    # both.
    s = systems.flexible_water(systems.config_water_8192())
    from heddlemd_amd.rng import maxwell_boltzmann
    kt = U.temperature(300.0)
    v0 = maxwell_boltzmann(s.mass, kt, seed=4, n_constraints=0).astype(np.float32)
    dt = U.time(5e-16)
    say("case: flexible water 8192 lossless reversal, 20 + 20 steps")
    e = engine_for(s, dt=dt, lossless=True, velocities=v0, r_skin=U.length(0.3e-9))
    e.force_eval(scalars=True)
    x0 = e.positions()
    reb0 = e.scalars()["rebuilds"]
    e.step(20)
    e.negate_velocities()
    e.step(20)
    x1 = e.positions()
    v1 = e.velocities()
    reb = e.scalars()["rebuilds"] - reb0
    e.close()
    # the (f32 + FP64 residual) sums are exact, and so reversible, except
    # for a coordinate that starts at or passes within a double rounding of
    # zero: there the FP64 residual of x + dx loses bits below 2^-53 |dx|
    # (the lattice puts 1,230 oxygen coordinates at exactly 0), and the
    # coordinate comes back as a few 1e-18 Bohr
    far = np.abs(x0[:, :3]) > 1e-9
    near = ~far
    say(f"  rebuilds between the passes {reb}; coordinates equal "
        f"{np.mean(x0[:, :3] == x1[:, :3]):.6f} ({near.sum()} within 1e-9 Bohr of zero, max "
        f"|x1 - x0| there {np.abs(x1[:, :3][near]).max():.1e}); velocities equal "
        f"{np.mean(-v1 == v0):.6f}")
    assert reb == 0, "a rebuild between the passes would change the f32 sum order"
    assert np.array_equal(x0[:, :3][far], x1[:, :3][far])
    assert np.abs(x1[:, :3][near].astype(np.float64) - x0[:, :3][near]).max() < 1e-15
    assert np.array_equal(x0[:, 3], x1[:, 3])
    assert np.array_equal(-v1, v0)

    say("case: flexible water 8192 NHC 300 K, compensated vs f32, 200 steps")
    runs = {}
    for lossless in (True, False):
        e = engine_for(s, dt=dt, lossless=lossless, velocities=v0, nhc_kt=kt,
                       nhc_tau=U.time(1e-13))
        e.force_eval(scalars=True)
        e.step(1, scalars_last=True)
        a = e.scalars()
        e.step(9)
        p10 = e.positions()
        e.step(190, scalars_last=True)
        b = e.scalars()
        e.close()
        h0 = a["ke"] + a["pe"] + a["nhc_chain"]
        h1 = b["ke"] + b["pe"] + b["nhc_chain"]
        runs[lossless] = (p10, abs(h1 - h0) / a["ke"], 2 * b["ke"] / (3 * s.n - 3) / kt)
    dp = np.abs(runs[True][0][:, :3].astype(np.float64) - runs[False][0][:, :3]).max()
    say(f"  after 10 steps max |dx| {dp:.2e} Bohr; conserved drift / KE: compensated "
        f"{runs[True][1]:.2e}, f32 {runs[False][1]:.2e}; T/T0 {runs[True][2]:.3f}")
    assert dp < 1e-3
    assert runs[True][1] < 1e-2 and runs[False][1] < 1e-2
    # (the unrelaxed lattice start releases its strain into heat faster
    # than the 100 fs chain removes it: T/T0 is about 2.6 at 100 fs)
    assert np.isfinite(runs[True][2]) and runs[True][2] < 5.0
