"""Tilted (triclinic) cells on the GPU, through the C ABI, against the FP64
oracle (whose minimum image and reciprocal rows take the tilt,
oracle.c or_min_image / recip_rows): the pair set of a sheared argon
lattice, its LJ forces, sheared SPC water with SPME (forces, E_recip,
W_recip, energy), and SETTLE water stepping in a sheared cell (rigid,
rerun byte-identical)."""
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


def test_tilted_argon_pair_set_and_forces(oracle):
    # the jittered lattice of the orthorhombic argon test (a perfect lattice
    # cancels most net forces, which measures only summation rounding)
    say("case: argon 10000 (jittered) in a cell tilted (0.2, 0.1, 0.15)")
    from tests.test_gpu_parity import _jittered_argon
    s = systems.tilted(_jittered_argon(), 0.2, 0.1, 0.15)
    e = engine_for(s, dt=U.time(5e-15))
    e.force_eval(scalars=True)
    f, sc = e.forces(), e.scalars()
    got = _keys(e.pairs())
    rs = float(s.lj.cutoff.max()) * 1.3
    e.close()
    want, band = oracle.pair_set(s.posq, s.lattice, rs)
    missing = np.setdiff1d(_keys(want), got)
    if len(band):
        missing = np.setdiff1d(missing, _keys(band))
    say(f"  list {len(got)} pairs, oracle P_s {len(want)}, missing {len(missing)}")
    assert len(missing) == 0
    F, E, W, fb = oracle.pair_forces(s.posq, s.types, s.lj.as_dict(), s.lattice, band=1e-6)
    check_forces(f, F, fb, "  tilted argon", say)
    assert sc["pe"] == pytest.approx(E.sum(), rel=1e-5)
    assert sc["virial"] == pytest.approx(W.sum(), rel=1e-5)


def test_tilted_water_spme(oracle):
    say("case: water 8192 in a cell tilted (0.15, -0.1, 0.12), SPME 64^3")
    s = systems.tilted(systems.config_water_8192(), 0.15, -0.1, 0.12)
    e = engine_for(s, dt=U.time(2e-15))
    e.force_eval(scalars=True)
    f, sc = e.forces(), e.scalars()
    e.close()
    Fp, Ep, Wp, band = oracle.pair_forces(s.posq, s.types, s.lj.as_dict(), s.lattice,
                                          coul=dict(alpha=s.spme.alpha, r_cut=s.spme.r_cut),
                                          excl=s.exclusion_csr(), band=1e-6)
    Fr, Er, Wr, er, wr = oracle.spme_recip(s.posq, s.lattice, s.spme.grid, 4, s.spme.alpha)
    say(f"  E_recip {sc['e_recip']:.9e} (oracle {er:.9e}), W_recip {sc['w_recip']:.9e} "
        f"(oracle {wr:.9e})")
    check_forces(f, Fp + Fr, band, "  tilted water", say)
    assert sc["e_recip"] == pytest.approx(er, rel=2e-6)
    assert sc["w_recip"] == pytest.approx(wr, rel=2e-6)
    assert sc["pe"] == pytest.approx((Ep + Er).sum(), rel=2e-6)


def test_tilted_settle_water_steps_rigid_and_repeatable():
    say("case: water 8192 SETTLE in a tilted cell, 100 steps twice")
    s = systems.tilted(systems.config_water_8192(), 0.1, 0.05, -0.08)
    from heddlemd_amd.rng import maxwell_boltzmann
    v = maxwell_boltzmann(s.mass, U.temperature(300.0), seed=8, n_constraints=s.n_constraints)
    runs = []
    for _ in range(2):
        e = engine_for(s, dt=U.time(1e-15), velocities=v, nhc_kt=U.temperature(300.0),
                       nhc_tau=U.time(1e-13))
        e.constrain_velocities()
        e.force_eval(scalars=False)
        e.step(100, scalars_last=True)
        runs.append((e.positions(images=True), e.velocities(), e.scalars()))
        e.close()
    (p, img), vel, sc = runs[0]
    assert p.tobytes() == runs[1][0][0].tobytes() and vel.tobytes() == runs[1][1].tobytes()
    g = s.water_groups
    L = s.lattice
    H = np.array([[L[0], L[3], L[4]], [0, L[1], L[5]], [0, 0, L[2]]])
    Hi = np.linalg.inv(H)
    d = p[g[:, 0], :3].astype(np.float64) - p[g[:, 1], :3]
    fr = d @ Hi.T
    fr -= np.floor(fr + 0.5)
    r_oh = np.linalg.norm(fr @ H.T, axis=1)
    d_oh = s.water_geom[0]
    T = 2 * sc["ke"] / (3 * s.n - s.n_constraints - 3) * U.TEMPERATURE_AU_TO_K
    say(f"  O-H {r_oh.min() / d_oh - 1:.2e} .. {r_oh.max() / d_oh - 1:.2e}, T {T:.1f} K")
    assert np.abs(r_oh / d_oh - 1).max() < 5e-6
    # the lattice start releases its strain energy as heat within the
    # first steps (no minimization here), so only finiteness is asked
    assert np.isfinite(T) and T < 3000.0
