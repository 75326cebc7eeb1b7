"""SPME on grids whose sides are not powers of two (the mixed-radix LDS
FFT: radix 4, 2, 3 and 5 stages), against the FP64 oracle on the same grid:
the 8,192-molecule water box on 48 x 40 x 96 (radix 3 along a, 5 along b,
the c axis packed to 48 complex points) and on 20 x 80 x 30 (radix 5 along
a, an odd 15-point packed c axis).  The reference sizes its example grids
this way for cuFFT (48 x 48 x 96, 20 x 20 x 80).  Also the order-5 B-spline
interpolation (the reference's chignolin example) on 64^3 and 40^3, with
both reciprocal transforms where they apply."""
import dataclasses
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


@pytest.mark.parametrize("grid", [(48, 40, 96), (20, 80, 30)])
def test_water_forces_on_mixed_radix_grids(oracle, grid):
    say(f"case: water 8192, SPME grid {grid}")
    s = systems.config_water_8192()
    s = dataclasses.replace(s, spme=dataclasses.replace(s.spme, grid=grid))
    e = engine_for(s, dt=U.time(2e-15))
    e.force_eval(scalars=True)
    f = e.forces()
    sc = e.scalars()
    rho_g, v_g = e.spme_grids(grid)
    e.close()
    Fp, Ep, Wp, band = oracle.pair_forces(s.posq, s.types, s.lj.as_dict(), s.lattice,
                                          coul=dict(alpha=s.spme.alpha, r_cut=s.spme.r_cut),
                                          excl=s.exclusion_csr(), band=1e-6)
    Fr, Er, Wr, er, wr, V = oracle.spme_recip(s.posq, s.lattice, grid, 4, s.spme.alpha,
                                              want_v=True)
    v_err = float(np.abs(v_g - V).max() / np.abs(V).max())
    say(f"  V err {v_err:.2e}, E_recip rel {abs(sc['e_recip'] / er - 1):.2e}, "
        f"W_recip rel {abs(sc['w_recip'] / wr - 1):.2e}")
    assert v_err <= 2e-5
    assert sc["e_recip"] == pytest.approx(er, rel=2e-6)
    assert sc["w_recip"] == pytest.approx(wr, rel=2e-6)
    check_forces(f, Fp + Fr, band, f"  grid {grid}", say)
    assert sc["pe"] == pytest.approx((Ep + Er).sum(), rel=2e-6)


def test_mixed_radix_grid_runs_steps_deterministically():
    # 20 graph-replayed steps on 48 x 48 x 96, twice: byte-identical
    say("case: water 8192 on 48 x 48 x 96, 20 steps twice")
    s = systems.config_water_8192()
    s = dataclasses.replace(s, spme=dataclasses.replace(s.spme, grid=(48, 48, 96)))
    out = []
    for _ in range(2):
        from heddlemd_amd.rng import maxwell_boltzmann
        kt = U.temperature(300.0)
        v = maxwell_boltzmann(s.mass, kt, seed=3, n_constraints=s.n_constraints)
        e = engine_for(s, dt=U.time(1e-15), velocities=v, nhc_kt=kt, nhc_tau=U.time(1e-13))
        e.constrain_velocities()
        e.step(20, scalars_last=True)
        out.append((e.positions(), e.velocities(), e.scalars()))
        e.close()
    assert np.isfinite(out[0][0]).all()
    assert out[0][0].tobytes() == out[1][0].tobytes()
    assert out[0][1].tobytes() == out[1][1].tobytes()
    assert out[0][2] == out[1][2]


@pytest.mark.parametrize("grid,dft", [((64, 64, 64), 0), ((64, 64, 64), 1), ((40, 40, 40), 0)])
def test_water_forces_spline_order_5(oracle, grid, dft):
    say(f"case: water 8192, SPME order 5, grid {grid}, {'split DFT' if dft else 'FFT'}")
    s = systems.config_water_8192()
    s = dataclasses.replace(s, spme=dataclasses.replace(s.spme, grid=grid, order=5))
    e = engine_for(s, dt=U.time(2e-15))
    e.set_spme_dft(dft)
    e.force_eval(scalars=True)
    f = e.forces()
    sc = e.scalars()
    rho_g, v_g = e.spme_grids(grid)
    e.close()
    Fp, Ep, Wp, band = oracle.pair_forces(s.posq, s.types, s.lj.as_dict(), s.lattice,
                                          coul=dict(alpha=s.spme.alpha, r_cut=s.spme.r_cut),
                                          excl=s.exclusion_csr(), band=1e-6)
    Fr, Er, Wr, er, wr, V = oracle.spme_recip(s.posq, s.lattice, grid, 5, s.spme.alpha,
                                              want_v=True)
    rho_o = oracle.spme_spread(s.posq, s.lattice, grid, 5)
    rho_err = float(np.abs(rho_g - rho_o).max() / np.abs(rho_o).max())
    v_err = float(np.abs(v_g - V).max() / np.abs(V).max())
    say(f"  rho err {rho_err:.2e}, V err {v_err:.2e}, E_recip rel {abs(sc['e_recip'] / er - 1):.2e}, "
        f"W_recip rel {abs(sc['w_recip'] / wr - 1):.2e}")
    assert rho_err <= 4e-6
    assert v_err <= 2e-5
    assert sc["e_recip"] == pytest.approx(er, rel=2e-6)
    assert sc["w_recip"] == pytest.approx(wr, rel=2e-6)
    check_forces(f, Fp + Fr, band, f"  order 5 grid {grid}", say)
    assert sc["pe"] == pytest.approx((Ep + Er).sum(), rel=2e-6)


@pytest.mark.parametrize("grid", [(32, 32, 32), (40, 40, 40), (48, 48, 48), (64, 64, 64),
                                  (40, 48, 64)])
def test_plane_fused_fft_matches_separate_passes(grid):
    # where an a-plane fits in LDS the c and b passes run per plane in one
    # kernel each way; mode 2 runs every pass as its own launch.  Each line
    # gets the same transform of the same values, but the two kernels may
    # round a few multiply-adds differently, so the bar is f32 rounding:
    # forces within 1e-6 F_rms per atom, E_recip and W_recip within 1e-7
    # relative, the potential grid within 1e-6 of its largest value
    say(f"case: water 8192, SPME grid {grid}, plane-fused FFT vs separate passes")
    s = systems.config_water_8192()
    s = dataclasses.replace(s, spme=dataclasses.replace(s.spme, grid=grid))
    out = []
    for mode in (0, 2):
        e = engine_for(s, dt=U.time(2e-15))
        e.set_spme_dft(mode)
        e.force_eval(scalars=True)
        out.append((e.forces().astype(np.float64), e.scalars(),
                    e.spme_grids(grid)[1].astype(np.float64)))
        e.close()
    (fa, sa, va), (fb, sb, vb) = out
    f_rms = np.sqrt((fb ** 2).sum(axis=0).mean())
    df = np.sqrt(((fa - fb) ** 2).sum(axis=0)).max() / f_rms
    dv = np.abs(va - vb).max() / np.abs(vb).max()
    de = abs(sa["e_recip"] / sb["e_recip"] - 1)
    dw = abs(sa["w_recip"] / sb["w_recip"] - 1)
    say(f"  max |dF|/F_rms {df:.2e}, grid {dv:.2e}, E_recip {de:.2e}, W_recip {dw:.2e}")
    assert df <= 1e-6
    assert dv <= 1e-6
    assert de <= 1e-7 and dw <= 1e-7
