"""Pins the oracle's SPME reciprocal space (src/forces/spme.rs:403-585,
kernels/spme_recip.cu, kernels/spme_spread_gather.cuh) against closed forms:
cardinal B-splines, numpy's FFT, charge conservation, the explicit Ewald
sum (the reference's own fixture, tests/spme_pipeline.rs:102-145), the
finite-difference gradient of the SPME energy and the volume derivative
that defines the virial."""
import json
import math
import os

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def explicit_ewald(pos, q, L, alpha, kmax):
    """U_recip = (1/2V) sum_{k!=0} (4 pi/k^2) exp(-k^2/4a^2) |S(k)|^2."""
    V = L[0] * L[1] * L[2]
    m = np.arange(-kmax, kmax + 1)
    ma, mb, mc = np.meshgrid(m, m, m, indexing="ij")
    k = 2 * np.pi * np.stack([ma / L[0], mb / L[1], mc / L[2]], -1).reshape(-1, 3)
    k2 = (k ** 2).sum(1)
    nz = k2 > 0
    k, k2 = k[nz], k2[nz]
    S = (q[None, :] * np.exp(1j * k @ pos.T)).sum(1)
    return float((4 * np.pi / k2 * np.exp(-k2 / (4 * alpha ** 2)) * np.abs(S) ** 2).sum() / (2 * V))


def test_cardinal_bspline_values(oracle):
    # M_4 at the integers: 1/6, 2/3, 1/6 (Essmann et al. 1995)
    assert oracle.bspline(4, 1.0) == pytest.approx(1 / 6)
    assert oracle.bspline(4, 2.0) == pytest.approx(2 / 3)
    assert oracle.bspline(4, 3.0) == pytest.approx(1 / 6)
    assert oracle.bspline(4, 0.0) == 0.0 and oracle.bspline(4, 4.0) == 0.0
    # partition of unity and symmetry M_p(x) = M_p(p - x)
    for p in (3, 4, 5, 6):
        for t in np.linspace(0, 1, 11, endpoint=False):
            assert sum(oracle.bspline(p, t + k) for k in range(p)) == pytest.approx(1.0, abs=1e-14)
            assert oracle.bspline(p, t + 1.3) == pytest.approx(oracle.bspline(p, p - t - 1.3))


def test_b_factors_closed_form(oracle):
    # |b(k)|^2 = 1/|sum_j M_p(j+1) e^{2 pi i jk/n}|^2 ; at k = 0 the sum is 1
    b = oracle.b_factors(32, 4)
    assert b[0] == pytest.approx(1.0)
    k = 16  # Nyquist for p = 4: 1/6 - 2/3 + 1/6 = -1/3 -> 9
    assert b[k] == pytest.approx(9.0)
    assert np.all(b > 0)


@pytest.mark.parametrize("shape", [(8, 8, 8), (16, 8, 32), (6, 10, 4), (12, 20, 30)])
def test_fft3d_matches_numpy(oracle, shape):
    rng = np.random.default_rng(0)
    x = rng.normal(size=shape)
    re, im = oracle.fft3d(x, np.zeros(shape), -1)
    ref = np.fft.fftn(x)
    assert np.allclose(re, ref.real, atol=1e-10) and np.allclose(im, ref.imag, atol=1e-10)
    # inverse is unnormalised: round trip = M * x
    r2, i2 = oracle.fft3d(re, im, +1)
    assert np.allclose(r2, x * np.prod(shape), atol=1e-8)


def test_spread_conserves_charge_and_locality(oracle):
    L = np.array([20.0, 20.0, 20.0, 0, 0, 0])
    posq = np.array([[1.3, -2.2, 0.7, 0.5], [-7.9, 9.9, -9.99, -1.25]], np.float32)
    rho = oracle.spme_spread(posq, L, (16, 16, 16), 4)
    assert rho.sum() == pytest.approx(float(posq[:, 3].sum()), abs=1e-12)
    single = oracle.spme_spread(posq[:1], L, (16, 16, 16), 4)
    assert np.count_nonzero(single) == 64  # p^3 support


def _golden(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return json.load(f)


def run_ewald_case(oracle, g, order=4):
    pos = np.array(g["pos_bohr"])
    q = np.array(g["q"])
    Lb = g["box_bohr"]
    L = np.array([Lb, Lb, Lb, 0, 0, 0])
    posq = np.zeros((len(q), 4), np.float32)
    posq[:, :3] = pos
    posq[:, 3] = q
    F, E, W, er, wr = oracle.spme_recip(posq, L, tuple(g["grid"]), order, g["alpha"])
    ref = explicit_ewald(posq[:, :3].astype(np.float64), posq[:, 3].astype(np.float64),
                         L[:3], g["alpha"], g["kmax"])
    assert ref == pytest.approx(g["e_recip_explicit"], rel=1e-9)
    assert abs(er - ref) / abs(ref) < g["rel_tol"]
    # per-atom energies add up to E_recip minus the self term
    self_e = g["alpha"] / math.sqrt(math.pi) * float((posq[:, 3].astype(np.float64) ** 2).sum())
    assert E.sum() == pytest.approx(er - self_e, rel=1e-10, abs=1e-12)


@pytest.mark.parametrize("case", ["two_charges", "four_charges", "four_charges_fine"])
def test_ewald_fixtures(oracle, case):
    run_ewald_case(oracle, _golden("spme_ewald_fixtures.json")[case])


@pytest.mark.parametrize("case", ["two_charges", "four_charges", "four_charges_fine"])
def test_ewald_fixtures_order5(oracle, case):
    # the order-5 interpolation (the reference's chignolin example) meets
    # the same explicit-Ewald bar, and is at least as close as order 4
    g = _golden("spme_ewald_fixtures.json")[case]
    run_ewald_case(oracle, g, order=5)
    posq = np.zeros((len(g["q"]), 4), np.float32)
    posq[:, :3] = np.array(g["pos_bohr"])
    posq[:, 3] = np.array(g["q"])
    L = np.array([g["box_bohr"]] * 3 + [0, 0, 0])
    ref = g["e_recip_explicit"]
    e4 = oracle.spme_recip(posq, L, tuple(g["grid"]), 4, g["alpha"])[3]
    e5 = oracle.spme_recip(posq, L, tuple(g["grid"]), 5, g["alpha"])[3]
    assert abs(e5 - ref) <= abs(e4 - ref) * 1.05 + 1e-15


def test_b_factors_order5_finite(oracle):
    # for odd p the modulus vanishes at k = n/2 (n even): M_5(1..4) = 1/24,
    # 11/24, 11/24, 1/24 alternate to 0 there; the modulus is then the mean
    # of its neighbours' (Essmann et al. 1995), equal by symmetry, so every
    # factor is finite
    for n in (20, 32, 48):
        b = oracle.b_factors(n, 5)
        assert b[0] == pytest.approx(1.0)
        assert np.all(np.isfinite(b)) and np.all(b > 0)
        if n % 2 == 0:
            assert b[n // 2] == pytest.approx(0.5 * (b[n // 2 - 1] + b[n // 2 + 1]), rel=1e-12)


def _water_like(n=60, L=18.0, seed=4):
    rng = np.random.default_rng(seed)
    posq = np.zeros((n, 4), np.float32)
    posq[:, :3] = rng.uniform(-L / 2, L / 2, size=(n, 3))
    posq[:, 3] = np.tile([-0.82, 0.41, 0.41], n // 3)
    return posq, np.array([L, L, L, 0, 0, 0])


@pytest.mark.parametrize("order", [4, 5])
def test_recip_force_is_energy_gradient(oracle, order):
    posq, L = _water_like()
    grid, alpha = (32, 32, 32), 0.35
    F, E, W, er, wr = oracle.spme_recip(posq, L, grid, order, alpha)
    h = 1.0 / 1024.0
    for atom in (0, 7, 31):
        for ax in range(3):
            p1 = posq.copy(); p1[atom, ax] += h
            p0 = posq.copy(); p0[atom, ax] -= h
            e1 = oracle.spme_recip(p1, L, grid, order, alpha)[3]
            e0 = oracle.spme_recip(p0, L, grid, order, alpha)[3]
            fd = -(e1 - e0) / (float(p1[atom, ax]) - float(p0[atom, ax]))
            assert F[atom, ax] == pytest.approx(fd, rel=2e-4, abs=2e-7)


def test_recip_virial_is_volume_derivative(oracle):
    # W_recip = -3 V dE/dV at fixed fractional coordinates (trace of the
    # Ewald virial tensor, factor (1 - k^2/2a^2) in spme_recip.cu:34-115)
    posq, L = _water_like(seed=9)
    grid, alpha = (32, 32, 32), 0.35
    F, E, W, er, wr = oracle.spme_recip(posq, L, grid, 4, alpha)
    lam = 1e-3

    def energy(s):
        p = posq.copy()
        p[:, :3] = (posq[:, :3].astype(np.float64) * s).astype(np.float32)
        Ls = L.copy()
        Ls[:3] *= s
        return oracle.spme_recip(p, Ls, grid, 4, alpha)[3]

    dE_dlam = (energy(1 + lam) - energy(1 - lam)) / (2 * lam)
    # V = lam^3 V0 -> dE/dV = dE/dlam / (3 V)
    assert wr == pytest.approx(-dE_dlam, rel=2e-3)
    assert np.allclose(W, wr / posq.shape[0])
