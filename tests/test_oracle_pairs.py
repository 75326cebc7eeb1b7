"""Pins the oracle's periodic boundaries, pair set and pair forces against
closed forms and brute force (reference behaviour: kernels/pbc.cuh,
src/forces/lj.rs:225-335, src/forces/spme.rs:708-783,
src/forces/neighbor_list.rs:1366-1506)."""
import math

import numpy as np
import pytest

from heddlemd_amd import systems


def _lj(ntypes=1, sigma=3.0, eps=0.01, cut=9.0, rsw=None):
    s = np.full((ntypes, ntypes), sigma)
    e = np.full((ntypes, ntypes), eps)
    c = np.full((ntypes, ntypes), cut)
    r = np.full((ntypes, ntypes), cut if rsw is None else rsw)
    return dict(ntypes=ntypes, sigma=s, epsilon=e, cutoff=c, r_switch=r)


def _two(r, L=40.0, q=(0.0, 0.0), axis=0):
    posq = np.zeros((2, 4), np.float32)
    posq[1, axis] = r
    posq[:, 3] = q
    return posq, np.array([L, L, L, 0, 0, 0], float)


@pytest.mark.parametrize("r", [2.9, 3.37, 4.5, 8.5])
def test_lj_two_body_closed_form(oracle, r):
    posq, L = _two(r)
    sig, eps = 3.0, 0.01
    F, E, W, _ = oracle.pair_forces(posq, np.zeros(2, np.uint32), _lj(sigma=sig, eps=eps), L)
    r = float(posq[1, 0])
    sr6 = (sig / r) ** 6
    e = 4 * eps * (sr6 * sr6 - sr6)
    fmag = 24 * eps / r * (2 * sr6 * sr6 - sr6)  # -dU/dr
    assert E.sum() == pytest.approx(e, rel=1e-12)
    assert E[0] == pytest.approx(E[1])
    # force on atom 1 points along +x when repulsive
    assert F[1, 0] == pytest.approx(fmag, rel=1e-12)
    assert F[0, 0] == pytest.approx(-fmag, rel=1e-12)
    # virial = r * F (sum over pairs of r_ij . F_ij)
    assert W.sum() == pytest.approx(fmag * r, rel=1e-12)


def test_lj_zero_epsilon_and_cutoff_are_inert(oracle):
    posq, L = _two(3.5)
    F, E, W, _ = oracle.pair_forces(posq, np.zeros(2, np.uint32), _lj(eps=0.0), L)
    assert np.all(F == 0) and np.all(E == 0)
    posq, L = _two(9.5)
    F, E, W, _ = oracle.pair_forces(posq, np.zeros(2, np.uint32), _lj(cut=9.0), L)
    assert np.all(F == 0) and np.all(E == 0)


def test_lj_switch_is_continuous_and_consistent(oracle):
    # cubic switch on r^2 between r_s and r_c: energy -> 0 at r_c, force = -dE/dr
    lj = _lj(sigma=3.0, eps=0.01, cut=9.0, rsw=7.0)
    zero = np.zeros(2, np.uint32)

    def energy(r):
        posq, L = _two(r)
        posq = posq.astype(np.float64)
        return oracle.pair_forces(posq.astype(np.float32), zero, lj, L)[1].sum()

    for r in (7.5, 8.25, 8.75):  # r and r +- h exactly representable in f32
        h = 1.0 / 256.0
        posq, L = _two(r)
        F, E, W, _ = oracle.pair_forces(posq, zero, lj, L)
        dE = (energy(r + h) - energy(r - h)) / (2 * h)
        assert F[1, 0] == pytest.approx(-dE, rel=1e-4)
    assert abs(energy(8.9999)) < 1e-9


@pytest.mark.parametrize("r", [2.0, 5.0, 12.0])
def test_coulomb_real_space_closed_form(oracle, r):
    qi, qj, alpha = -0.82, 0.41, 0.3
    posq, L = _two(r, q=(qi, qj))
    lj = _lj(eps=0.0, cut=15.0)
    F, E, W, _ = oracle.pair_forces(posq, np.zeros(2, np.uint32), lj, L,
                                    coul=dict(alpha=alpha, r_cut=15.0))
    r = float(posq[1, 0])
    qi, qj = float(posq[0, 3]), float(posq[1, 3])  # the f32 charges the oracle sees
    e = qi * qj * math.erfc(alpha * r) / r
    f = qi * qj / r ** 2 * (math.erfc(alpha * r) + 2 * alpha * r / math.sqrt(math.pi)
                            * math.exp(-(alpha * r) ** 2))
    assert E.sum() == pytest.approx(e, rel=1e-12)
    assert F[1, 0] == pytest.approx(f, rel=1e-12)


def test_hastings_erfc_within_published_bound(oracle):
    # A&S 7.1.26 has |error| <= 1.5e-7 on erfc
    posq, L = _two(5.0, q=(1.0, 1.0))
    lj = _lj(eps=0.0, cut=15.0)
    worst = 0.0
    for r in np.linspace(0.5, 14.0, 60):
        posq[1, 0] = r
        a = oracle.pair_forces(posq, np.zeros(2, np.uint32), lj, L, coul=dict(alpha=0.3, r_cut=15.0))[1].sum()
        oracle.set_erfc_mode(True)
        b = oracle.pair_forces(posq, np.zeros(2, np.uint32), lj, L, coul=dict(alpha=0.3, r_cut=15.0))[1].sum()
        oracle.set_erfc_mode(False)
        rr = float(posq[1, 0])
        worst = max(worst, abs(a - b) * rr)  # |erfc error|
    assert worst <= 1.5e-7


def test_exclusions_remove_pairs(oracle):
    posq, L = _two(3.0, q=(0.5, -0.5))
    excl = (np.array([0, 1, 2], np.uint32), np.array([1, 0], np.uint32), np.zeros(2))
    F, E, W, _ = oracle.pair_forces(posq, np.zeros(2, np.uint32), _lj(), L,
                                    coul=dict(alpha=0.3, r_cut=9.0), excl=excl)
    assert np.all(F == 0) and np.all(E == 0) and np.all(W == 0)


def test_minimum_image_across_boundary(oracle):
    # two atoms 1 Bohr apart through the periodic boundary
    L = np.array([20.0, 20.0, 20.0, 0, 0, 0])
    posq = np.array([[9.5, 0, 0, 0], [-9.5, 0, 0, 0]], np.float32)
    lj = _lj(sigma=1.0, eps=0.01, cut=3.0)
    F, E, W, _ = oracle.pair_forces(posq, np.zeros(2, np.uint32), lj, L)
    posq2 = np.array([[0.5, 0, 0, 0], [-0.5, 0, 0, 0]], np.float32)
    F2, E2, W2, _ = oracle.pair_forces(posq2, np.zeros(2, np.uint32), lj, L)
    assert E.sum() == pytest.approx(E2.sum(), rel=1e-12)
    # atom 0 at +9.5 is pushed towards +x (away from its image partner at +10.5)
    assert F[0, 0] == pytest.approx(-F2[0, 0], rel=1e-12)


def _random_system(n, L, seed):
    rng = np.random.default_rng(seed)
    posq = np.zeros((n, 4), np.float32)
    posq[:, :3] = rng.uniform(-L / 2, L / 2, size=(n, 3))
    posq[:, 3] = rng.choice([-0.5, 0.5], size=n)
    return posq, np.array([L, L, L, 0, 0, 0], float)


@pytest.mark.parametrize("box", [30.0, 25.0])  # 3 cells per axis vs the all-pairs path
def test_pair_forces_newton_and_brute_force(oracle, box):
    posq, L = _random_system(300, box, 3)
    lj = _lj(sigma=2.5, eps=0.001, cut=9.0)
    F, E, W, _ = oracle.pair_forces(posq, np.zeros(300, np.uint32), lj, L,
                                    coul=dict(alpha=0.35, r_cut=9.0))
    assert np.abs(F.sum(axis=0)).max() < 1e-10 * np.abs(F).max() * 300
    # brute force in numpy over all pairs
    p = posq[:, :3].astype(np.float64)
    q = posq[:, 3].astype(np.float64)
    d = p[:, None, :] - p[None, :, :]
    d -= L[0] * np.floor(d / L[0] + 0.5)
    r = np.sqrt((d ** 2).sum(-1))
    np.fill_diagonal(r, np.inf)
    m = r <= 9.0
    sr6 = (2.5 / r) ** 6
    elj = np.where(m, 4e-3 * (sr6 ** 2 - sr6), 0)
    from scipy.special import erfc
    ec = np.where(m, q[:, None] * q[None, :] * erfc(0.35 * r) / r, 0)
    assert E.sum() == pytest.approx(0.5 * (elj + ec).sum(), rel=1e-10)
    flj = np.where(m, 24e-3 / r ** 2 * (2 * sr6 ** 2 - sr6), 0)
    fc = np.where(m, q[:, None] * q[None, :] / r ** 2 * (erfc(0.35 * r) / r + 2 * 0.35 / np.sqrt(np.pi)
                                                            * np.exp(-(0.35 * r) ** 2)), 0)
    Fb = ((flj + fc)[:, :, None] * d).sum(1)
    assert np.abs(F - Fb).max() < 1e-10 * np.abs(Fb).max()
    # virial = sum r_ij . F_ij
    Wb = 0.5 * np.where(m, (flj + fc) * r ** 2, 0).sum()
    assert W.sum() == pytest.approx(Wb, rel=1e-10)


def test_pair_set_matches_brute_force(oracle):
    posq, L = _random_system(400, 25.0, 5)
    rs = 7.0
    pairs, band = oracle.pair_set(posq, L, rs)
    p = posq[:, :3].astype(np.float64)
    d = p[:, None, :] - p[None, :, :]
    d -= L[0] * np.floor(d / L[0] + 0.5)
    r2 = (d ** 2).sum(-1)
    iu = np.triu_indices(400, 1)
    inside = r2[iu] <= rs * rs
    want = set(zip(iu[0][inside].tolist(), iu[1][inside].tolist()))
    got = set(map(tuple, pairs.tolist())) | set(map(tuple, band.tolist()))
    assert want - got == set()
    assert set(map(tuple, pairs.tolist())) <= want


def test_angle_forces_are_energy_gradient(oracle):
    # harmonic angle (src/forces/angle.rs:258-313): F = -grad E by central
    # differences, Newton's third law, zero force at the rest angle
    from heddlemd_amd import systems
    s = systems.polymer_melt(n_chains=2, n_beads=6, seed=3)
    rng = np.random.default_rng(0)
    posq = s.posq.copy()
    posq[:, :3] += rng.normal(0, 0.2, size=(s.n, 3)).astype(np.float32)
    ijk, kth, th0 = s.angles
    F, E, W = oracle.angle_forces(posq, s.lattice, ijk, kth, th0)
    assert np.allclose(F.sum(0), 0.0, atol=1e-12)
    h = 1.0 / 1024
    for a in range(s.n):
        for d in range(3):
            pp = posq.copy()
            pm = posq.copy()
            pp[a, d] += np.float32(h)
            pm[a, d] -= np.float32(h)
            ep = oracle.angle_forces(pp, s.lattice, ijk, kth, th0)[1].sum()
            em = oracle.angle_forces(pm, s.lattice, ijk, kth, th0)[1].sum()
            dx = float(pp[a, d]) - float(pm[a, d])
            assert -(ep - em) / dx == pytest.approx(F[a, d], rel=1e-4, abs=1e-9)
    # at the rest geometry (the generator's exact 114 degrees) the force vanishes
    F0, E0, _ = oracle.angle_forces(s.posq, s.lattice, ijk, kth, th0)
    assert np.abs(F0).max() < 1e-5 * np.abs(F).max()


def test_exclusion_span_gives_1_2_and_1_3_pairs():
    from heddlemd_amd import systems
    s = systems.polymer_melt(n_chains=3, n_beads=7, seed=4)
    off, part, _ = s.exclusion_csr()
    for c in range(3):
        for b in range(7):
            i = 7 * c + b
            want = sorted(7 * c + x for x in range(max(0, b - 2), min(7, b + 3)) if x != b)
            assert part[off[i]:off[i + 1]].tolist() == want
