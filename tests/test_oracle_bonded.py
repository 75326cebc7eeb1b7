"""The oracle's bonded terms and per-term exclusion scales against closed
forms, central differences of the energy and invariants, on the geometries
of the reference's own tests (tests/forces_dihedral.rs:111-122,
tests/forces_harmonic_bond.rs:208-240)."""
import numpy as np
import pytest

from oracle import oracle

BOX = np.array([10.0, 10.0, 10.0, 0.0, 0.0, 0.0])


def _posq(xyz, q=None):
    xyz = np.asarray(xyz, np.float64)
    p = np.zeros((len(xyz), 4), np.float32)
    p[:, :3] = xyz + 5.0  # away from the box faces
    if q is not None:
        p[:, 3] = q
    return p


def unit_dihedral(phi):
    # j at the origin, k on +x, i on +y, l at angle phi about the j-k axis
    return _posq([[0.0, 1.0, 0.0], [0.0, 0.0, 0.0], [1.0, 0.0, 0.0],
                  [1.0, np.cos(phi), np.sin(phi)]])


def _dih(phi, k=1.0, n=1.0, phi0=0.0):
    return oracle.dihedral_forces(unit_dihedral(phi), BOX, [0, 1, 2, 3], [[k, n, phi0]])


@pytest.mark.parametrize("phi", [0.3, 1.1, 2.5, -0.8, -2.9])
def test_dihedral_angle_and_sign(phi):
    # U = k (1 + cos(n phi - phi0)): with phi0 = 0 the energy fixes |phi|,
    # with phi0 = pi/2 (U = k (1 + sin phi)) its sign
    _, E0, _ = _dih(phi, k=1.0, n=1.0, phi0=0.0)
    _, E1, _ = _dih(phi, k=1.0, n=1.0, phi0=np.pi / 2)
    assert E0.sum() == pytest.approx(1.0 + np.cos(phi), abs=1e-6)
    assert E1.sum() == pytest.approx(1.0 + np.sin(phi), abs=1e-6)
    _, E3, _ = _dih(phi, k=0.7, n=3.0, phi0=0.4)
    assert E3.sum() == pytest.approx(0.7 * (1.0 + np.cos(3 * phi - 0.4)), abs=1e-6)


def test_dihedral_equilibrium_has_no_force():
    F, _, _ = _dih(0.5, k=1.0, n=1.0, phi0=0.5)
    assert np.abs(F).max() < 1e-6


@pytest.mark.parametrize("phi", [0.2, 0.7, 1.9, -1.3])
def test_dihedral_newton_torque_and_central_difference(phi):
    p = unit_dihedral(phi).astype(np.float64)
    p[:, :3] += np.array([[0.01, -0.02, 0.03], [0.0, 0.01, -0.01], [0.02, 0.0, 0.01],
                          [-0.01, 0.02, 0.0]])  # off the symmetric placement
    p = p.astype(np.float32)
    prm = [[0.9, 2.0, 0.3]]
    F, E, W = oracle.dihedral_forces(p, BOX, [0, 1, 2, 3], prm)
    assert np.abs(F.sum(0)).max() < 1e-12
    x = p[:, :3].astype(np.float64)
    torque = np.cross(x - x.mean(0), F).sum(0)
    assert np.abs(torque).max() < 1e-10
    # central differences of the energy in double: move one coordinate by
    # a step the f32 positions represent exactly
    h = 2.0 ** -12
    for a in range(4):
        for d in range(3):
            pp, pm = p.copy(), p.copy()
            pp[a, d] += h
            pm[a, d] -= h
            ep = oracle.dihedral_forces(pp, BOX, [0, 1, 2, 3], prm)[1].sum()
            em = oracle.dihedral_forces(pm, BOX, [0, 1, 2, 3], prm)[1].sum()
            assert F[a, d] == pytest.approx(-(ep - em) / (2 * h), abs=2e-6)
    # virial W = sum_a r_a . F_a (translation invariant since sum F = 0)
    assert W.sum() == pytest.approx(float((x * F).sum()), abs=1e-10)


@pytest.mark.parametrize("kind,prm", [(0, (2.0, 1.5, 0.0)), (1, (0.3, 1.7, 1.4))])
@pytest.mark.parametrize("r", [1.2, 1.5, 1.9])
def test_bond_closed_form_and_difference(kind, prm, r):
    p = _posq([[0.0, 0.0, 0.0], [r * 0.6, r * 0.8, 0.0]])
    F, E, W = oracle.bond_forces(p, BOX, [0, 1], [kind], [prm])
    d = (p[0, :3].astype(np.float64) - p[1, :3])
    rr = np.linalg.norm(d)
    if kind == 0:
        u, dudr = 0.5 * prm[0] * (rr - prm[1]) ** 2, prm[0] * (rr - prm[1])
    else:
        e = np.exp(-prm[1] * (rr - prm[2]))
        u, dudr = prm[0] * (1 - e) ** 2, 2 * prm[0] * prm[1] * (1 - e) * e
    assert E.sum() == pytest.approx(u, rel=1e-12, abs=1e-15)
    np.testing.assert_allclose(F[0], -dudr * d / rr, rtol=1e-12, atol=1e-15)
    np.testing.assert_allclose(F[1], dudr * d / rr, rtol=1e-12, atol=1e-15)
    assert W.sum() == pytest.approx(-dudr * rr, rel=1e-12, abs=1e-15)


def test_bond_signs_match_the_reference_tests():
    # compressed: atom 0 pushed to -x, stretched: pulled to +x
    Fc, _, _ = oracle.bond_forces(_posq([[0, 0, 0], [1.0, 0, 0]]), BOX, [0, 1], [0], [(2.0, 1.5, 0)])
    Fs, _, _ = oracle.bond_forces(_posq([[0, 0, 0], [2.0, 0, 0]]), BOX, [0, 1], [0], [(2.0, 1.5, 0)])
    assert Fc[0, 0] < 0 < Fc[1, 0] and Fs[1, 0] < 0 < Fs[0, 0]
    F0, _, _ = oracle.bond_forces(_posq([[0, 0, 0], [1.5, 0, 0]]), BOX, [0, 1], [0], [(2.0, 1.5, 0)])
    assert np.abs(F0).max() < 1e-6


def test_bond_across_the_box_face_uses_the_minimum_image():
    p = _posq([[0, 0, 0], [0, 0, 0]])
    p[0, 0], p[1, 0] = 0.2, 9.6  # 0.6 apart through the face
    F, E, _ = oracle.bond_forces(p, BOX, [0, 1], [0], [(1.0, 0.5, 0)])
    assert E.sum() == pytest.approx(0.5 * (0.6 - 0.5) ** 2, rel=1e-5)
    assert F[0, 0] < 0  # stretched: atom 0 pulled towards atom 1 through the face


def test_split_exclusion_scales_apply_per_term():
    # two charged LJ atoms 1.8 apart: scale_lj scales the LJ term only,
    # scale_coul the real-space Coulomb term only
    p = _posq([[0, 0, 0], [1.8, 0, 0]], q=[0.5, -0.4])
    types = np.zeros(2, np.uint32)
    lj = dict(ntypes=1, sigma=[[1.6]], epsilon=[[0.01]], cutoff=[[4.0]], r_switch=[[4.0]])
    coul = dict(alpha=0.8, r_cut=4.0)
    off = np.array([0, 1, 2], np.uint32)
    par = np.array([1, 0], np.uint32)
    full = oracle.pair_forces(p, types, lj, BOX, coul=coul)
    lj_only = oracle.pair_forces(p, types, lj, BOX)
    coul_only = oracle.pair_forces(p, types, dict(lj, epsilon=[[0.0]]), BOX, coul=coul)
    for slj, sc in ((0.5, 1.0 / 1.2), (0.0, 1.0), (1.0, 0.0), (0.25, 0.75)):
        got = oracle.pair_forces(p, types, lj, BOX, coul=coul,
                                 excl=(off, par, np.full(2, slj), np.full(2, sc)))
        want_F = slj * lj_only[0] + sc * coul_only[0]
        want_E = slj * lj_only[1] + sc * coul_only[1]
        np.testing.assert_allclose(got[0], want_F, rtol=1e-12, atol=1e-15)
        np.testing.assert_allclose(got[1], want_E, rtol=1e-12, atol=1e-15)
    # the single-scale form applies one scale to both terms
    one = oracle.pair_forces(p, types, lj, BOX, coul=coul, excl=(off, par, np.full(2, 0.5)))
    np.testing.assert_allclose(one[0], 0.5 * full[0], rtol=1e-12)
