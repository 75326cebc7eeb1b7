"""Bonded terms and scaled exclusions on the GPU, through the C ABI, against
the FP64 oracle: harmonic and Morse bonds, harmonic angles, periodic
dihedrals, full 1-2 / 1-3 exclusions and scaled 1-4 pairs (LJ 0.5, Coulomb
1/1.2) on 120 flexible charged chains of 20 beads with SPME; and the
exclusion-list rules hd_set_exclusions accepts and refuses."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from heddlemd_amd import systems, units as U  # noqa: E402
from heddlemd_amd.engine import Engine, HeddleError, engine_for  # noqa: E402
from tests.forcecheck import check_forces  # noqa: E402


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _reference(oracle, s):
    Fp, Ep, Wp, band = oracle.pair_forces(s.posq, s.types, s.lj.as_dict(), s.lattice,
                                          coul=dict(alpha=s.spme.alpha, r_cut=s.spme.r_cut),
                                          excl=s.exclusion_csr(), band=1e-6)
    Fr, Er, Wr, er, wr = oracle.spme_recip(s.posq, s.lattice, s.spme.grid, 4, s.spme.alpha)
    Fa, Ea, Wa = oracle.angle_forces(s.posq, s.lattice, *s.angles)
    Fb, Eb, Wb = oracle.bond_forces(s.posq, s.lattice, *s.bonds)
    Fd, Ed, Wd = oracle.dihedral_forces(s.posq, s.lattice, *s.dihedrals)
    F = Fp + Fr + Fa + Fb + Fd
    E = (Ep + Er + Ea + Eb + Ed).sum()
    W = (Wp + Wr + Wa + Wb + Wd).sum()
    parts = dict(bond=Eb.sum(), angle=Ea.sum(), dihedral=Ed.sum(), pair=Ep.sum(), recip=er)
    return F, E, W, band, parts


def test_bonded_melt_forces_energy_virial(oracle):
    say("case: bonded melt 120 x 20, bonds + Morse + angles + dihedrals + scaled 1-4, SPME 32^3")
    s = systems.bonded_melt()
    e = engine_for(s, dt=U.time(1e-15))
    e.force_eval(scalars=True)
    f, sc = e.forces(), e.scalars()
    e.close()
    F, E, W, band, parts = _reference(oracle, s)
    say("  oracle terms: " + ", ".join(f"{k} {v:.6e}" for k, v in parts.items()))
    say(f"  pe {sc['pe']:.9e} (oracle {E:.9e}), virial {sc['virial']:.9e} (oracle {W:.9e})")
    check_forces(f, F, band, "  bonded melt", say)
    assert sc["pe"] == pytest.approx(E, rel=2e-6)
    assert sc["virial"] == pytest.approx(W, rel=2e-5, abs=1e-5 * abs(parts["pair"]))


def test_scaled_pairs_move_the_result_by_the_scaled_terms(oracle):
    # the same chains with every 1-4 scale set to 1 differ from the scaled
    # run by exactly (1 - s) of the 1-4 terms: the correction pass acts
    say("case: bonded melt, 1-4 scales 1 vs (0.5, 1/1.2)")
    s = systems.bonded_melt(n_chains=60)
    out = []
    for scaled in (True, False):
        ij, slj, sco = s.exclusions
        if not scaled:
            keep = (slj == 0.0) & (sco == 0.0)
            s.exclusions = (ij[keep], slj[keep], sco[keep])
        e = engine_for(s, dt=U.time(1e-15))
        e.force_eval(scalars=True)
        out.append((e.forces(), e.scalars(), s.exclusion_csr()))
        e.close()
    for (f, sc, csr) in out:
        Fp, Ep, _, band = oracle.pair_forces(s.posq, s.types, s.lj.as_dict(), s.lattice,
                                             coul=dict(alpha=s.spme.alpha, r_cut=s.spme.r_cut),
                                             excl=csr, band=1e-6)
        Fr = oracle.spme_recip(s.posq, s.lattice, s.spme.grid, 4, s.spme.alpha)[0]
        Fa = oracle.angle_forces(s.posq, s.lattice, *s.angles)[0]
        Fb = oracle.bond_forces(s.posq, s.lattice, *s.bonds)[0]
        Fd = oracle.dihedral_forces(s.posq, s.lattice, *s.dihedrals)[0]
        check_forces(f, Fp + Fr + Fa + Fb + Fd, band, "  ", say)
    d_dev = out[0][1]["pe"] - out[1][1]["pe"]
    assert abs(d_dev) > 0


def _tiny_engine(n):
    e = Engine(0)
    posq = np.zeros((n, 4), np.float32)
    posq[:, 0] = np.arange(n, dtype=np.float32) * 3.0
    e.set_system(posq, np.ones(n, np.float32), np.zeros(n, np.uint32), None)
    return e


def test_exclusion_rules_accepted_and_refused():
    say("case: hd_set_exclusions key derivation")
    # molecules (cliques) of 3, and chains of 6 excluded up to distance 2
    e = _tiny_engine(12)
    tri = [(a, b) for m in range(4) for a, b in ((3 * m, 3 * m + 1), (3 * m, 3 * m + 2),
                                                 (3 * m + 1, 3 * m + 2))]
    e.set_exclusions(np.array(tri), np.zeros(len(tri)), np.zeros(len(tri)))
    e.close()
    e = _tiny_engine(12)
    ch = [(c * 6 + b, c * 6 + b + d) for c in range(2) for b in range(6) for d in (1, 2)
          if b + d < 6]
    e.set_exclusions(np.array(ch), np.zeros(len(ch)), np.zeros(len(ch)))
    e.close()
    # a chain missing one 1-2 pair is neither form: its molecule is left out
    # by key and its other pairs added back
    e = _tiny_engine(12)
    e.set_exclusions(np.array(ch[1:]), np.zeros(len(ch) - 1), np.zeros(len(ch) - 1))
    e.close()
    # scales outside [0, 1], a repeated pair, a self pair
    for ij, sl in (([(0, 1)], [1.5]), ([(0, 1), (1, 0)], [0.5, 0.5]), ([(2, 2)], [0.5])):
        e = _tiny_engine(12)
        with pytest.raises(HeddleError):
            e.set_exclusions(np.array(ij), np.array(sl), np.array(sl))
        e.close()


def test_irregular_exclusions_are_added_back(oracle):
    # chains whose full exclusions are 1-2, 1-3 and one extra pair (bead 0 -
    # bead 5): neither cliques nor fixed-span runs, so every intramolecular
    # pair is left out by key and the non-excluded ones (1-4 scaled, the
    # rest full) come back through the correction pass
    say("case: bonded melt 60 x 20 with an irregular exclusion per chain")
    s = systems.bonded_melt(n_chains=60)
    ij, slj, sco = s.exclusions
    nb = 20
    extra = np.array([(c * nb, c * nb + 5) for c in range(60)], np.uint32)
    s.exclusions = (np.concatenate([ij, extra]), np.concatenate([slj, np.zeros(60)]),
                    np.concatenate([sco, np.zeros(60)]))
    e = engine_for(s, dt=U.time(1e-15))
    e.force_eval(scalars=True)
    f, sc = e.forces(), e.scalars()
    e.close()
    F, E, W, band, parts = _reference(oracle, s)
    say(f"  pe {sc['pe']:.9e} (oracle {E:.9e})")
    check_forces(f, F, band, "  irregular exclusions", say)
    assert sc["pe"] == pytest.approx(E, rel=2e-6)
    assert sc["virial"] == pytest.approx(W, rel=2e-5, abs=1e-5 * abs(parts["pair"]))
