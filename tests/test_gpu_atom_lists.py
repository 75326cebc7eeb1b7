"""The per-atom full neighbour lists (hd_set_pair_list(ctx, 1)) against the
oracle and against the packed cluster list on the same state: pair set
(exactly P_s outside the 4-ulp boundary band), forces, energy and virial at
the tolerances of the cluster-list tests, agreement of the two list forms
to f32 rounding, run-to-run and graph-replay byte identity, and a switch
of the list form after set-up."""
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


def _jittered_argon(seed=1):
    s = systems.config_argon()
    rng = np.random.default_rng(seed)
    p = s.posq[:, :3].astype(np.float64) + rng.normal(0, 0.3, size=(s.n, 3))
    s.posq[:, :3] = systems.wrap_into_box(p, s.lattice).astype(np.float32)
    return s




def test_atom_lists_pair_set_and_argon_forces(oracle):
    say("case: argon 10000, atom lists: pair set + forces")
    s = _jittered_argon(seed=3)
    e = engine_for(s, dt=U.time(5e-15), pair_list=1)
    gp = e.pairs()
    useful = e.count_pairs()
    e.force_eval(scalars=True)
    f = e.forces()
    pe, _ = e.per_atom()
    sc = e.scalars()
    e.close()
    rs = float(s.lj.cutoff[0, 0]) * 1.3
    want, band = oracle.pair_set(s.posq, s.lattice, rs)
    got = set(zip(gp[:, 0].tolist(), gp[:, 1].tolist()))
    wantset = set(zip(want[:, 0].tolist(), want[:, 1].tolist()))
    bandset = set(zip(band[:, 0].tolist(), band[:, 1].tolist())) if len(band) else set()
    say(f"  pairs {len(got)} (oracle {len(wantset)}, band {len(bandset)})")
    assert wantset - bandset <= got
    assert got - wantset <= bandset
    within_rc, _ = oracle.pair_set(s.posq, s.lattice, float(s.lj.cutoff[0, 0]))
    assert abs(useful - len(within_rc)) <= 4
    F, E, W, fband = oracle.pair_forces(s.posq, s.types, s.lj.as_dict(), s.lattice, band=1e-6)
    check_forces(f, F, fband, "  argon atom lists", say)
    assert sc["pe"] == pytest.approx(E.sum(), rel=1e-5)
    assert sc["virial"] == pytest.approx(W.sum(), rel=1e-5)
    assert pe.sum() == pytest.approx(E.sum(), rel=1e-5)


def test_atom_lists_water_forces_and_agreement_with_cluster_list(oracle):
    say("case: water 8192, atom lists vs oracle and vs the cluster list")
    s = systems.config_water_8192()
    out = {}
    for mode in (0, 1):
        e = engine_for(s, dt=U.time(2e-15), pair_list=mode)
        e.force_eval(scalars=True)
        out[mode] = (e.forces(), e.scalars(), e.per_atom())
        e.close()
    Fp, Ep, Wp, band = oracle.pair_forces(s.posq, s.types, s.lj.as_dict(), s.lattice,
                                          coul=dict(alpha=s.spme.alpha, r_cut=s.spme.r_cut),
                                          excl=s.exclusion_csr(), band=1e-6)
    Fr, Er, Wr, er, wr = oracle.spme_recip(s.posq, s.lattice, s.spme.grid, 4, s.spme.alpha)
    f1, sc1, (pe1, w1) = out[1]
    check_forces(f1, Fp + Fr, band, "  water atom lists vs oracle", say)
    assert sc1["pe"] == pytest.approx((Ep + Er).sum(), rel=2e-6)
    assert sc1["virial"] == pytest.approx((Wp + Wr).sum(), rel=2e-6)
    f0, sc0, (pe0, w0) = out[0]
    rms01, mx01, _ = force_errors(f1, f0.astype(np.float64))
    say(f"  atom lists vs cluster list: rms {rms01:.3e} max {mx01:.3e}; "
        f"pe {sc1['pe']:.9e} vs {sc0['pe']:.9e}")
    assert rms01 <= 1e-6 and mx01 <= 2e-5
    assert sc1["pe"] == pytest.approx(sc0["pe"], rel=1e-6)
    assert sc1["virial"] == pytest.approx(sc0["virial"], rel=1e-6)
    # per-atom sums of ~600 f32 terms that cancel to a few percent
    assert np.allclose(pe1, pe0, rtol=1e-4, atol=1e-6 * np.abs(pe0).max())


def test_atom_lists_dynamics_repeat_graph_and_switch():
    say("case: water 8192, atom lists: 40 steps x2, graph vs eager, switch after set-up")
    s = systems.config_water_8192()
    kt = U.temperature(300.0)
    from heddlemd_amd.rng import maxwell_boltzmann
    v = maxwell_boltzmann(s.mass, kt, seed=3, n_constraints=s.n_constraints)
    res = []
    for graph in (True, True, False):
        e = engine_for(s, dt=U.time(1e-15), velocities=v, nhc_kt=kt, nhc_tau=U.time(1e-13),
                       pair_list=1)
        e.constrain_velocities()
        e.use_graph(graph)
        e.step(40, scalars_last=True)
        p, _ = e.positions(images=True)
        res.append((p, e.velocities(), e.scalars()["rebuilds"]))
        e.close()
    say(f"  rebuilds {res[0][2]}")
    assert res[0][2] >= 1
    for k in (1, 2):
        assert np.array_equal(res[0][0], res[k][0]) and np.array_equal(res[0][1], res[k][1])
    # switching the list form after set-up: same state, same forces as a
    # context built in that form
    e = engine_for(s, dt=U.time(1e-15), pair_list=0)
    e.force_eval(scalars=True)
    f0 = e.forces()
    e.set_pair_list(1)
    e.force_eval(scalars=True)
    f1 = e.forces()
    e.close()
    e = engine_for(s, dt=U.time(1e-15), pair_list=1)
    e.force_eval(scalars=True)
    f1b = e.forces()
    e.close()
    assert np.array_equal(f1, f1b)
    assert not np.array_equal(f0, f1)
