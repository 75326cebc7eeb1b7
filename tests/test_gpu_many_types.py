"""Pair forces with more than four Lennard-Jones types, the pair kernels'
LDS-table form, against the FP64 oracle: seven types with Lorentz-Berthelot
sigma and epsilon per pair (one type with epsilon 0), on jittered argon
positions with alternating charges and SPME.  Four paths: the regrouped loop
with LJ + Coulomb on one cutoff, the regrouped LJ-only loop, the general
loop (cubic switch from 0.9 r_c), and per-atom lists."""
import copy
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from heddlemd_amd import systems, units as U  # noqa: E402
from heddlemd_amd.engine import HeddleError, engine_for  # noqa: E402
from tests.forcecheck import check_forces  # noqa: E402
from tests.test_gpu_parity import _jittered_argon  # noqa: E402


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def many_types(nt=7, charged=True, switch=False, seed=5):
    s = copy.deepcopy(_jittered_argon(n_atoms=4000, seed=seed))
    rng = np.random.default_rng(seed)
    s.types = rng.integers(0, nt, size=s.n).astype(np.uint32)
    sig_t = U.length(rng.uniform(0.30e-9, 0.40e-9, nt))
    eps_t = U.energy(rng.uniform(0.5, 1.5, nt) * 1.654e-21)
    eps_t[nt - 1] = 0.0
    sig = 0.5 * (sig_t[:, None] + sig_t[None, :])
    eps = np.sqrt(eps_t[:, None] * eps_t[None, :])
    rc = U.length(0.85e-9)
    cut = np.full((nt, nt), rc)
    rsw = np.full((nt, nt), 0.9 * rc if switch else rc)
    s.lj = systems.LJTable(nt, sig, eps, cut, rsw)
    if charged:
        s.posq[:, 3] = np.where(np.arange(s.n) % 2 == 0, 0.3, -0.3).astype(np.float32)
        s.spme = systems.SpmeParams(systems.ewald_alpha(rc, 1e-5), rc, (64, 64, 64))
    else:
        s.posq[:, 3] = 0.0
        s.spme = None
    return s


@pytest.mark.parametrize("charged,switch,pair_list", [
    (True, False, 0), (False, False, 0), (True, True, 0), (True, False, 1)])
def test_seven_types_match_the_oracle(oracle, charged, switch, pair_list):
    tag = (f"7 types, {'LJ + Coulomb' if charged else 'LJ only'}, "
           f"{'switched' if switch else 'one cutoff'}, "
           f"{'atom lists' if pair_list else 'cluster list'}")
    say("case: " + tag)
    s = many_types(charged=charged, switch=switch)
    e = engine_for(s, dt=U.time(2e-15), pair_list=pair_list)
    e.force_eval(scalars=True)
    f = e.forces()
    sc = e.scalars()
    e.close()
    coul = dict(alpha=s.spme.alpha, r_cut=s.spme.r_cut) if charged else None
    F, E, W, band = oracle.pair_forces(s.posq, s.types, s.lj.as_dict(), s.lattice, coul=coul,
                                       band=1e-6)
    if charged:
        Fr, Er, Wr = oracle.spme_recip(s.posq, s.lattice, s.spme.grid, 4, s.spme.alpha)[:3]
        F, E, W = F + Fr, E + Er, W + Wr
    check_forces(f, F, band, tag, say)
    assert sc["pe"] == pytest.approx(E.sum(), rel=1e-5)
    assert sc["virial"] == pytest.approx(W.sum(), rel=1e-5)


def test_type_count_limit():
    s = many_types(nt=33, charged=False)
    with pytest.raises(HeddleError, match="at most 32 types"):
        engine_for(s, dt=U.time(2e-15))
