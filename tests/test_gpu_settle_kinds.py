"""SETTLE with more than one water geometry in a context (hd_add_settle):
the reference keeps d_OH, d_HH and the O/H masses per group
(src/integrator/settle.rs:196-240).  The 8,192-molecule SPC box with half
of its waters given a second geometry (0.9572 A O-H, 104.52 deg): every
water is held to its own geometry through 10 steps, and adding the second
half with the first half's geometry reproduces the one-geometry run byte
for byte."""
import math
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from heddlemd_amd import systems, units as U  # noqa: E402
from heddlemd_amd.engine import Engine  # noqa: E402


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _engine(s, second=None, split=True):
    g = s.water_groups
    half = len(g) // 2
    d_oh, d_hh = s.water_geom
    m_o, m_h = float(s.mass[g[0, 0]]), float(s.mass[g[0, 1]])
    e = Engine(0)
    e.set_system(s.posq, s.mass, s.types, s.excl_key)
    e.set_lattice(s.lattice)
    e.set_lj(s.lj)
    e.set_spme(s.spme)
    if split:
        e.set_settle(g[:half], d_oh, d_hh, m_o, m_h)
        e.add_settle(g[half:], *(second or (d_oh, d_hh)), m_o, m_h)
    else:
        e.set_settle(g, d_oh, d_hh, m_o, m_h)
    e.set_integrator(U.time(1e-15))
    e.setup()
    return e


def _dist(p, L, i, j):
    d = p[i, :3].astype(np.float64) - p[j, :3]
    d -= L[:3] * np.round(d / L[:3])
    return np.linalg.norm(d, axis=1)


def test_same_geometry_added_reproduces_one_geometry():
    say("case: water 8192, SETTLE groups set in two calls with one geometry")
    s = systems.config_water_8192()
    out = []
    for split in (False, True):
        e = _engine(s, split=split)
        e.force_eval(scalars=False)
        e.step(10, scalars_last=True)
        p, im = e.positions(images=True)
        out.append((p.tobytes(), im.tobytes(), e.velocities().tobytes()))
        e.close()
    assert out[0] == out[1]


def test_two_geometries_each_water_keeps_its_own():
    say("case: water 8192, half SPC (1.0 A, 109.47 deg), half 0.9572 A / 104.52 deg, 10 steps")
    s = systems.config_water_8192()
    d_oh2 = U.length(0.9572e-10)
    d_hh2 = 2.0 * d_oh2 * math.sin(math.radians(104.52) / 2.0)
    e = _engine(s, second=(d_oh2, d_hh2))
    e.force_eval(scalars=False)
    e.step(10, scalars_last=True)
    p = e.positions()
    L = e.lattice()
    sc = e.scalars()
    e.close()
    g = s.water_groups
    half = len(g) // 2
    d_oh1, d_hh1 = s.water_geom
    for sel, (oh, hh), tag in ((slice(0, half), (d_oh1, d_hh1), "SPC"),
                               (slice(half, None), (d_oh2, d_hh2), "second")):
        gg = g[sel]
        r1 = _dist(p, L, gg[:, 0], gg[:, 1]) / oh - 1
        r2 = _dist(p, L, gg[:, 0], gg[:, 2]) / oh - 1
        r3 = _dist(p, L, gg[:, 1], gg[:, 2]) / hh - 1
        err = max(np.abs(r1).max(), np.abs(r2).max(), np.abs(r3).max())
        say(f"  {tag}: max relative bond error {err:.2e}")
        assert err < 2e-5
    assert np.isfinite(sc["ke"])
