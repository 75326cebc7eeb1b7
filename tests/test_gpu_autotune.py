"""The step autotuner (heddlemd_amd/autotune.py) on the GPU: it times every
candidate (neighbour skin x pair-list form x reciprocal transform) of
8,192-molecule SPC water, keeps the fastest, and every candidate gives the
same forces at the starting state to the f32 force bar."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from heddlemd_amd import systems, units as U  # noqa: E402
from heddlemd_amd.autotune import forces_agree, tune_step  # noqa: E402
from heddlemd_amd.engine import engine_for  # noqa: E402
from tests.forcecheck import force_errors  # noqa: E402


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_autotuner_picks_the_fastest_and_forces_do_not_move():
    say("case: water 8192, autotune skin x pair list x transform")
    s = systems.config_water_8192()

    def make(skin):
        return engine_for(s, dt=U.time(2e-15), r_skin=skin)

    skins = [U.length(x * 1e-9) for x in (0.05, 0.1, 0.2)]
    best, table = tune_step(make, skins, grid=s.spme.grid, steps=30, warmup=5)
    for r in table:
        say(f"  skin {r['r_skin'] / U.length(1e-9):.2f} nm, list {r['pair_list']}, dft "
            f"{r['dft_mode']}: {r.get('ms_per_step', float('nan')):.4f} ms/step {r.get('error', '')}")
    say(f"  best: {best}")
    timed = [r for r in table if "ms_per_step" in r]
    assert len(timed) == len(table) == 12
    assert best["ms_per_step"] == min(r["ms_per_step"] for r in timed)
    res = forces_agree(make, table, grid=s.spme.grid)
    ref = res[0][1]
    worst = 0.0
    for cfg, f in res[1:]:
        rms, mx, rel = force_errors(f, ref)
        worst = max(worst, rel)
    say(f"  largest per-atom force difference between candidates {worst:.2e}")
    assert worst < 1e-5
