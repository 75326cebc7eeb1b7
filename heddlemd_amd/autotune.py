"""Step autotuner: times the step of one system over the launch-level
choices that leave its physics unchanged and keeps the fastest.

The choices are the neighbour skin (every skin gives the same forces: the
rebuild is triggered by the displacement check, so the list always holds
every pair within the cutoff), the pair-list form (`hd_set_pair_list`:
packed 64-atom clusters or per-atom full lists) and the reciprocal-space
transform (`hd_set_spme_dft`: FP32 LDS FFT or split-precision matrix-core
DFT, the latter only on power-of-two axes of 32 to 128 points).  The
reference has no autotuner; this is MI355X-side tuning of the same step.

`make_engine(r_skin)` must return a set-up `engine.Engine` at the state to
time from; every candidate starts from that state, runs `warmup` steps and
is timed over `steps` steps between device synchronisations.
"""
from __future__ import annotations

import itertools
import time

import numpy as np

from ._lib import HeddleError


def _dft_ok(grid):
    return all(g in (32, 64, 128) for g in grid)


def tune_step(make_engine, skins, pair_lists=(0, 1), dft_modes=(0, 1), grid=None, steps=50,
              warmup=10):
    """Returns (best, table): best = dict(r_skin, pair_list, dft_mode,
    ms_per_step); table = one such dict per candidate (failed candidates
    carry an 'error' entry instead of a time)."""
    modes = [m for m in dft_modes if m == 0 or (grid is not None and _dft_ok(grid))]
    table = []
    for skin in skins:
        for pl, dm in itertools.product(pair_lists, modes):
            row = dict(r_skin=float(skin), pair_list=int(pl), dft_mode=int(dm))
            e = make_engine(skin)
            try:
                e.set_pair_list(pl)
                if grid is not None:
                    e.set_spme_dft(dm)
                e.force_eval(scalars=False)
                e.step(warmup)
                e.sync()
                t0 = time.perf_counter()
                e.step(steps)
                e.sync()
                row["ms_per_step"] = 1e3 * (time.perf_counter() - t0) / steps
            except HeddleError as err:
                row["error"] = str(err)
            finally:
                e.close()
            table.append(row)
    timed = [r for r in table if "ms_per_step" in r]
    if not timed:
        raise HeddleError("no candidate ran")
    best = min(timed, key=lambda r: r["ms_per_step"])
    return best, table


def forces_agree(make_engine, configs, grid=None):
    """Forces at the starting state under each configuration (dicts with
    r_skin, pair_list, dft_mode): the launch-level choices must not change
    them beyond f32 rounding.  Returns the list of (config, forces)."""
    out = []
    for cfg in configs:
        e = make_engine(cfg["r_skin"])
        try:
            e.set_pair_list(cfg["pair_list"])
            if grid is not None:
                e.set_spme_dft(cfg["dft_mode"])
            e.force_eval(scalars=False)
            out.append((cfg, np.array(e.forces())))
        finally:
            e.close()
    return out
