"""The reference's own example bundles through `heddlemd run` on the GPU
(inputs under tests/golden/examples): lj-10000-argon (100 NVE steps, frames
with velocities every 10, log every 5), spc-water-256 (flexible SPC
water: Morse O-H bonds, harmonic angles, switched LJ, SPME on a 20 x 20 x 80
grid, 100 steps of 0.5 fs), and chignolin (a peptide with 758 SETTLE waters,
2,446 atoms: bonds, angles, periodic dihedrals, explicit exclusions with scaled
1-4 pairs, 16 LJ types, SPME order 5 on 32^3; steepest descent, then 2,000
CSVR steps twice), spc-water-8192 and spc-water-65536 (SETTLE water on
48 x 48 x 96 and 96 x 96 x 192 grids: steepest descent, CSVR NVT, CSVR +
C-rescale NPT).  Each runs twice; the output files must be byte-identical
across the two runs."""
import os
import re
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from tests.cli_inputs import example_bundle, run_cli  # noqa: E402


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("name,root,frames,rows", [
    ("lj-10000-argon", "argon", 11, 21),
    ("spc-water-256", "spc", 11, 21),
])
def test_example_bundle_runs_and_reruns_identically(tmp_path, name, root, frames, rows):
    say(f"case: heddlemd run examples/{name}, twice")
    outs = []
    for rep in range(2):
        d = tmp_path / f"r{rep}"
        d.mkdir()
        cfg = example_bundle(name, str(d))
        r = run_cli("run", cfg, timeout=300)
        say("  " + (r.stdout + r.stderr).strip().replace("\n", "\n  "))
        assert r.returncode == 0
        assert re.fullmatch(rf"\[heddlemd\] phase `run`: 100 steps in \d+ (ms|µs) "
                            rf"\(frames: {frames}, log rows: {rows}\)", r.stdout.splitlines()[0])
        outs.append({ext: open(os.path.join(str(d), f"{root}.out.run.{ext}"), "rb").read()
                     for ext in ("xyz", "log")})
    assert outs[0]["xyz"] == outs[1]["xyz"]
    assert outs[0]["log"] == outs[1]["log"]
    log = outs[0]["log"].decode().splitlines()
    assert log[0] == "step,time,kinetic_energy,temperature"
    temps = np.array([float(x.split(",")[3]) for x in log[1:]])
    say(f"  T from {temps[0]:.1f} to {temps[-1]:.1f} K (max {temps.max():.1f})")
    assert np.all(np.isfinite(temps)) and temps.max() < 1500.0


@pytest.mark.parametrize("name,root,steps,tmax", [
    ("chignolin", "chignolin", (2000, 2000), 600.0),
    ("spc-water-8192", "water", (2500, 1000), 1500.0),
    ("spc-water-65536", "water", (250, 125), 1500.0),
])
def test_phased_bundle_runs_and_reruns_identically(tmp_path, name, root, steps, tmax):
    # minimization, CSVR NVT, then (water) CSVR + C-rescale NPT
    say(f"case: heddlemd run examples/{name}, twice")
    outs = []
    for rep in range(2):
        d = tmp_path / f"r{rep}"
        d.mkdir()
        cfg = example_bundle(name, str(d))
        r = run_cli("run", cfg, timeout=600)
        say("  " + (r.stdout + r.stderr).strip().replace("\n", "\n  "))
        assert r.returncode == 0
        lines = r.stdout.splitlines()
        assert re.fullmatch(r"\[heddlemd\] phase `min`: \d+ iters in \d+ (ms|µs) \(converged: "
                            r"(force_tolerance|energy_tolerance|force_zero|max_iterations), "
                            r"frames: \d+, log rows: \d+\)", lines[0]), lines[0]
        assert lines[1].startswith(f"[heddlemd] phase `equil`: {steps[0]} steps")
        assert lines[2].startswith(f"[heddlemd] phase `prod`: {steps[1]} steps")
        outs.append({f: open(os.path.join(str(d), f), "rb").read()
                     for f in sorted(os.listdir(str(d))) if ".out." in f
                     and not f.endswith(".timings")})
    assert sorted(outs[0]) == sorted(outs[1])
    for f in outs[0]:
        assert outs[0][f] == outs[1][f], f
    minlog = outs[0][f"{root}.out.min.minlog"].decode().splitlines()
    e_min = [float(x.split(",")[1]) for x in minlog[1:]]
    assert e_min[-1] < e_min[0]
    for ph in ("equil", "prod"):
        log = outs[0][f"{root}.out.{ph}.log"].decode().splitlines()
        temps = np.array([float(x.split(",")[3]) for x in log[1:]])
        say(f"  {ph}: T from {temps[0]:.1f} to {temps[-1]:.1f} K (mean {temps.mean():.1f})")
        assert np.all(np.isfinite(temps)) and temps.max() < tmax
