"""Headline benchmark: one MI355X running the 196,608-atom SPC/SPME MD step.

Workload (one "step" = one full velocity-Verlet MD step on the GPU):
  SPC water, 65,536 molecules = 196,608 atoms, cubic box at 1.0 g/cm^3
  (L = 12.516 nm), O on the first sites of a 42^3 grid with Shoemake-uniform
  orientations (seed 7); LJ O-O sigma 0.3166 nm / eps 0.650 kJ/mol and
  real-space Ewald with r_c = 0.9 nm (erfc(alpha r_c) = 1e-5); SPME 128^3,
  order 4; SETTLE; Nose-Hoover chain at 300 K; C-rescale at 1 bar; dt 2 fs.
Untimed set-up: Maxwell-Boltzmann velocities (ChaCha8 seed 7), an
equilibration phase with a tight thermostat (the lattice start is far
from equilibrium), then `--warmup` steps at the production settings.
Timed: exactly `--steps` steps replayed as captured hipGraphs, bracketed by
device synchronisation (each step's SPME pipeline runs on a second stream
beside the neighbour build and the pair kernel).  Afterwards a second pass
of the same number of steps runs without graphs, every stage alone on one
stream with hipEvents around it, which gives the per-kernel durations used
for the roofline entry.

Prints one JSON line: metric, value, unit, the run's shape and the
roofline and CPU-baseline objects.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "ns/day (196k-atom SPC SPME) & batched atom-steps/s vs plain, same error"
# MI355X published peaks used for the roofline fraction
PEAK_FP32_TFLOPS = 157.3
PEAK_HBM_GBS = 8000.0
# Flops per useful pair of the production pair kernel (LJ + Hastings-erfc
# Coulomb force, i and j force sums, virial; a transcendental or an FMA's
# multiply-add counts once per operation), a conservative count
FLOPS_PER_PAIR = 55.0
# The production pair kernel: virial only (a barostat step that logs
# nothing), two LJ types, no switch
PAIR_KERNEL = "void k_pair<2, 2, false>"



def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def build_engine(system, U, args):
    """NVT engine at dt = 1 fs with a tight thermostat (the equilibration
    phase); the production settings are switched on after equilibration."""
    from heddlemd_amd.engine import engine_for
    from heddlemd_amd.rng import maxwell_boltzmann

    kt = U.temperature(args.temperature)
    v = maxwell_boltzmann(system.mass, kt, seed=7, n_constraints=system.n_constraints)
    e = engine_for(system, dt=U.time(1e-15), nhc_kt=kt,
                   nhc_tau=U.time(args.equil_tau_fs * 1e-15), velocities=v,
                   r_skin=U.length(args.skin_nm * 1e-9), pair_list=args.pair_list)
    e.constrain_velocities()
    e.force_eval(scalars=True)
    return e, kt


def autotune_skin(e, system, U, kt, args):
    """Time the production step from the equilibrated state at several
    neighbour skins (heddlemd_amd.autotune; the forces do not depend on the
    skin) and continue with the fastest.  Returns the engine to measure and
    the tuning record."""
    import dataclasses
    from heddlemd_amd.autotune import tune_step
    from heddlemd_amd.engine import engine_for

    state = dataclasses.replace(system, posq=e.positions(), lattice=e.lattice())
    v = e.velocities()
    baro = None if args.no_barostat else dict(
        pressure=U.pressure(1.0e5), tau=U.time(1.0e-12), compressibility=U.inv_pressure(4.5e-10),
        kt=kt, seed=13)

    def make(skin):
        return engine_for(state, dt=U.time(args.dt_fs * 1e-15), nhc_kt=kt,
                          nhc_tau=U.time(args.tau_fs * 1e-15), velocities=v, r_skin=skin,
                          crescale=baro, pair_list=args.pair_list)

    skins_nm = sorted(set([args.skin_nm, 0.08, 0.1, 0.14, 0.2]))
    best, table = tune_step(make, [U.length(x * 1e-9) for x in skins_nm],
                            pair_lists=(args.pair_list,), dft_modes=(0,), steps=100, warmup=20)
    nm = lambda b: round(b / U.length(1e-9), 3)
    rec = {"candidates": [{"skin_nm": nm(r["r_skin"]),
                           "ms_per_step": round(r.get("ms_per_step", float("nan")), 4)}
                          for r in table],
           "chosen_skin_nm": nm(best["r_skin"]), "steps_per_candidate": 100,
           "method": "heddlemd_amd.autotune.tune_step from the equilibrated state"}
    log(f"autotune: {rec}")
    if abs(nm(best["r_skin"]) - args.skin_nm) > 1e-9:
        e.close()
        args.skin_nm = nm(best["r_skin"])
        e = make(best["r_skin"])
        e.constrain_velocities()
        e.force_eval(scalars=True)
    return e, rec


def minimize(e, U, args):
    """Steepest descent from the lattice start, with the reference examples'
    settings (examples/spc-water-8192/water.in.toml [[minimization]]:
    force tolerance 1e-9 N, energy tolerance 1e-5, default steps)."""
    f_tol = 1.0e-9 / U.FORCE_AU_TO_N
    r = e.minimize_begin(U.length(1.0e-12))
    e0 = prev = r["energy"]
    why, it = "max_iterations", 0
    for it in range(1, args.min_iters + 1):
        r = e.minimize_step(U.length(1.0e-10))
        if r["max_force"] == 0.0:
            why = "force_zero"
            break
        if r["max_force"] <= f_tol:
            why = "force_tolerance"
            break
        if r["accepted"] and abs(r["energy"] - prev) <= 1e-5 * max(abs(r["energy"]), abs(prev),
                                                                    1e-30):
            why = "energy_tolerance"
            break
        if r["accepted"]:
            prev = r["energy"]
    log(f"minimization: {it} iterations ({why}), PE {e0:.2f} -> {r['energy']:.2f} Eh")


def equilibrate(e, system, U, kt, args):
    """The lattice start holds a large amount of potential energy: relax it
    with steepest descent, drain what remains with Maxwell-Boltzmann redraws
    at 1 fs, then settle at the production time step with a tight
    thermostat."""
    from heddlemd_amd.rng import maxwell_boltzmann

    ndof = 3 * system.n - system.n_constraints - 3

    def temp():
        sc = e.scalars()
        return 2 * sc["ke"] / ndof * U.TEMPERATURE_AU_TO_K, sc

    if args.min_iters > 0:
        minimize(e, U, args)
        e.force_eval(scalars=False)
    tau = U.time(args.tau_fs * 1e-15)
    for k in range(args.equil_redraws):
        e.step(50, scalars_last=True)
        T, sc = temp()
        log(f"equil redraw {k}: T = {T:.1f} K before redraw, PE = {sc['pe']:.2f} Eh")
        e.set_velocities(maxwell_boltzmann(system.mass, kt, seed=100 + k,
                                           n_constraints=system.n_constraints))
        e.constrain_velocities()
        e.set_nhc(kt, tau, n_constraints=system.n_constraints)  # fresh chain
    e.set_integrator(U.time(args.dt_fs * 1e-15))
    done = 0
    while done < args.equil:
        k = min(250, args.equil - done)
        e.step(k, scalars_last=True)
        done += k
        T, sc = temp()
        log(f"equil {done}/{args.equil}: T = {T:.1f} K, PE = {sc['pe']:.2f} Eh, "
            f"rebuilds {int(sc['rebuilds'])}")
    e.set_nhc(kt, U.time(args.tau_fs * 1e-15), n_constraints=system.n_constraints)
    if not args.no_barostat:
        e.set_crescale(pressure=U.pressure(1.0e5), tau=U.time(1.0e-12),
                       compressibility=U.inv_pressure(4.5e-10), kt=kt, seed=13)


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or platform.machine()


def cpu_baseline(system, v, args):
    """The reference algorithm on the host: the CPU restatement of the same
    MD step (oracle/cpu_step.c: cell-list half list, LJ + Hastings-erfc pair
    terms, SPME order 4 on the same grid, SETTLE, NHC, C-rescale; f32 with
    FP64 reductions, OpenMP over all the threads this process may use),
    started from the GPU run's final state and timed over --cpu-steps steps
    after 2 warm-up steps."""
    from heddlemd_amd import units as U
    from oracle.cpu_step import CpuStep

    kt = U.temperature(args.temperature)
    c = CpuStep(system, dt=U.time(args.dt_fs * 1e-15), r_skin=U.length(args.skin_nm * 1e-9),
                kt=kt, tau=U.time(args.tau_fs * 1e-15),
                pressure=0.0 if args.no_barostat else U.pressure(1.0e5),
                tau_p=0.0 if args.no_barostat else U.time(1.0e-12),
                kappa=U.inv_pressure(4.5e-10))
    c.set_velocities(v)
    c.step(2)
    t0 = time.perf_counter()
    c.step(args.cpu_steps)
    t = time.perf_counter() - t0
    _, _, st = c.state()
    c.close()
    threads = int(st["threads"])
    ns_day = args.dt_fs * 1e-6 * (args.cpu_steps / t) * 86400.0
    return {"value": round(ns_day, 4), "unit": "ns/day", "cores": threads, "kind": "port",
            "ms_per_step": round(t / args.cpu_steps * 1e3, 2),
            "atom_steps_per_s": round(system.n * args.cpu_steps / t, 1),
            "sample": f"{args.cpu_steps} MD steps (after 2 warm-up steps) of the same "
                      f"{system.n}-atom system from the GPU run's final state, "
                      f"oracle/cpu_step.c (f32, FP64 reductions, OpenMP) on {threads} threads "
                      f"of {os.cpu_count()} logical CPUs ({_cpu_model()}): {t:.2f} s, "
                      f"{int(st['rebuilds'])} list builds"}


def fp64_reference(system):
    """FP64 oracle forces, reciprocal energy and virial of the benchmark's
    initial state (for the precision block; not timed)."""
    from oracle import oracle as O

    F, E, W, band = O.pair_forces(system.posq, system.types, system.lj.as_dict(),
                                  system.lattice,
                                  coul=dict(alpha=system.spme.alpha, r_cut=system.spme.r_cut),
                                  excl=system.exclusion_csr(), band=1e-6)
    Fr, Er, Wr, er, wr = O.spme_recip(system.posq, system.lattice, system.spme.grid,
                                      system.spme.order, system.spme.alpha)
    return {"F": F + Fr, "Fr": Fr, "e_recip": er, "w_recip": wr, "band": band}


def precision_check(system, ref):
    """Force / reciprocal-energy / virial error of both reciprocal transforms
    (FP32 radix FFT, FP16 split matrix-core DFT) against the FP64 oracle on
    the benchmark's initial state."""
    import numpy as np
    from heddlemd_amd.engine import engine_for

    out = {}
    F = ref["F"]
    frms = float(np.sqrt(np.mean(np.sum(F ** 2, axis=1))))
    for mode, name in ((0, "fp32_fft"), (1, "fp16_split_dft")):
        e = engine_for(system, dt=1.0)
        e.set_spme_dft(mode)
        e.force_eval(scalars=True)
        f = e.forces().astype(np.float64)
        sc = e.scalars()
        e.close()
        # atoms of a pair whose r^2 lies within 1e-6 of a cutoff^2 are left
        # out: which side of the cutoff such a pair falls depends on rounding
        ok = ~ref["band"]
        d = np.linalg.norm(f - F, axis=1)[ok]
        fi = np.linalg.norm(F, axis=1)[ok]
        out[name] = {"force_err_rms_rel": float(np.sqrt(np.mean(d ** 2)) / frms),
                     "force_err_max_rel": float(d.max() / frms),
                     # per atom |dF_i| / max(|F_i|, F_rms): the FP32-mode bar (<= 1e-5)
                     "force_err_max_rel_per_atom": float((d / np.maximum(fi, frms)).max()),
                     "e_recip_rel": float(abs(sc["e_recip"] / ref["e_recip"] - 1.0)),
                     "w_recip_rel": float(abs(sc["w_recip"] / ref["w_recip"] - 1.0))}
    return out


def batched_replicas(args):
    """Batched-replica workload: uneven argon and SPC water replicas with
    their own neighbour lists, SPME grids and Langevin thermostats, timed
    as looped single-replica launches (hd_step per replica) and as one
    graph launch per step for all replicas (hd_batch_step)."""
    from heddlemd_amd import systems, units as U
    from heddlemd_amd.engine import Batch, engine_for
    from heddlemd_amd.rng import maxwell_boltzmann

    specs = systems.replica_specs(count=args.replicas)
    t0 = time.time()
    engines = []
    for r, sp in enumerate(specs):
        s = systems.replica_system(sp)
        kt = U.temperature(sp["temperature_K"])
        v = maxwell_boltzmann(s.mass, kt, seed=100 + r, n_constraints=s.n_constraints)
        e = engine_for(s, dt=U.time(1e-15), velocities=v,
                       langevin=dict(friction=10.0 / U.time(1e-12), kt=kt, seed=1000 + r))
        e.constrain_velocities()
        engines.append((e, kt, s.n))
    atoms = sum(n for _, _, n in engines)
    log(f"batched: {len(engines)} replicas, {atoms} atoms, set-up {time.time() - t0:.1f} s")
    batch = Batch([e for e, _, _ in engines])
    batch.step(args.batch_equil)  # relax the lattice starts at 1 fs, strong friction
    for r, (e, kt, _) in enumerate(engines):
        e.set_integrator(U.time(args.dt_fs * 1e-15))
        e.set_langevin(1.0 / U.time(1e-12), kt, 1000 + r)
    batch.step(args.warmup)
    lanes, tune = batch.autotune(steps=10)
    log(f"batched: graph shape autotuned to lanes={lanes} "
        f"({', '.join(f'{k}: {v * 100:.1f} ms/step' for k, v in tune.items())})")
    for e, _, _ in engines:
        e.step(2)  # capture each replica's own step graph before timing
    t0 = time.perf_counter()
    for e, _, _ in engines:
        e.step(args.batch_steps)
    t_loop = time.perf_counter() - t0
    t0 = time.perf_counter()
    batch.step(args.batch_steps)
    t_batch = time.perf_counter() - t0
    batch.close()
    for e, _, _ in engines:
        e.close()
    out = {"replicas": len(engines), "atoms": atoms, "steps": args.batch_steps,
           "dt_fs": args.dt_fs,
           "looped_atom_steps_per_s": round(atoms * args.batch_steps / t_loop, 1),
           "batched_atom_steps_per_s": round(atoms * args.batch_steps / t_batch, 1),
           "speedup": round(t_loop / t_batch, 3),
           "graph_lanes": lanes,
           "workload": "half LJ argon (2.1e28 m^-3) / half SPC water (1 g/cm^3), atom counts "
                       "log-uniform in [256, 8192] (seed 11), 280-380 K Langevin (1/ps), "
                       "SPME grid sides from {32, 40, 48, 64} at ~1 Angstrom"}
    log(f"batched: {out}")
    return out


def polymer_shake(args):
    """Polymer melt with every bond SHAKE-constrained (200 chains x 50
    beads, 10,000 united atoms, 9,800 constraints in 49-constraint chains):
    plain SHAKE against the learned warm start.  The predictor is trained on
    the device from the device-resident log of the plain run's own solves."""
    from heddlemd_amd import systems, units as U
    from heddlemd_amd.engine import ShakePredictor, engine_for
    from heddlemd_amd.rng import maxwell_boltzmann

    s = systems.config_polymer()
    kt = U.temperature(400.0)
    v = maxwell_boltzmann(s.mass, kt, seed=5, n_constraints=s.n_constraints)
    e = engine_for(s, dt=U.time(1e-15), velocities=v, nhc_kt=kt, nhc_tau=U.time(1e-13))
    tol = 5e-5
    e.set_shake_tolerance(tol, 1e-8, 500)
    e.constrain_velocities()
    e.force_eval(scalars=False)
    t0 = time.time()
    e.step(args.polymer_equil)  # 1 fs from the packed random-walk start
    e.set_integrator(U.time(args.dt_fs * 1e-15))
    e.step(args.warmup, scalars_last=True)
    ndof = 3 * s.n - s.n_constraints - 3
    log(f"polymer: {s.n} atoms, {s.n_constraints} constraints, relaxed in {time.time() - t0:.1f} s,"
        f" T = {2 * e.scalars()['ke'] / ndof * U.TEMPERATURE_AU_TO_K:.0f} K")
    n_groups = len(s.shake_groups)

    def phase(n):
        e.use_graph(True)
        e.step(2)
        e.sync()
        t0 = time.perf_counter()
        e.step(n, scalars_last=True)
        e.sync()
        el = time.perf_counter() - t0
        st = e.constraint_stats()
        T = 2 * e.scalars()["ke"] / ndof * U.TEMPERATURE_AU_TO_K
        e.use_graph(False)
        e.set_timing(True)
        e.step(50)
        e.sync()
        e.set_timing(False)
        stages = dict(e.stage_times())
        e.use_graph(True)
        return {"ns_per_day": round(args.dt_fs * 1e-6 * (n / el) * 86400.0, 3),
                "ms_per_step": round(el / n * 1e3, 4),
                "mean_sweeps": round(st["shake_sweeps"] / (n_groups * n), 3),
                "max_sweeps": st["shake_max"], "capped": st["shake_capped"],
                "shake_ms": round(stages.get("shake_positions", float("nan")), 4),
                "predict_ms": round(stages.get("shake_predict", 0.0), 4),
                "temperature_K": round(T, 1),
                "stages_ms": {k: round(v, 4) for k, v in stages.items()}}

    plain = phase(args.polymer_steps)
    log(f"polymer plain SHAKE: {plain}")
    t0 = time.perf_counter()
    e.shake_log_start(args.polymer_log)
    e.step(args.polymer_log)
    e.sync()
    e.shake_log_stop()
    t_log = time.perf_counter() - t0
    steps, rows = e.shake_log_size()
    log(f"polymer: logged {steps} steps x {rows} constraints in {t_log:.1f} s")
    m = ShakePredictor(seed=1)
    t0 = time.perf_counter()
    loss = m.train_on_log(e, epochs=args.polymer_epochs, batch=4096, lr=2e-3, seed=2)
    t_train = time.perf_counter() - t0
    log(f"polymer: trained {args.polymer_epochs} epoch(s) in {t_train:.1f} s, loss {loss}")
    e.set_shake_predictor(m)
    learned = phase(args.polymer_steps)
    log(f"polymer learned warm start: {learned}")
    m.close()
    e.close()
    return {
        "workload": "200 chains x 50 united-atom beads (14 amu), 0.8 g/cm^3 (L 6.624 nm), LJ "
                    "sigma 0.395 nm eps 0.382 kJ/mol r_c 1.0 nm, harmonic angles 114 deg "
                    "(520 kJ/mol/rad^2), 1-2/1-3 exclusions, all 9,800 bonds SHAKE "
                    f"(|r^2 - d^2| <= {tol} Bohr^2), NHC 400 K, dt {args.dt_fs} fs",
        "atoms": s.n, "constraints": s.n_constraints, "steps": args.polymer_steps,
        "plain": plain, "learned": learned,
        "training": {"logged_steps": steps, "rows": steps * rows, "epochs": args.polymer_epochs,
                     "batch": 4096, "optimizer": "Adam lr 2e-3", "log_seconds": round(t_log, 2),
                     "train_seconds": round(t_train, 2),
                     "loss_per_epoch": [round(float(x), 6) for x in loss]},
        "sweep_reduction": round(1.0 - learned["mean_sweeps"] / plain["mean_sweeps"], 4),
    }


def traffic_from_profile(kernel_prefix):
    """HBM bytes per launch of a kernel from the committed rocprofv3 pmc
    summary (FETCH_SIZE / WRITE_SIZE scaled by the copy calibration)."""
    path = os.path.join(ROOT, "profiles", "water196k_step_traffic.md")
    try:
        with open(path) as f:
            for line in f:
                if line.startswith("| `" + kernel_prefix):
                    return float(line.rstrip().rstrip("|").split("|")[-1]) * 1e6
    except OSError:
        pass
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--equil", type=int, default=1000,
                    help="untimed equilibration steps at the production dt, tight thermostat")
    ap.add_argument("--min-iters", type=int, default=500,
                    help="steepest-descent iterations at most before the MD equilibration")
    ap.add_argument("--equil-redraws", type=int, default=4,
                    help="50-step blocks at 1 fs, each ended by a velocity redraw")
    ap.add_argument("--equil-tau-fs", type=float, default=100.0)
    ap.add_argument("--tau-fs", type=float, default=100.0)
    ap.add_argument("--dt-fs", type=float, default=2.0)
    ap.add_argument("--temperature", type=float, default=300.0)
    ap.add_argument("--molecules", type=int, default=65536)
    ap.add_argument("--no-barostat", action="store_true")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=20,
                    help="timed steps of the CPU baseline (after 2 warm-up steps)")
    ap.add_argument("--no-roofline-pass", action="store_true")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched-replica pass")
    ap.add_argument("--no-autotune", action="store_true",
                    help="measure at --skin-nm instead of the fastest tuned skin")
    ap.add_argument("--replicas", type=int, default=256)
    ap.add_argument("--batch-steps", type=int, default=500)
    ap.add_argument("--batch-equil", type=int, default=100)
    ap.add_argument("--skin-nm", type=float, default=0.1,
                    help="neighbour-list skin (the reference's default is 0.3 r_c = 0.27 nm; "
                         "a thinner skin evaluates fewer pairs and rebuilds more often)")
    ap.add_argument("--pair-list", type=int, default=0, choices=(0, 1),
                    help="0: packed 64-atom cluster list; 1: per-atom full lists")
    ap.add_argument("--no-polymer", action="store_true",
                    help="skip the polymer-melt SHAKE / learned warm-start pass")
    ap.add_argument("--polymer-equil", type=int, default=1000)
    ap.add_argument("--polymer-steps", type=int, default=5000)
    ap.add_argument("--polymer-log", type=int, default=50000,
                    help="plain-SHAKE steps logged for training the predictor")
    ap.add_argument("--polymer-epochs", type=int, default=1)
    ap.add_argument("--no-split-dft", action="store_true",
                    help="skip the split-precision matrix-core DFT pass")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (the reference kernel that "
                         "scales FETCH_SIZE / WRITE_SIZE counter readings to bytes) and exit")
    args = ap.parse_args()

    if args.copy_calibration:
        import torch

        n = args.copy_calibration * (1 << 20) // 4
        src = torch.arange(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        ok = bool(torch.equal(dst, src))
        print(json.dumps({"copy_calibration_bytes_each_way": n * 4, "dispatches": 3, "ok": ok}),
              flush=True)
        sys.exit(0 if ok else 1)

    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); "
              "multi-GPU domain decomposition is not built", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    from heddlemd_amd import systems, units as U

    if args.molecules == 65536:
        system = systems.config_water_65536()
    else:
        side = int(math.ceil(args.molecules ** (1.0 / 3.0)))
        system = systems.spc_water(args.molecules, side,
                                   grid=(64, 64, 64) if args.molecules <= 16384 else (128,) * 3)
    log(f"system {system.name}: {system.n} atoms, L = {system.lattice[0] * 0.0529177:.4f} nm, "
        f"grid {system.spme.grid}")
    t0 = time.time()
    e, kt = build_engine(system, U, args)
    log(f"setup {time.time() - t0:.1f} s; neighbour list {e.neighbor_stats()}")

    # the equilibration re-parameterises the thermostat and redraws the
    # velocities often; every change drops and re-captures the step graphs
    t0 = time.time()
    equilibrate(e, system, U, kt, args)
    log(f"equilibration {time.time() - t0:.1f} s")
    tuned = None
    if not args.no_autotune:
        e, tuned = autotune_skin(e, system, U, kt, args)
    e.step(args.warmup, scalars_last=True)
    e.sync()
    reb0 = e.scalars()["rebuilds"]

    # ---- timed region ------------------------------------------------------
    t0 = time.perf_counter()
    e.step(args.steps, scalars_last=True)
    e.sync()
    elapsed = time.perf_counter() - t0
    sc = e.scalars()
    ms_per_step = elapsed / args.steps * 1e3
    ns_day = args.dt_fs * 1e-6 * (args.steps / elapsed) * 86400.0
    atom_steps = system.n * args.steps / elapsed
    rebuilds = int(sc["rebuilds"] - reb0)
    T = 2 * sc["ke"] / (3 * system.n - system.n_constraints - 3) * U.TEMPERATURE_AU_TO_K
    log(f"timed {args.steps} steps: {elapsed:.3f} s, {ms_per_step:.3f} ms/step, {ns_day:.2f} ns/day,"
        f" T = {T:.1f} K, rebuilds {rebuilds}")

    result = {
        "metric": METRIC, "value": round(ns_day, 3), "unit": "ns/day", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": (f"SPC water {system.n // 3} molecules ({system.n} atoms), SPME "
                         f"{'x'.join(map(str, system.spme.grid))} order 4, r_c 0.9 nm, SETTLE, "
                         f"NHC 300 K" + ("" if args.no_barostat else " + C-rescale 1 bar")
                         + f", dt {args.dt_fs} fs, neighbour skin {args.skin_nm} nm"
                         + (" (autotuned)" if tuned else "")),
            "atoms": system.n, "parallelism": "single GPU",
        },
        "atom_steps_per_s": round(atom_steps, 1),
        "neighbour_rebuilds_in_timed_region": rebuilds,
        "temperature_K": round(T, 2),
    }
    if tuned:
        result["autotune"] = tuned

    # ---- roofline pass: same steps, no graph, hipEvents per stage ------------
    if not args.no_roofline_pass:
        e.use_graph(False)
        e.set_timing(True)
        t0 = time.perf_counter()
        e.step(args.steps, scalars_last=True)
        e.sync()
        el2 = time.perf_counter() - t0
        e.set_timing(False)
        stages = e.stage_times()
        pairs = e.count_pairs()
        pair_ms = dict(stages).get("pair_force")
        flops = pairs * FLOPS_PER_PAIR
        achieved = flops / (pair_ms * 1e-3) / 1e12
        result["roofline"] = {
            "kernel": "k_pair<2, 2, false> (fused LJ + real-space Ewald forces and virial, "
                      "packed wave64 cluster list; FP32 vector ALU, no matrix-core form)",
            "bound": "fp32-valu", "achieved": round(achieved, 3), "peak": PEAK_FP32_TFLOPS,
            "unit": "TFLOP/s", "frac": round(achieved / PEAK_FP32_TFLOPS, 4),
            "traffic": traffic_from_profile(PAIR_KERNEL),
            "traffic_source": "profiles/water196k_step_traffic.md (rocprofv3 --pmc, calibrated)",
            "useful_pairs": pairs, "flops_per_pair": FLOPS_PER_PAIR,
            "mean_launch_ms": round(pair_ms, 4),
            "timing": f"hipEvents on the context stream over {args.steps} non-graph steps "
                      f"({el2 / args.steps * 1e3:.3f} ms/step in that pass)",
        }
        result["stages_ms"] = {k: round(v, 4) for k, v in stages}
        log(f"stages: {result['stages_ms']}")
    # ---- precision form: the same production steps with the split-precision
    # matrix-core DFT in place of the FP32 radix FFT ---------------------------
    if not args.no_split_dft:
        e.use_graph(True)
        e.set_spme_dft(1)
        e.step(args.warmup, scalars_last=True)
        e.sync()
        t0 = time.perf_counter()
        e.step(args.steps, scalars_last=True)
        e.sync()
        el3 = time.perf_counter() - t0
        e.use_graph(False)
        e.set_timing(True)
        e.step(min(args.steps, 100), scalars_last=True)
        e.sync()
        e.set_timing(False)
        st3 = dict(e.stage_times())
        e.set_spme_dft(0)
        result["split_dft"] = {
            "ns_per_day": round(args.dt_fs * 1e-6 * (args.steps / el3) * 86400.0, 3),
            "ms_per_step": round(el3 / args.steps * 1e3, 4),
            "spme_recip_ms": round(st3.get("spme_recip", float("nan")), 4),
            "fp32_fft_spme_recip_ms": round(result.get("stages_ms", {}).get("spme_recip", float("nan")), 4),
        }
        log(f"split-precision DFT: {result['split_dft']}")
    if not args.no_cpu_baseline:
        import dataclasses
        state = dataclasses.replace(system, posq=e.positions(), lattice=e.lattice())
        v_final = e.velocities()
    e.close()
    if not args.no_batched:
        result["batched"] = batched_replicas(args)
    if not args.no_polymer:
        result["polymer_shake"] = polymer_shake(args)
    if not args.no_cpu_baseline:
        log("cpu baseline (the step restated for the host's cores) ...")
        result["cpu_baseline"] = cpu_baseline(state, v_final, args)
        log(f"cpu baseline: {result['cpu_baseline']}")
        log("precision check against the FP64 oracle ...")
        result["precision"] = precision_check(system, fp64_reference(system))
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
