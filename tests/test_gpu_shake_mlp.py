"""The learned SHAKE warm start on the device against the oracle: the
network's forward pass, gradient and Adam step (FP64 oracle), the feature
rows and converged multipliers of the step log (f32 restatement, bit for
bit), the warm-started SHAKE (bit for bit given the device's predicted
multipliers), and the sweep count on the polymer melt before and after
training on its own plain-SHAKE log.

Tolerances: network outputs within 2e-5 s (1 + |o|) of the FP64 forward
(f32 sums of 64 terms and device tanhf); gradient within 1e-4 of the
FP64 gradient in norm; the first Adam step within 1e-3 lr of
lr g / (|g| + eps / sqrt(1 - b2))."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from heddlemd_amd import systems, units as U  # noqa: E402
from heddlemd_amd.engine import ShakePredictor, engine_for  # noqa: E402


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _random_params(n, rng, s):
    p = rng.normal(0, 0.3, size=n).astype(np.float32)
    p[-1] = s
    return p


def _rows(rng, n, s):
    x = rng.normal(0, s, size=(n, 8)).astype(np.float32)
    x[:, 5:7] = rng.uniform(-0.5, 0.5, size=(n, 2))
    # a smooth target: the first correction of the linearised chain system
    t = (x[:, 0] - x[:, 5] * x[:, 1] - x[:, 6] * x[:, 2]).astype(np.float32)
    return x, t


def test_predictor_forward_and_gradient_match_fp64(oracle):
    say("case: predictor forward/gradient, 1000 rows")
    rng = np.random.default_rng(11)
    m = ShakePredictor(seed=3)
    s = 0.01
    p = _random_params(m.n_params, rng, s)
    m.params = p
    assert np.array_equal(m.params, p)
    x, t = _rows(rng, 1000, s)
    nu = m.predict(x).astype(np.float64)
    ref = oracle.mlp_forward(p.astype(np.float64), x)
    o = (ref - x[:, 0]) / s
    err = np.abs(nu - ref) / (s * (1 + np.abs(o)))
    say(f"  forward max err {err.max():.2e} (units of s (1 + |o|))")
    assert err.max() < 2e-5
    loss, g = m.gradient(x, t)
    lref, gref = oracle.mlp_loss_grad(p.astype(np.float64), x, t)
    rel = np.linalg.norm(g - gref) / np.linalg.norm(gref)
    say(f"  loss {loss:.6e} vs {lref:.6e}; gradient rel err {rel:.2e}")
    assert loss == pytest.approx(lref, rel=1e-5)
    assert rel < 1e-4
    m.close()


def test_predictor_adam_step_and_deterministic_training(oracle):
    say("case: first Adam step, training determinism")
    rng = np.random.default_rng(12)
    s = 0.01
    x, t = _rows(rng, 4096, s)
    m = ShakePredictor(seed=5)
    p0 = _random_params(m.n_params, rng, s)
    m.params = p0
    lr = 1e-3
    m.train(x, t, epochs=1, batch=4096, lr=lr, seed=1)  # one step over all rows
    dp = (m.params - p0).astype(np.float64)[:-1]
    _, g = oracle.mlp_loss_grad(p0.astype(np.float64), x, t)
    # first Adam step: m = (1-b1) g, v = (1-b2) g^2; with the bias
    # correction the update is lr g / (|g| + eps / sqrt(1 - b2))
    want = -lr * g / (np.abs(g) + 1e-8 / np.sqrt(1 - 0.999))
    big = np.abs(g) > 1e-5
    assert big.sum() > 1000
    assert np.abs(dp - want).max() < 1e-3 * lr
    assert m.params[-1] == s
    m.close()
    runs = []
    for rep in range(2):
        m = ShakePredictor(seed=9)
        loss = m.train(x, t, epochs=40, batch=256, lr=3e-3, seed=4)
        runs.append((m.params, loss))
        m.close()
    say(f"  losses {np.array2string(runs[0][1], precision=3)}")
    assert np.array_equal(runs[0][0], runs[1][0])
    assert np.array_equal(runs[0][1], runs[1][1])
    # (the same schedule in FP64 numpy: 0.158 -> 0.003; the loss stays flat
    # for ~8 epochs while the zero-initialised output layer grows)
    assert runs[0][1][-1] < 0.1 * runs[0][1][0]
    # the scale came from the data: rms of the nu1 column
    assert runs[0][0][-1] == pytest.approx(np.sqrt(np.mean(x[:, 0].astype(np.float64) ** 2)),
                                           rel=1e-6)


def _chains(n_ch=24, n_b=50, seed=23):
    rng = np.random.default_rng(seed)
    d, L = 2.89, 120.0
    pos = []
    for c in range(n_ch):
        xx = rng.uniform(-L / 4, L / 4, size=3)
        chain = [xx]
        for k in range(1, n_b):
            u = rng.normal(size=3)
            chain.append(chain[-1] + d * u / np.linalg.norm(u))
        pos.extend(chain)
    pos = np.asarray(pos)
    n = n_ch * n_b
    posq = np.zeros((n, 4), np.float32)
    posq[:, :3] = systems.wrap_into_box(pos, [L, L, L])
    lj = systems.LJTable(ntypes=1, sigma=np.array([[7.46]]), epsilon=np.array([[0.0]]),
                         cutoff=np.array([[18.9]]), r_switch=np.array([[18.9]]))
    s = systems.System(name="chains", posq=posq, mass=np.full(n, 25520.0, np.float32),
                       types=np.zeros(n, np.uint32),
                       lattice=np.array([L, L, L, 0, 0, 0], np.float64), lj=lj,
                       excl_key=np.repeat(np.arange(n_ch, dtype=np.uint32), n_b),
                       shake_groups=[(list(range(c * n_b, (c + 1) * n_b)),
                                      [(k, k + 1, d) for k in range(n_b - 1)])
                                     for c in range(n_ch)])
    v0 = rng.normal(0, 2.2e-4, size=(3, n)).astype(np.float32)
    return s, v0


def _pre_shake_state(oracle, s, posq0, img0, v0, f0, dt):
    p = posq0.copy()
    img = img0.T.copy()
    v = v0.copy()
    L32 = s.lattice.astype(np.float32)
    snap = p.copy()
    oracle.vv_kick_drift_f32(p, img, v, f0, s.mass, L32, dt)
    return p, v, snap, L32


def test_shake_log_and_warm_started_step_match_oracle(oracle):
    say("case: 24 chains x 50 beads, SHAKE log + warm-started step")
    s, v0 = _chains()
    dt = np.float32(U.time(2e-15))
    tol = 5e-5
    G = oracle.ShakeGroups(s.shake_groups)
    # plain SHAKE with the log on
    e = engine_for(s, dt=float(dt), velocities=v0)
    e.set_shake_tolerance(tol, 1e-8, 200)
    e.use_graph(False)
    e.force_eval(scalars=False)
    f0 = e.forces().T.copy()
    posq0, img0 = e.positions(images=True)
    e.shake_log_start(4)
    e.step(1)
    assert e.shake_log_size() == (1, len(G.r2))
    x_log, t_log = e.shake_log()
    _, nu_dev = e.shake_multipliers()
    # a second step: its rows carry the first step's multipliers
    posq1, img1 = e.positions(images=True)
    v1 = e.velocities()
    e.step(1)
    x2, t2 = e.shake_log(1, 1)
    e.close()
    p2, v2, snap2, L32 = _pre_shake_state(oracle, s, posq1, img1, v1,
                                          np.zeros_like(f0), dt)
    feat2 = oracle.shake_features(G, p2, snap2, s.mass, L32, nu_prev=t_log)
    say(f"  step 2 features: {np.sum(feat2 != x2)} differ")
    assert np.array_equal(feat2, x2)
    assert np.array_equal(x2[:, 3], t_log)
    p, v, snap, L32 = _pre_shake_state(oracle, s, posq0, img0, v0, f0, dt)
    feat = oracle.shake_features(G, p, snap, s.mass, L32)
    say(f"  features: {np.sum(feat != x_log)} of {feat.size} differ, "
        f"max {np.abs(feat - x_log).max():.2e}")
    assert np.array_equal(feat, x_log)
    nu = np.zeros(len(G.r2), np.float32)
    oracle.shake_positions(G, p, v, snap, s.mass, L32, dt, tol=tol, max_iter=200, nu_out=nu)
    assert np.array_equal(nu, t_log) and np.array_equal(nu, nu_dev)
    # a predictor trained on this one step, then the same step warm-started
    m = ShakePredictor(seed=2)
    m.train(x_log, t_log, epochs=40, batch=64, lr=3e-3, seed=3)
    params = m.params
    m.close()
    e = engine_for(s, dt=float(dt), velocities=v0)
    e.set_shake_tolerance(tol, 1e-8, 200)
    e.use_graph(False)
    e.force_eval(scalars=False)
    e.set_shake_predictor(params)
    e.step(1)
    nu0_dev, nu_dev = e.shake_multipliers()
    posq1, _ = e.positions(images=True)
    st = e.constraint_stats()
    e.close()
    ref0 = oracle.mlp_forward(params.astype(np.float64), feat)
    sc = float(params[-1])
    err = np.abs(nu0_dev - ref0).max() / sc
    say(f"  predicted nu0 max err {err:.2e} s; sweeps {st['shake_sweeps']}")
    assert err < 1e-4
    p, v, snap, L32 = _pre_shake_state(oracle, s, posq0, img0, v0, f0, dt)
    nu_w = np.zeros(len(G.r2), np.float32)
    _, it = oracle.shake_positions(G, p, v, snap, s.mass, L32, dt, tol=tol, max_iter=200,
                                   nu0=nu0_dev, nu_out=nu_w)
    assert st["shake_sweeps"] == int(it.sum())
    assert np.array_equal(nu_w, nu_dev)
    assert np.array_equal(p, posq1)


def test_shake_log_in_graph_replays_fills_and_stops():
    say("case: SHAKE log under graph replay, capacity 3 of 5 steps")
    s, v0 = _chains(n_ch=8)
    e = engine_for(s, dt=U.time(2e-15), velocities=v0)
    e.set_shake_tolerance(5e-5, 1e-8, 200)
    e.force_eval(scalars=False)
    e.shake_log_start(3)
    e.step(2)
    assert e.shake_log_size()[0] == 2
    e.step(3)
    steps, rows = e.shake_log_size()
    assert (steps, rows) == (3, 8 * 49)
    x, t = e.shake_log()
    assert np.isfinite(x).all() and np.isfinite(t).all()
    # three different steps were logged
    assert not np.array_equal(t[:rows], t[rows:2 * rows])
    assert not np.array_equal(t[rows:2 * rows], t[2 * rows:])
    e.shake_log_stop()
    e.step(1)
    assert e.shake_log_size()[0] == 3
    # a restart with a different size re-allocates the log: the step graphs
    # must pick up the new buffers
    for cap in (4, 2):
        e.shake_log_start(cap)
        e.step(cap + 1)
        steps, rows = e.shake_log_size()
        assert steps == cap
        x, t = e.shake_log()
        t = t.reshape(cap, rows)
        assert (np.abs(t).max(axis=1) > 0).all()
        assert (np.abs(x).reshape(cap, -1).max(axis=1) > 0).all()
    e.close()


def test_polymer_melt_predictor_cuts_shake_sweeps():
    # united-atom melt with every bond constrained: log plain-SHAKE solves,
    # train the predictor on the device from the log, run on warm-started
    say("case: polymer melt 40 x 50, plain SHAKE log -> train -> warm start")
    s = systems.polymer_melt(n_chains=40, n_beads=50)
    kt = U.temperature(400.0)
    from heddlemd_amd.rng import maxwell_boltzmann
    v = maxwell_boltzmann(s.mass, kt, seed=5, n_constraints=s.n_constraints)
    e = engine_for(s, dt=U.time(1e-15), velocities=v, nhc_kt=kt, nhc_tau=U.time(1e-13))
    tol = 5e-5
    e.set_shake_tolerance(tol, 1e-8, 500)
    e.constrain_velocities()
    e.force_eval(scalars=False)
    e.step(300, scalars_last=True)  # relax the packed start at 1 fs
    ndof = 3 * s.n - s.n_constraints - 3
    say(f"  T after 300 fs: {2 * e.scalars()['ke'] / ndof * U.TEMPERATURE_AU_TO_K:.0f} K")
    e.set_integrator(U.time(2e-15))
    n_groups = 40
    n_steps = 200
    L = s.lattice[0]
    d0 = U.length(0.153e-9)

    def residual():
        # max |r^2 / d^2 - 1| over the 49 x 40 constrained bonds, the
        # quantity the SHAKE tolerance bounds, from the FP64 difference of
        # the unwrapped f32 positions
        p, _ = e.positions(images=True)
        pc = p[:, :3].astype(np.float64).reshape(40, 50, 3)
        r = pc[:, 1:] - pc[:, :-1]
        r -= L * np.floor(r / L + 0.5)
        return float(np.abs((r ** 2).sum(axis=2) / d0 ** 2 - 1).max())

    e.shake_log_start(n_steps)
    e.step(n_steps)
    st0 = e.constraint_stats()
    res0 = residual()
    e.shake_log_stop()
    plain = st0["shake_sweeps"] / (n_groups * n_steps)
    m = ShakePredictor(seed=1)
    loss = m.train_on_log(e, epochs=8, batch=1024, lr=2e-3, seed=2)
    say(f"  log {e.shake_log_size()}, training loss {np.array2string(loss, precision=4)}")
    assert loss[-1] < loss[0]
    e.set_shake_predictor(m)
    e.step(n_steps)
    st1 = e.constraint_stats()
    res1 = residual()
    warm = st1["shake_sweeps"] / (n_groups * n_steps)
    # untrained (first-order) start for comparison
    m0 = ShakePredictor(seed=1)
    m0.params = np.concatenate([np.zeros(m0.n_params - 1, np.float32), [np.float32(1.0)]])
    e.set_shake_predictor(m0)
    e.step(n_steps)
    st2 = e.constraint_stats()
    res2 = residual()
    first = st2["shake_sweeps"] / (n_groups * n_steps)
    e.close()
    say(f"  mean SHAKE sweeps per chain: plain {plain:.2f}, learned {warm:.2f}, "
        f"first-order {first:.2f}; max {st0['shake_max']} / {st1['shake_max']} / "
        f"{st2['shake_max']}")
    say(f"  constraint residual max |r^2/d^2 - 1|: plain {res0:.2e}, learned {res1:.2e}, "
        f"first-order {res2:.2e} (tolerance {tol / d0 ** 2:.2e})")
    assert st0["shake_capped"] == 0 and st1["shake_capped"] == 0 and st2["shake_capped"] == 0
    # the learned start at least halves the sweeps of the plain solve and
    # beats the first-order (previous-step multiplier) start
    assert 2.0 * warm <= plain
    assert warm < first
    # every phase leaves the bonds at the same residual: the start changes
    # the sweep count, not the converged answer (f32 positions hold
    # r^2/d^2 to about 4e-5)
    for res in (res0, res1, res2):
        assert res < 4e-5
