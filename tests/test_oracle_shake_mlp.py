"""Pins the oracle's SHAKE predictor pieces (no reference counterpart):
the warm-started SHAKE and its multiplier output against closed forms, the
feature rows against geometry and the linearised constraint system, and
the FP64 network's analytic gradient against finite differences."""
import numpy as np

M_BEAD = np.float32(14 * 1822.888)
D_BOND = 2.89


def _chain(n, seed, amp):
    rng = np.random.default_rng(seed)
    x = np.zeros((n, 3))
    for k in range(1, n):
        u = rng.normal(size=3)
        x[k] = x[k - 1] + D_BOND * u / np.linalg.norm(u)
    snap = np.zeros((n, 4), np.float32)
    snap[:, :3] = x
    posq = snap.copy()
    posq[:, :3] += (amp * rng.normal(size=(n, 3))).astype(np.float32)
    return snap, posq


def _groups(oracle, n_chains, n):
    return oracle.ShakeGroups([(np.arange(c * n, (c + 1) * n),
                                [(k, k + 1, D_BOND) for k in range(n - 1)])
                               for c in range(n_chains)])


L_BIG = np.array([1e4, 1e4, 1e4, 0, 0, 0], np.float32)


def test_converged_nu_of_one_bond_is_the_quadratic_root(oracle):
    # one constraint: x_i - x_j moves to r - nu g with |r - nu g| = d, the
    # smaller root of |g|^2 nu^2 - 2 (g.r) nu + |r|^2 - d^2 = 0
    rng = np.random.default_rng(1)
    for trial in range(20):
        snap = np.zeros((2, 4), np.float32)
        u = rng.normal(size=3)
        snap[1, :3] = D_BOND * u / np.linalg.norm(u)
        posq = snap.copy()
        posq[:, :3] += rng.normal(0, 0.05, size=(2, 3)).astype(np.float32)
        G = oracle.ShakeGroups([([0, 1], [(0, 1, D_BOND)])])
        mass = np.array([M_BEAD, 2 * M_BEAD], np.float32)
        v = np.zeros((3, 2), np.float32)
        p = posq.copy()
        nu = np.zeros(1, np.float32)
        oracle.shake_positions(G, p, v, snap, mass, L_BIG, 1.0, tol=1e-6, max_iter=100,
                               nu_out=nu)
        g = snap[0, :3].astype(np.float64) - snap[1, :3]
        r = posq[0, :3].astype(np.float64) - posq[1, :3]
        d2 = float(G.r2[0])
        a, b, c = g @ g, g @ r, r @ r - d2
        root = (b - np.sqrt(b * b - a * c)) / a
        assert abs(nu[0] - root) <= 2e-5 * max(abs(root), 1e-3)
        # the first Gauss-Seidel multiplier is feature 0
        f = oracle.shake_features(G, posq, snap, mass, L_BIG)
        p1 = posq.copy()
        nu1 = np.zeros(1, np.float32)
        oracle.shake_positions(G, p1, v.copy(), snap, mass, L_BIG, 1.0, tol=1e-12, max_iter=1,
                               nu_out=nu1)
        assert abs(nu1[0] - f[0, 0]) <= 1e-6 * abs(f[0, 0])


def test_features_of_a_uniformly_stretched_straight_chain(oracle):
    # collinear beads stretched by (1 + e): nu1 = ((1+e)^2 - 1) / (2 (1+e)),
    # couplings to both neighbours -1/2, rho = e, zero past the ends; the
    # previous multipliers are passed through
    n, e = 12, 0.01
    snap = np.zeros((n, 4), np.float32)
    snap[:, 0] = np.arange(n) * D_BOND
    posq = snap.copy()
    posq[:, 0] *= np.float32(1 + e)
    G = _groups(oracle, 1, n)
    mass = np.full(n, M_BEAD, np.float32)
    prev = np.linspace(-0.01, 0.02, n - 1).astype(np.float32)
    f = oracle.shake_features(G, posq, snap, mass, L_BIG, nu_prev=prev).astype(np.float64)
    nu1 = ((1 + e) ** 2 - 1) / (2 * (1 + e))
    nc = n - 1
    # |r|^2 - d^2 cancels ~2 of f32's 7 digits at these coordinates
    tol = 3e-4
    assert np.allclose(f[:, 0], nu1, rtol=tol)
    assert np.allclose(f[1:, 1], nu1, rtol=tol) and f[0, 1] == 0
    assert np.allclose(f[:-1, 2], nu1, rtol=tol) and f[-1, 2] == 0
    assert np.array_equal(f[:, 3], prev.astype(np.float64))
    # second neighbours summed: two inside, one within two of an end
    assert np.allclose(f[2:-2, 4], 2 * nu1, rtol=tol)
    assert np.allclose(f[[0, 1, -2, -1], 4], nu1, rtol=tol)
    assert np.allclose(f[1:, 5], -0.5, rtol=1e-5) and f[0, 5] == 0
    assert np.allclose(f[:-1, 6], -0.5, rtol=1e-5) and f[nc - 1, 6] == 0
    assert np.allclose(f[:, 7], e, rtol=1e-3)


def test_converged_multipliers_solve_the_linearised_system(oracle):
    # the couplings define nu_k + sum_l kappa_kl nu_l = nu1_k up to terms of
    # second order in the displacement: the residual falls ~4x when the
    # displacement halves (a wrong sign, mass or neighbour makes it first
    # order)
    n = 30
    G = _groups(oracle, 1, n)
    mass = np.full(n, M_BEAD, np.float32)
    mass[::3] *= 2  # unequal masses exercise the 1/m factors
    res = []
    for amp in (0.04, 0.02):
        snap, posq = _chain(n, 5, amp)
        f = oracle.shake_features(G, posq, snap, mass, L_BIG).astype(np.float64)
        p = posq.copy()
        nu = np.zeros(n - 1, np.float32)
        # 5e-5: the f32 floor of |r|^2 - d^2 at ~30 Bohr coordinates
        _, it = oracle.shake_positions(G, p, np.zeros((3, n), np.float32), snap, mass, L_BIG,
                                       1.0, tol=5e-5, max_iter=2000, nu_out=nu)
        assert it[0] < 2000
        nu = nu.astype(np.float64)
        lhs = nu.copy()
        lhs[1:] += f[1:, 5] * nu[:-1]
        lhs[:-1] += f[:-1, 6] * nu[1:]
        r = np.abs(lhs - f[:, 0]).max()
        assert r < 0.05 * np.abs(f[:, 0]).max()
        res.append(r)
    assert 3.0 < res[0] / res[1] < 5.5


def test_warm_start_from_zero_is_plain_shake_and_from_the_solution_converges(oracle):
    n, nch = 50, 6
    G = _groups(oracle, nch, n)
    mass = np.full(n * nch, M_BEAD, np.float32)
    snaps, posqs = zip(*[_chain(n, 10 + c, 0.03) for c in range(nch)])
    snap = np.concatenate(snaps)
    posq = np.concatenate(posqs)
    tol = 5e-5
    p0, v0 = posq.copy(), np.zeros((3, n * nch), np.float32)
    nu = np.zeros(nch * (n - 1), np.float32)
    w0, it0 = oracle.shake_positions(G, p0, v0, snap, mass, L_BIG, 82.7, tol=tol, max_iter=500,
                                     nu_out=nu)
    p1, v1 = posq.copy(), np.zeros((3, n * nch), np.float32)
    nu_z = np.zeros_like(nu)
    w1, it1 = oracle.shake_positions(G, p1, v1, snap, mass, L_BIG, 82.7, tol=tol, max_iter=500,
                                     nu0=np.zeros_like(nu), nu_out=nu_z)
    assert np.array_equal(p0, p1) and np.array_equal(v0, v1) and np.array_equal(it0, it1)
    assert np.array_equal(nu, nu_z)
    p2, v2 = posq.copy(), np.zeros((3, n * nch), np.float32)
    nu_w = np.zeros_like(nu)
    _, it2 = oracle.shake_positions(G, p2, v2, snap, mass, L_BIG, 82.7, tol=tol, max_iter=500,
                                    nu0=nu, nu_out=nu_w)
    # from the converged multipliers only f32 rounding is left to sweep out
    assert it0.min() > 8 and it2.max() <= 4 and it2.sum() * 4 < it0.sum()
    assert np.abs(p2 - p0).max() < 1e-4
    # both solves stop within the tolerance: |d sigma| <= 5e-5 moves nu by
    # ~ 5e-5 / (2 d^2) = 3e-6 per bond, a few times that along the chain
    assert np.abs(nu_w - nu).max() < 2e-5


def _random_params(oracle, rng, s=0.01):
    p = rng.normal(0, 0.3, size=oracle.MLP_NPARAM)
    p[-1] = s
    return p


def test_mlp_forward_closed_forms(oracle):
    rng = np.random.default_rng(2)
    x = rng.normal(0, 0.01, size=(7, 8))
    p = np.zeros(oracle.MLP_NPARAM)
    p[-1] = 0.02
    p[oracle.MLP_NTRAIN - 1] = 0.7  # b3
    assert np.allclose(oracle.mlp_forward(p, x), x[:, 0] + 0.02 * 0.7)
    # one path: W1[3][5] (a coupling input), W2[1][3], w3[1]
    W1, b1, W2, b2, w3, b3, s = oracle.mlp_unpack(np.arange(oracle.MLP_NPARAM, dtype=float))
    p = np.zeros(oracle.MLP_NPARAM)
    p[-1] = 0.02
    p[int(W1[3, 5])] = 1.3
    p[int(W2[1, 3])] = -0.8
    p[int(w3[1])] = 2.0
    want = x[:, 0] + 0.02 * 2.0 * np.tanh(-0.8 * np.tanh(1.3 * x[:, 5]))
    assert np.allclose(oracle.mlp_forward(p, x), want)
    # a scaled input: W1[0][7] sees rho / s
    p = np.zeros(oracle.MLP_NPARAM)
    p[-1] = 0.02
    p[int(W1[0, 7])] = 0.5
    p[int(W2[0, 0])] = 1.0
    p[int(w3[0])] = 1.0
    want = x[:, 0] + 0.02 * np.tanh(np.tanh(0.5 * x[:, 7] / 0.02))
    assert np.allclose(oracle.mlp_forward(p, x), want)


def test_mlp_gradient_matches_finite_differences(oracle):
    rng = np.random.default_rng(3)
    p = _random_params(oracle, rng)
    x = rng.normal(0, 0.01, size=(50, 8))
    x[:, 5:7] = rng.uniform(-0.5, 0.5, size=(50, 2))
    t = x[:, 0] + rng.normal(0, 0.005, size=50)
    loss, g = oracle.mlp_loss_grad(p, x, t)
    W1, b1, W2, b2, w3, b3, s = oracle.mlp_unpack(np.arange(oracle.MLP_NPARAM, dtype=float))
    idx = list(rng.choice(oracle.MLP_NTRAIN, 60, replace=False))
    idx += [int(W1[0, 0]), int(W1[5, 6]), int(b1[9]), int(W2[2, 63]), int(b2[0]), int(w3[7]),
            int(b3)]
    h = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (oracle.mlp_loss_grad(pp, x, t)[0] - oracle.mlp_loss_grad(pm, x, t)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g[i])), i
    # the loss is the mean squared scaled residual of the forward pass
    nu = oracle.mlp_forward(p, x)
    assert np.isclose(loss, np.mean(((nu - t) / p[-1]) ** 2))
