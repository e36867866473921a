"""Pins of the surrogate's CPU oracle: backward against central finite differences of the
loss, features against an independent brute-force neighbour search."""
import numpy as np

from oracle import mlp_oracle as M


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(1)
    p = M.init_params(3)
    X = rng.normal(size=(7, M.NFEAT))
    T = rng.normal(size=(7, M.NOUT))
    loss, g = M.loss_and_grad(p, X, T)
    for key in ("W1", "b1", "W2", "b2", "W3", "b3"):
        a = p[key]
        flat = a.reshape(-1)
        for idx in rng.choice(flat.size, size=min(6, flat.size), replace=False):
            old = flat[idx]
            h = 1e-6
            flat[idx] = old + h
            lp, _ = M.loss_and_grad(p, X, T)
            flat[idx] = old - h
            lm, _ = M.loss_and_grad(p, X, T)
            flat[idx] = old
            fd = (lp - lm) / (2 * h)
            assert abs(fd - g[key].reshape(-1)[idx]) <= 1e-6 * max(1.0, abs(fd)), key


def test_silu_derivative():
    z = np.linspace(-6, 6, 101)
    h = 1e-6
    fd = (M.silu(z + h) - M.silu(z - h)) / (2 * h)
    assert np.allclose(fd, M.dsilu(z), atol=1e-8)


def test_knn_features_against_brute_force():
    rng = np.random.default_rng(2)
    n, L = 300, 6.0
    x = rng.uniform(0, L, (n, 3))
    sig = rng.uniform(0.8, 1.2, n)
    sp = rng.integers(0, 2, n)
    F = M.knn_features(x, sig, sp, [L, L, L])
    for i in (0, 17, 299):
        d = x - x[i]
        d -= L * np.round(d / L)
        r = np.sqrt((d * d).sum(1))
        r[i] = np.inf
        order = np.lexsort((np.arange(n), r))[:16]
        # decode the rows back to (d, sigma_j, species_j) and compare with the brute force
        rows = F[i, :64].reshape(16, 4)
        m = np.sqrt((rows[:, :3] ** 2).sum(1))
        spj = (m > 3.0).astype(int)
        sgj = m - 4.0 * spj
        rj = (rows[:, 3] / M.GAP_SCALE + 1.0) * 0.5 * (sig[i] + sgj)
        dj = rows[:, :3] / m[:, None] * rj[:, None]
        assert np.array_equal(spj, sp[order])
        assert np.allclose(sgj, sig[order])
        assert np.allclose(dj, d[order])
        assert F[i, 64] == sig[i] and F[i, 65] == sp[i]
        assert np.all(F[i, 66:] == 0)


def test_adam_reduces_loss():
    rng = np.random.default_rng(4)
    p = M.init_params(5)
    X = rng.normal(size=(256, M.NFEAT))
    W = rng.normal(size=(M.NOUT, M.NFEAT)) * 0.1
    T = np.tanh(X @ W.T)
    st = {}
    l0, _ = M.loss_and_grad(p, X, T)
    for _ in range(200):
        _, g = M.loss_and_grad(p, X, T)
        M.adam_step(p, g, st, lr=3e-3)
    l1, _ = M.loss_and_grad(p, X, T)
    assert l1 < 0.2 * l0
