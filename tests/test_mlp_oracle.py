"""Pins of the surrogate's CPU oracle: backward against central finite differences of the
loss, features against an independent brute-force neighbour search, the macro-step inputs
(one-pass drift, noise sum) against the oracle's own Euler steps."""
import math

import numpy as np

from cil_cuda_amd import abi, configs
from oracle import mlp_oracle as M
from oracle.oracle import Oracle


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


def _mobility_dt(p, col):
    out = []
    for s, sp in enumerate(col.species):
        n = sp["ncells"]
        area = math.pi * (sp["sigma"][:n, 0] / 2) ** 2
        out.append(p.dt / (p.species[s].zeta[0] * area))
    return np.concatenate(out)


def _keys(col):
    return [((s << 28) | c) << 1 for s, sp in enumerate(col.species) for c in range(sp["ncells"])]


def test_noise_sum_is_the_oracle_steps_displacement():
    """Discs too far apart to interact: 8 oracle Euler steps move each by its FDT noise sum."""
    p, box, col = configs.bulk_3d_mixed(n=8, phi=1e-4, seed=3, kT=0.05, dt=1e-4, precision=abi.PREC_FP64)
    o = Oracle(p, box)
    o.forces(col)
    assert np.all(col.flat("frc") == 0.0)
    x0 = col.flat("pos").reshape(-1, 3).copy()
    o.step(col, M.NSUB, 40)
    d = col.flat("pos").reshape(-1, 3) - x0
    L = np.array(box.length[:3])
    d -= L * np.round(d / L)
    temp = _mobility_dt(p, col)
    kT = np.concatenate([np.full(sp["ncells"], p.species[s].kT) for s, sp in enumerate(col.species)])
    z = M.fdt_noise_sum(p.seed, _keys(col), np.sqrt(2 * kT * temp), 40)
    assert np.abs(z).min() > 0
    assert np.allclose(d, z, rtol=0, atol=1e-12)  # positions ~ L carry ~1e-14 rounding per step
    # the draws of other steps do not match (the step index enters the counter)
    assert not np.allclose(d, M.fdt_noise_sum(p.seed, _keys(col), np.sqrt(2 * kT * temp), 41))
    o.close()


def test_one_pass_drift_is_one_euler_step_of_nsub_dt():
    """Features 66-68 x BASE_SCALE equal the displacement of one noiseless oracle step of 8 dt."""
    p, box, col = configs.bulk_3d_mixed(n=300, phi=0.3, seed=4, kT=0.0, dt=1e-4, precision=abi.PREC_FP64)
    p.noise = abi.NOISE_NONE
    o = Oracle(configs.relaxation_params(p), box)  # random start: push overlapping discs apart first
    o.forces(col)
    o.step(col, 100, 0)
    o.close()
    o = Oracle(p, box)
    o.forces(col)
    frc = col.flat("frc").reshape(-1, 3).copy()
    feats = M.macro_inputs(frc, _mobility_dt(p, col), np.zeros_like(frc))
    o.close()
    q = abi.Params.from_buffer_copy(p)
    q.dt = M.NSUB * p.dt
    o = Oracle(q, box)
    x0 = col.flat("pos").reshape(-1, 3).copy()
    o.step(col, 1, 0)
    d = col.flat("pos").reshape(-1, 3) - x0
    L = np.array(box.length[:3])
    d -= L * np.round(d / L)
    assert np.abs(d).max() > 0
    assert np.allclose(feats[:, :3] * M.BASE_SCALE, d, rtol=1e-10, atol=1e-14)
    assert np.all(feats[:, 3:] == 0)
    o.close()
