"""CPU oracle of the drift surrogate (TEST INFRASTRUCTURE ONLY, numpy FP64).

Restates what the HIP surrogate kernels compute, so they can be checked:
  - features: for each single-disc particle, its 16 nearest neighbours (minimum image,
    ties broken by particle index) in ascending distance, each as (u_x m, u_y m, u_z m, 20 g)
    with u = d / r the unit vector towards the neighbour, m = sigma_j for species 0 and
    sigma_j + 4 for species 1, g = r / sigma_bar - 1 (sigma_bar = (sigma_i + sigma_j) / 2);
    then sigma_i, species_i (features 64, 65); the macro-step's one-pass drift
    b = nsub dt mu F(x0) (features 66-68) and the sum z of the FDT noise displacements of its
    nsub steps (69-71), both in units of BASE_SCALE -> 72 features.  (The neighbour part is a
    bijective re-expression of (dx, dy, dz, sigma_j, species_j) that BF16 stores to ~0.4 %
    relative.)
  - MLP 72 -> 128 -> 128 -> 3, SiLU after the hidden layers, output scaled by OUT_SCALE: the
    correction to b, so the macro-step moves x0 + b + OUT_SCALE y + z;
  - loss: mean over samples and components of (y - t)^2 / 2 on the scaled target;
  - backward and an Adam step (Kingma & Ba 2015).
The surrogate has no reference counterpart: parity of these functions is pinned by finite
differences (backward) and by the GPU kernels agreeing with them, not by reference vectors.
"""
from __future__ import annotations

import numpy as np

NNB = 16
NFEAT = 72
HIDDEN = 128
NOUT = 3
OUT_SCALE = 0.002
BASE_SCALE = 0.01
NSUB = 8
ST_FDT = 2


def knn_features(pos, sigma, species, box_len, k=NNB):
    """Features of single-disc particles in a periodic cubic-or-rectangular 3D box."""
    from scipy.spatial import cKDTree

    L = np.asarray(box_len, dtype=float)
    x = np.mod(pos, L)
    tree = cKDTree(x, boxsize=L)
    # ask for a few extra neighbours so distance ties can be ordered by index
    d, j = tree.query(x, k=k + 4)
    n = len(x)
    feats = np.zeros((n, NFEAT))
    for i in range(n):
        dj = d[i][1:]
        jj = j[i][1:]
        order = np.lexsort((jj, dj))[:k]
        jj = jj[order]
        dx = x[jj] - x[i]
        dx -= L * np.round(dx / L)
        feats[i, : 4 * k] = encode_neighbours(dx, sigma[i], sigma[jj], species[jj]).reshape(-1)
        feats[i, 64] = sigma[i]
        feats[i, 65] = species[i]
    return feats


def macro_inputs(frc, mobility_dt, noise, nsub=NSUB):
    """Features 66-71 of particles with forces frc (n, 3), dt / (zeta area) mobility_dt (n,)
    and noise sums (n, 3)."""
    b = nsub * mobility_dt[:, None] * frc
    return np.concatenate([b, noise], axis=1) / BASE_SCALE


def fdt_noise_sum(seed, keys, amp, ts0, nsub=NSUB, dim=3):
    """Sum over steps ts0..ts0+nsub-1 of the FDT displacements amp * z of particles with
    counter keys `keys` ((cell key << 1) | disc), z the first `dim` of the four normals of the
    step's draws (sub 0 and 1), as the oracle's step draws them (cil_oracle.c, FDT noise)."""
    from oracle import oracle as O

    out = np.zeros((len(keys), dim))
    for i, k in enumerate(keys):
        for t in range(ts0, ts0 + nsub):
            z = O.philox_normal2(seed, k, t, ST_FDT, 0) + O.philox_normal2(seed, k, t, ST_FDT, 1)
            out[i] += amp[i] * np.array(z[:dim])
    return out


SPECIES1_SHIFT = 4.0
GAP_SCALE = 20.0


def encode_neighbours(dx, sigma_i, sigma_j, species_j):
    """(k, 3) displacements -> (k, 4) rows (u m, g) as described in the module header."""
    r = np.sqrt((dx * dx).sum(1))
    m = np.where(species_j == 0, sigma_j, sigma_j + SPECIES1_SHIFT)
    sc = np.where(r > 0, m / np.where(r > 0, r, 1.0), 0.0)
    g = GAP_SCALE * (r / (0.5 * (sigma_i + sigma_j)) - 1.0)
    return np.concatenate([dx * sc[:, None], g[:, None]], axis=1)


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def init_params(seed=0):
    rng = np.random.default_rng(seed)
    p = {
        "W1": rng.normal(0, np.sqrt(2.0 / NFEAT), (HIDDEN, NFEAT)),
        "b1": np.zeros(HIDDEN),
        "W2": rng.normal(0, np.sqrt(2.0 / HIDDEN), (HIDDEN, HIDDEN)),
        "b2": np.zeros(HIDDEN),
        "W3": rng.normal(0, np.sqrt(1.0 / HIDDEN), (NOUT, HIDDEN)),
        "b3": np.zeros(NOUT),
    }
    return p


def forward(p, X):
    z1 = X @ p["W1"].T + p["b1"]
    h1 = silu(z1)
    z2 = h1 @ p["W2"].T + p["b2"]
    h2 = silu(z2)
    y = h2 @ p["W3"].T + p["b3"]
    return y, (X, z1, h1, z2, h2)


def loss_and_grad(p, X, T):
    """T: targets already divided by OUT_SCALE.  Returns loss and gradients."""
    y, (X, z1, h1, z2, h2) = forward(p, X)
    n = X.shape[0]
    dy = (y - T) / (n * NOUT)
    loss = 0.5 * np.sum((y - T) ** 2) / (n * NOUT)
    g = {}
    g["W3"] = dy.T @ h2
    g["b3"] = dy.sum(0)
    dh2 = dy @ p["W3"]
    dz2 = dh2 * dsilu(z2)
    g["W2"] = dz2.T @ h1
    g["b2"] = dz2.sum(0)
    dh1 = dz2 @ p["W2"]
    dz1 = dh1 * dsilu(z1)
    g["W1"] = dz1.T @ X
    g["b1"] = dz1.sum(0)
    return loss, g


def adam_step(p, g, state, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
    t = state.get("t", 0) + 1
    state["t"] = t
    for k in p:
        m = state.setdefault("m_" + k, np.zeros_like(p[k]))
        v = state.setdefault("v_" + k, np.zeros_like(p[k]))
        m[:] = b1 * m + (1 - b1) * g[k]
        v[:] = b2 * v + (1 - b2) * g[k] ** 2
        mh = m / (1 - b1 ** t)
        vh = v / (1 - b2 ** t)
        p[k] = p[k] - lr * mh / (np.sqrt(vh) + eps)
    return p
