"""Drift-surrogate workflow over the C ABI: training from plain trajectories and the noise
floor of a predictor that sees only the macro-step's start state.

The surrogate replaces 8 Euler steps by x += b + C + z: b = 8 dt mu F(x0) the one-pass drift
(one force pass), z the exact sum of the 8 noise increments, and C the network's correction,
predicted from the neighbourhood of x0, b and z.  The plain method's deterministic displacement
over those 8 steps, sum_k dt mu F(x_k), depends on the noise through the intermediate positions
x_k, so no function of x0 alone can match it better than its spread over noise realisations.
`noise_floor` measures that spread by running the same start state with two noise seeds:
E|D_a - D_b|^2 / 2 is the conditional variance, and its square root is the smallest RMS error
an x0-only predictor can reach.  The network also sees z, which carries the part of that spread
that is linear in the noise sum, so it can go below this floor.
"""
from __future__ import annotations

import numpy as np

from . import abi
from .colony import HipMD


def train(md: HipMD, ts: int, n_samples: int, nmacro: int, steps: int, batch: int = 8192, lr: float = 1e-3,
          init_seed: int = 1, log=None):
    """Collect `n_samples` (>= 1 per particle row) from `nmacro` plain macro-steps and run
    `steps` Adam steps (70 % at lr, 30 % at lr / 5).  Returns (ts, losses)."""
    n = sum(md.num_cells()) * md.nelem
    stride = max(1, (n * nmacro) // max(1, n_samples))
    ts, got = md.surrogate_collect(nmacro, stride, n_samples, ts)
    if log:
        log(f"surrogate: {got} samples from {nmacro} plain macro-steps (stride {stride})")
    md.mlp_init(init_seed)
    k1 = int(steps * 0.7)
    losses = np.concatenate([md.surrogate_train(k1, batch=batch, lr=lr, seed=init_seed + 1),
                             md.surrogate_train(steps - k1, batch=batch, lr=lr / 5, seed=init_seed + 2)])
    return ts, losses


def noise_floor(params: abi.Params, box: abi.Box, colony, ts: int = 0, seeds=(101, 202)):
    """RMS (per particle, length units) of the deterministic 8-step displacement's spread over
    noise realisations from the same start state, and the RMS of that displacement itself.
    (The training targets are that displacement minus the one-pass drift b, the same for both
    seeds; b is added back from the features, BF16-rounded.)"""
    T = []
    for seed in seeds:
        q = abi.Params.from_buffer_copy(params)
        q.seed = seed
        md = HipMD(q, box, [sp["capacity"] for sp in colony.species])
        try:
            md.upload(colony)
            md.forces(want=False)
            md.surrogate_collect(1, 1, colony.total_cells(), ts)
            X, t = md.surrogate_get_samples()
            T.append(t.astype(np.float64) + abi.SURR_BASE_SCALE * X[:, 66:69].astype(np.float64))
        finally:
            md.close()
    d = T[0] - T[1]
    floor = float(np.sqrt(np.mean((d * d).sum(1)) / 2.0))
    drift = float(np.sqrt(np.mean(0.5 * ((T[0] ** 2).sum(1) + (T[1] ** 2).sum(1)))))
    return floor, drift


def noise_order_floor(params: abi.Params, box: abi.Box, colony, ts: int = 0):
    """RMS (per particle, length units) of the part of the 8-step drift that the macro-step's
    inputs do not determine: the order of the 8 FDT increments.

    8 plain steps run from the same start state twice, with the same increments in step order
    and in reverse order (step indices ts+7 .. ts: the counter draws are keyed by the index).
    The noise sum z and the one-pass drift b are the same for both, so half the RMS difference
    of the end positions measures how the drift depends on the order.  For a linear force field
    F = F0 + K x it is exactly the spread of the drift given (x0, z): the drift's noise term is
    dt mu K sum_k (7 - k) eta_k, whose part beyond its mean given z, dt mu K sum_k (3.5 - k) eta_k,
    changes sign under the reversal.  No predictor of (x0, b, z) can be more accurate."""
    ends = []
    for order in (1, -1):
        md = HipMD(params, box, [sp["capacity"] for sp in colony.species])
        try:
            md.upload(colony)
            md.forces(want=False)
            if order == 1:
                md.step(8, ts)
            else:
                for k in range(7, -1, -1):
                    md.step(1, ts + k)
            c = colony.copy()
            md.download(c)
            ends.append(c.flat("pos").reshape(-1, 3).astype(np.float64))
        finally:
            md.close()
    L = np.array(box.length[:3], dtype=np.float64)
    d = ends[0] - ends[1]
    d -= L * np.round(d / L)
    return float(np.sqrt(np.mean((d * d).sum(1))) / 2.0)


def base_predictor_errors(params: abi.Params, box: abi.Box, colony, ts: int = 0, alpha: float = 3.5 / 8):
    """RMS errors (per particle, length units) against 8 plain steps from `colony` of three
    network-free predictors of the 8-step displacement, each plus the exact noise sum z:
      - zero drift: x0 + z;
      - one Euler step of 8 dt (the surrogate's base b): x0 + b + z, b = 8 dt mu F(x0);
      - midpoint: x0 + b_m + z, b_m = 8 dt mu F(x0 + alpha (b + z)), a second force pass.
    For a linear force field the drift's mean given z is 8 dt mu F(x0 + (3.5 / 8) (b + z)) to first
    order, so alpha = 3.5 / 8 leaves only the order term that noise_order_floor measures."""
    import math

    caps = [sp["capacity"] for sp in colony.species]
    L = np.array(box.length[:3], dtype=np.float64)

    def mi(d):
        return d - L * np.round(d / L)

    x0 = colony.flat("pos").reshape(-1, 3).astype(np.float64)
    sig = colony.flat("sigma").reshape(-1)
    spec = np.concatenate([np.full(sp["ncells"], s) for s, sp in enumerate(colony.species)])
    zeta = np.array([params.species[s].zeta[0] for s in spec])
    mudt = params.dt / (zeta * math.pi * (sig / 2) ** 2)

    def forces(col):
        md = HipMD(params, box, caps)
        try:
            md.upload(col)
            return md.forces()[0]
        finally:
            md.close()

    def run8(p, col):
        md = HipMD(p, box, caps)
        try:
            md.upload(col)
            md.forces(want=False)
            md.step(8, ts)
            c = col.copy()
            md.download(c)
            return c.flat("pos").reshape(-1, 3).astype(np.float64)
        finally:
            md.close()

    b = 8 * mudt[:, None] * forces(colony)
    free = abi.Params.from_buffer_copy(params)  # the same draws with every law off: z
    for i in range(abi.MAX_TYPES):
        for j in range(abi.MAX_TYPES):
            free.law[i][j] = abi.LAW_NONE
    z = mi(run8(free, colony) - x0)
    x8 = run8(params, colony)
    cm = colony.copy()
    xm = x0 + alpha * (b + z)
    for s, sp in enumerate(cm.species):
        sp["pos"][: sp["ncells"], 0] = np.mod(xm[spec == s], L)
    bm = 8 * mudt[:, None] * forces(cm)

    def err(pred):
        return float(np.sqrt(np.mean((mi(pred - x8) ** 2).sum(1))))

    return {"zero_drift": err(x0 + z), "euler_8dt": err(x0 + b + z), "midpoint": err(x0 + bm + z)}

