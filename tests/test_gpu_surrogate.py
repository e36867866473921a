"""GPU tests of the drift surrogate through the C ABI.

Bars (BF16 operands with FP32 accumulation against the FP64 oracle on the same BF16-rounded
weights and inputs):
  - forward: |y_gpu - y_ref| <= 0.03 * rms(y_ref) + 1e-6 per output (the GPU also rounds
    hidden activations to BF16: ~0.4 % per layer);
  - gradient: per tensor ||g_gpu - g_ref|| <= 0.05 * ||g_ref||, loss within 2 %;
  - Adam: one step moves every parameter whose reference gradient is not tiny by lr
    (first Adam step with bias correction), to 1e-3 * lr, against the gradient's sign for
    >= 99.9 % of them;
  - kNN features 0-65: equal to the oracle's rows rounded to BF16 (one BF16 ulp) for every
    particle without a near-tie (relative gap < 1e-4) among its 17 nearest distances, in
    stencil mode and in neighbour-list mode after macro-steps and plain steps have moved the
    particles away from the list's build (candidate lists, and with a skin that overfills them,
    the fallback search over bins the particles have left);
  - features 66-71 (one-pass drift from the oracle's FP64 forces, noise sum from the oracle's
    draws): one BF16 ulp, plus 1e-3 of the drift's RMS for the FP32 forces;
  - training on plain trajectories (kT = 0.005): the trained network's 8-step displacement
    error is below that of one Euler step of 8 dt (the network's starting point: its
    correction added to that step) and at least 10 % below the zero-drift error.
"""
import sys

import numpy as np
import pytest

from cil_cuda_amd import abi, configs
from cil_cuda_amd.colony import HipMD
from oracle import mlp_oracle as M
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def bf16(x):
    """Round to BF16 (nearest even) and back, as the library does on entry."""
    a = np.ascontiguousarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def flatten(p):
    return np.concatenate([p[k].reshape(-1) for k in ("W1", "b1", "W2", "b2", "W3", "b3")]).astype(np.float32)


def unflatten(v):
    out, o = {}, 0
    for k, shape in (("W1", (M.HIDDEN, M.NFEAT)), ("b1", (M.HIDDEN,)), ("W2", (M.HIDDEN, M.HIDDEN)),
                     ("b2", (M.HIDDEN,)), ("W3", (M.NOUT, M.HIDDEN)), ("b3", (M.NOUT,))):
        n = int(np.prod(shape))
        out[k] = np.asarray(v[o:o + n], dtype=np.float64).reshape(shape)
        o += n
    assert o == abi.MLP_NPARAM
    return out


def rounded(p):
    """BF16 weights (as the MFMA operands) with FP32 biases."""
    q = {k: v.copy() for k, v in p.items()}
    for k in ("W1", "W2", "W3"):
        q[k] = bf16(q[k])
    for k in ("b1", "b2", "b3"):
        q[k] = q[k].astype(np.float32).astype(np.float64)
    return q


def small_ctx(n=3000, seed=5):
    p, box, col = configs.bulk_3d_mixed(n=n, seed=seed, precision=abi.PREC_FP32)
    md = HipMD(p, box, [sp["capacity"] for sp in col.species])
    md.upload(col)
    return md, p, box, col


def test_abi_exports_surrogate(gpu):
    L = abi.lib()
    assert L.cil_hip_mlp_nparams() == abi.MLP_NPARAM == 26243


def random_params(seed):
    rng = np.random.default_rng(seed)
    p = M.init_params(seed)
    for k in ("b1", "b2", "b3"):
        p[k] = rng.normal(0, 0.1, p[k].shape)
    return p


def test_mlp_forward_matches_oracle(gpu):
    md, *_ = small_ctx(600)
    rng = np.random.default_rng(11)
    p = random_params(3)
    n = 1000  # not a multiple of the 32-row tile
    X = rng.normal(0, 1, (n, M.NFEAT))
    md.mlp_set_params(flatten(p))
    y = md.mlp_forward(X).astype(np.float64)
    yr, _ = M.forward(rounded(p), bf16(X))
    yr *= M.OUT_SCALE
    err = np.abs(y - yr).max()
    scale = np.sqrt(np.mean(yr ** 2))
    say(f"forward: max |dy| {err:.3e}, rms(y) {scale:.3e}")
    assert err <= 0.03 * scale + 1e-6
    # a planted transposition of W2 must fail the same bar (the bar separates layouts)
    q = rounded(p)
    q["W2"] = q["W2"].T.copy()
    yt, _ = M.forward(q, bf16(X))
    assert np.abs(y - M.OUT_SCALE * yt).max() > 0.03 * scale + 1e-6
    # round trip of the parameters
    assert np.array_equal(md.mlp_get_params(), flatten(p))
    md.close()


def test_mlp_grad_matches_oracle(gpu):
    md, *_ = small_ctx(600)
    rng = np.random.default_rng(12)
    p = random_params(4)
    n = 777
    X = rng.normal(0, 1, (n, M.NFEAT))
    T = rng.normal(0, 0.01, (n, 3))
    md.mlp_set_params(flatten(p))
    g, loss = md.mlp_grad(X, T)
    lr, gr = M.loss_and_grad(rounded(p), bf16(X), T.astype(np.float32).astype(np.float64) / M.OUT_SCALE)
    gg = unflatten(g)
    say(f"loss gpu {loss:.6e} oracle {lr:.6e}")
    assert abs(loss - lr) <= 0.02 * abs(lr)
    for k in ("W1", "b1", "W2", "b2", "W3", "b3"):
        rel = np.linalg.norm(gg[k] - gr[k]) / np.linalg.norm(gr[k])
        say(f"grad {k}: rel err {rel:.3e}")
        assert rel <= 0.05, k
    md.close()


def test_adam_first_step(gpu):
    md, *_ = small_ctx(600)
    rng = np.random.default_rng(13)
    p = random_params(6)
    x = rng.normal(0, 1, (1, M.NFEAT))
    t = rng.normal(0, 0.01, (1, 3))
    flat = flatten(p)
    md.mlp_set_params(flat)
    md.surrogate_set_samples(x, t)
    lr = 1e-3
    losses = md.surrogate_train(1, batch=64, lr=lr, seed=1)
    after = md.mlp_get_params().astype(np.float64)
    _, gr = M.loss_and_grad(rounded(p), bf16(x), t.astype(np.float32).astype(np.float64) / M.OUT_SCALE)
    g = np.concatenate([gr[k].reshape(-1) for k in ("W1", "b1", "W2", "b2", "W3", "b3")])
    big = np.abs(g) > 1e-3 * np.abs(g).max()
    step = after - flat.astype(np.float64)
    say(f"adam: {big.sum()} parameters with clear gradient sign, loss {losses[0]:.4e}")
    assert big.sum() > 1000
    tol = 1e-3 * lr + 1e-7 * np.abs(flat[big])
    # magnitude: exactly lr for every clear gradient (independent of the BF16 gradient error)
    assert np.all(np.abs(np.abs(step[big]) - lr) <= tol)
    # direction: against the reference gradient's sign (BF16 may flip a rare near-cancelling one)
    agree = np.abs(step[big] + lr * np.sign(g[big])) <= tol
    say(f"adam: sign agreement {agree.mean():.5f}")
    assert agree.mean() >= 0.999
    md.close()


def _near_tie(d):
    d = np.sort(d)[:17]
    gaps = np.diff(d) / np.maximum(d[1:], 1e-12)
    return np.any(gaps < 1e-4)


def _check_knn_rows(F, col, box):
    """Features 0-65 of every row without a near-tie against the oracle's kNN rows."""
    n = col.total_cells()
    x = col.flat("pos").reshape(-1, 3)
    sig = col.flat("sigma").reshape(-1)
    spec = np.concatenate([np.full(sp["ncells"], s) for s, sp in enumerate(col.species)])
    L = np.array(box.length[:3])
    ref = M.knn_features(x, sig, spec, L)
    xw = np.mod(x, L)
    checked = 0
    for i in range(n):
        d = xw - xw[i]
        d -= L * np.round(d / L)
        r = np.sqrt((d * d).sum(1))
        r[i] = np.inf
        if _near_tie(r):
            continue
        checked += 1
        want = bf16(ref[i, :66])
        tol = 2.0 ** -7 * np.abs(want) + 1e-5
        if not np.all(np.abs(F[i, :66] - want) <= tol):
            bad = np.nonzero(np.abs(F[i, :66] - want) > tol)[0]
            raise AssertionError(f"row {i}: features {bad[:8]} gpu {F[i][bad[:8]]} oracle {want[bad[:8]]}")
    say(f"knn: {checked}/{n} rows compared")
    assert checked >= 0.95 * n


def test_knn_features_match_oracle(gpu):
    p, box, col = configs.relaxed_bulk_3d(3000, seed=7, precision=abi.PREC_FP32)
    md = HipMD(p, box, [sp["capacity"] for sp in col.species])
    md.upload(col)
    n = col.total_cells()
    F = md.knn_features(n).astype(np.float64)  # at step 0, forces computed by the call
    _check_knn_rows(F, col, box)
    # features 66-71 of every row: one-pass drift and noise sum of the macro-step from step 0
    o = Oracle(p, box)
    o.forces(col)
    o.close()
    temp = np.concatenate([p.dt / (p.species[s].zeta[0] * np.pi * (sp["sigma"][:sp["ncells"], 0] / 2) ** 2)
                           for s, sp in enumerate(col.species)])
    kT = np.concatenate([np.full(sp["ncells"], p.species[s].kT) for s, sp in enumerate(col.species)])
    keys = [((s << 28) | c) << 1 for s, sp in enumerate(col.species) for c in range(sp["ncells"])]
    z = M.fdt_noise_sum(p.seed, keys, np.sqrt(2 * kT * temp), 0)
    want = M.macro_inputs(col.flat("frc").reshape(-1, 3), temp, z)
    tol = 2.0 ** -7 * np.abs(want) + 1e-6
    tol[:, :3] += 1e-3 * np.sqrt(np.mean(want[:, :3] ** 2))
    err = np.abs(F[:, 66:] - want)
    say(f"drift and noise features: worst error / bar {np.max(err / tol):.3f}; rms drift {np.sqrt(np.mean(want[:, :3] ** 2)):.3e}, "
        f"rms noise {np.sqrt(np.mean(want[:, 3:] ** 2)):.3e} (x {M.BASE_SCALE})")
    assert np.all(err <= tol)
    md.close()


@pytest.mark.parametrize("skin", [0.3, 1.2])
def test_knn_features_list_mode_after_moves(gpu, skin):
    """Neighbour-list mode: features at the list's build, then after 3 macro-steps and 5 plain
    steps (positions off the build by up to skin / 2, candidate lists reused).  Skin 1.2 makes
    most candidate lists overfull (about 150 particles within rb = 3.2 sigma), so most rows come
    from the fallback bin walk with its staleness margin."""
    p, box, col = configs.relaxed_bulk_3d(3000, seed=8, precision=abi.PREC_FP32)
    md = HipMD(p, box, [sp["capacity"] for sp in col.species])
    md.upload(col)
    md.forces(want=False)
    md.set_neighbour_list(True, skin, 512)
    n = col.total_cells()
    _check_knn_rows(md.knn_features(n).astype(np.float64), col, box)
    md.mlp_set_params(np.zeros(abi.MLP_NPARAM, dtype=np.float32))
    ts, st = md.surrogate_step(3, 0, judge_every=0, tol=1.0)
    say(f"list mode skin {skin}: {st['knn_list_builds']} candidate-list builds, {st['knn_fallbacks']} fallback rows")
    if skin > 1.0:
        assert st["knn_fallbacks"] >= 2 * n  # most rows of the calls so far
    else:
        assert st["knn_list_builds"] >= 1
    ts = md.step(5, ts)
    F = md.knn_features(n).astype(np.float64)
    c = col.copy()
    md.download(c)
    L = np.array(box.length[:3])
    d = c.flat("pos").reshape(-1, 3) - col.flat("pos").reshape(-1, 3)
    moved = np.abs(d - L * np.round(d / L)).max()
    say(f"list mode skin {skin}: largest move {moved:.3e} sigma after {ts} steps")
    assert moved > 1e-2
    _check_knn_rows(F, c, box)
    md.close()


@pytest.mark.parametrize("skin", [0.3, 0.1])
def test_list_mode_macro_steps_match_stencil_mode(gpu, skin):
    """The same sequence of macro-steps, judged macro-steps (plain result kept) and plain steps
    with a zero network, in list mode (macro-steps keep the list's bookkeeping, the list is
    rebuilt by its own skin test) and in stencil mode (re-sorted every macro-step): the two
    differ only by the FP32 force summation order.  With skin 0.1 (rebuild past 0.05 sigma) the
    skin test of the macro-steps themselves triggers most rebuilds."""
    p, box, col = configs.relaxed_bulk_3d(4096, seed=6, precision=abi.PREC_FP32)
    ends = []
    for list_on in (False, True):
        md = HipMD(p, box, [sp["capacity"] for sp in col.species])
        md.upload(col)
        md.forces(want=False)
        if list_on:
            md.set_neighbour_list(True, skin, 256)
        md.mlp_set_params(np.zeros(abi.MLP_NPARAM, dtype=np.float32))
        ts, st = md.surrogate_step(12, 0, judge_every=3, tol=1e9)
        assert st["macro_steps"] == 8 and st["judged"] == 4
        ts = md.step(9, ts)
        ts, st = md.surrogate_step(5, ts, judge_every=0, tol=1.0)
        assert ts == 8 * 17 + 9
        if list_on:
            say(f"list mode skin {skin}: {st['knn_list_builds']} candidate-list builds in 17 macro-steps")
            if skin == 0.1:
                assert st["knn_list_builds"] >= 5
        c = col.copy()
        md.download(c)
        md.close()
        ends.append(c.flat("pos").reshape(-1, 3))
    L = np.array(box.length[:3])
    d = ends[1] - ends[0]
    d -= L * np.round(d / L)
    moved = ends[0] - col.flat("pos").reshape(-1, 3)
    moved -= L * np.round(moved / L)
    say(f"list (skin {skin}) vs stencil mode after 145 steps: max |dx| {np.abs(d).max():.3e}, "
        f"rms move {np.sqrt((moved ** 2).mean()):.3e}")
    assert np.abs(d).max() < 1e-5


def test_long_list_mode_surrogate_run(gpu):
    """100 macro-steps of a 16,384-particle bulk at kT 0.05 with a zero network (every
    macro-step an Euler step of 8 dt; judged every 4th), in list mode and in stencil mode: the
    list's own skin test drives its rebuilds, the candidate lists follow them, and the two runs
    end within 1e-4 sigma of each other (FP32 summation order, amplified over 800 steps)."""
    p, box, col = configs.relaxed_bulk_3d(16384, seed=11, precision=abi.PREC_FP32)
    ends, stats = [], []
    for list_on in (False, True):
        md = HipMD(p, box, [sp["capacity"] for sp in col.species])
        md.upload(col)
        md.forces(want=False)
        if list_on:
            md.set_neighbour_list(True, 0.3, 256)
        md.mlp_set_params(np.zeros(abi.MLP_NPARAM, dtype=np.float32))
        ts, st = md.surrogate_step(100, 0, judge_every=4, tol=1e9)
        assert ts == 800
        c = col.copy()
        md.download(c)
        md.close()
        ends.append(c.flat("pos").reshape(-1, 3))
        stats.append(st)
    L = np.array(box.length[:3])
    d = ends[1] - ends[0]
    d -= L * np.round(d / L)
    say(f"long run: list-mode stats {stats[1]}; list vs stencil max |dx| {np.abs(d).max():.3e}")
    assert np.all(np.isfinite(ends[1]))
    assert stats[1]["macro_steps"] == 75 and stats[1]["judged"] == 25
    assert stats[1]["knn_list_builds"] >= 3
    assert np.abs(d).max() < 1e-4


def test_surrogate_requires_supported_setup(gpu):
    p, box, col = configs.plain_2d_sc(n_cells=64, L=16.0, seed=2)
    md = HipMD(p, box, [sp["capacity"] for sp in col.species])
    md.upload(col)
    with pytest.raises(abi.CilError, match="surrogate"):
        md.knn_features(col.total_cells())
    md.close()


@pytest.fixture(scope="module")
def trained():
    """A 4096-particle relaxed bulk at kT = 0.005 (where the noise floor of start-state
    predictors, ~3e-4 sigma, is well below the drift), 24 plain macro-steps of training data,
    4000 Adam steps."""
    p, box, col = configs.relaxed_bulk_3d(4096, seed=3, precision=abi.PREC_FP32, kT=0.005)
    md = HipMD(p, box, [sp["capacity"] for sp in col.species])
    md.upload(col)
    md.forces(want=False)
    ts, nsamp = md.surrogate_collect(24, 1, 1 << 20, 0)
    assert nsamp == 24 * 4096 and ts == 24 * 8
    md.mlp_init(7)
    l1 = md.surrogate_train(3000, batch=4096, lr=1e-3, seed=1)
    l2 = md.surrogate_train(1000, batch=4096, lr=2e-4, seed=2)
    losses = np.concatenate([l1, l2])
    say(f"training loss: first {losses[0]:.4e}, last {losses[-1]:.4e}")
    yield md, ts, losses
    md.close()


def test_training_reduces_loss(gpu, trained):
    md, ts, losses = trained
    assert np.all(np.isfinite(losses))
    assert losses[-20:].mean() < 0.2 * losses[:5].mean()


@pytest.mark.parametrize("list_on", [False, True])
def test_zero_correction_is_one_euler_step(gpu, list_on):
    """With a network that outputs 0 the macro-step is one Euler step of 8 dt (the one-pass drift
    is kept in FP32: differences at FP32 rounding of the drift)."""
    p, box, col = configs.relaxed_bulk_3d(2048, seed=5, precision=abi.PREC_FP32, kT=0.005)
    md = HipMD(p, box, [sp["capacity"] for sp in col.species])
    md.upload(col)
    md.forces(want=False)
    if list_on:
        md.set_neighbour_list(True, 0.3, 256)
    md.mlp_set_params(np.zeros(abi.MLP_NPARAM, dtype=np.float32))
    ts, rms, mx = md.surrogate_eval(0)
    say(f"zero correction: rms network {rms[0]:.6e}, one Euler step {rms[2]:.6e}, zero drift {rms[1]:.6e}")
    assert ts == 8
    assert abs(rms[0] - rms[2]) <= 1e-5 * rms[2]
    assert rms[2] < rms[1]
    md.close()


def test_trained_network_beats_one_euler_step(gpu, trained):
    md, ts, _ = trained
    ts, rms, mx = md.surrogate_eval(ts)
    say(f"8-step displacement error vs plain (rms / max, sigma): network {rms[0]:.3e} / {mx[0]:.3e}, "
        f"zero drift {rms[1]:.3e} / {mx[1]:.3e}, one Euler step of 8 dt {rms[2]:.3e} / {mx[2]:.3e}")
    assert rms[0] < rms[2]
    assert rms[0] < 0.9 * rms[1]


def test_noise_floor_of_start_state_predictors(gpu):
    """The spread of the 8-step deterministic displacement over noise seeds bounds every
    start-state predictor from below; it must be below the displacement itself and above 0."""
    from cil_cuda_amd import surrogate

    p, box, col = configs.relaxed_bulk_3d(4096, seed=3, precision=abi.PREC_FP32)
    floor, drift = surrogate.noise_floor(p, box, col)
    say(f"noise floor {floor:.3e} sigma, deterministic displacement rms {drift:.3e} sigma")
    assert 0.0 < floor < drift


def test_noise_order_floor_bounds_the_network(gpu, trained):
    """The drift's dependence on the order of the 8 increments, which the macro-step's inputs do
    not carry, is positive and below the trained network's error on the same kind of state."""
    from cil_cuda_amd import surrogate

    md, ts, _ = trained
    p, box, col = configs.relaxed_bulk_3d(4096, seed=3, precision=abi.PREC_FP32, kT=0.005)
    floor = surrogate.noise_order_floor(p, box, col)
    ts, rms, mx = md.surrogate_eval(ts)
    say(f"noise-order floor {floor:.3e} sigma; trained network {rms[0]:.3e} sigma")
    assert 0.0 < floor < rms[0]


def test_network_free_predictors(gpu):
    """Of the predictors with the exact noise sum, a second force pass at x0 + (3.5/8)(b + z)
    beats one Euler step of 8 dt, which beats zero drift, and none goes below the noise-order
    floor."""
    from cil_cuda_amd import surrogate

    p, box, col = configs.relaxed_bulk_3d(4096, seed=3, precision=abi.PREC_FP32, kT=0.05)
    e = surrogate.base_predictor_errors(p, box, col)
    floor = surrogate.noise_order_floor(p, box, col)
    say(f"network-free predictors (kT 0.05): {e}; noise-order floor {floor:.3e}")
    assert floor < e["midpoint"] < e["euler_8dt"] < e["zero_drift"]


def test_surrogate_step_judge_and_fallback(gpu, trained):
    md, ts, _ = trained
    md.surrogate_stats_reset()
    ts1, st = md.surrogate_step(8, ts, judge_every=4, tol=1e9)
    say(f"tolerant judge: {st}")
    assert ts1 == ts + 64
    assert st["judged"] == 2 and st["failed"] == 0 and st["macro_steps"] == 6 and st["plain_steps"] == 16
    md.surrogate_stats_reset()
    ts2, st = md.surrogate_step(6, ts1, judge_every=2, tol=0.0)
    say(f"strict judge: {st}")
    # m0 surrogate, m1 judged+failed, m2 m3 fallback, m4 judged+failed, m5 fallback
    assert ts2 == ts1 + 48
    assert st["judged"] == 2 and st["failed"] == 2 and st["macro_steps"] == 1
    assert st["plain_steps"] == 8 * 5
    # state stays usable by the plain integrator afterwards
    ts3 = md.step(3, ts2)
    assert ts3 == ts2 + 3


def test_surrogate_step_deterministic(gpu, trained):
    md, ts, _ = trained
    params = md.mlp_get_params()
    p, box, col = configs.relaxed_bulk_3d(2048, seed=9, precision=abi.PREC_FP32)
    outs = []
    for _ in range(2):
        m2 = HipMD(p, box, [sp["capacity"] for sp in col.species])
        m2.upload(col)
        m2.forces(want=False)
        m2.mlp_set_params(params)
        m2.surrogate_step(5, 0, judge_every=0, tol=1.0)
        c = col.copy()
        m2.download(c)
        outs.append(c.flat("pos").copy())
        m2.close()
    assert np.array_equal(outs[0], outs[1])
    assert np.all(np.isfinite(outs[0]))
