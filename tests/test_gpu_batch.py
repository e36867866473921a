"""GPU tests of colony batches: every colony of a batch must evolve bit-identically to the
same colony stepped alone (same seed, parameters and tuning), for heterogeneous sizes, boxes,
seeds, noise temperatures, motilities and growth rates, through both the graph-captured and
the eager step paths, with and without division."""
import sys

import numpy as np
import pytest

from cil_cuda_amd import abi, configs
from cil_cuda_amd.colony import HipBatch, HipMD

pytestmark = pytest.mark.gpu

FIELDS = ("pos", "vel", "sigma", "rest_len", "state", "t0", "tswitch", "sigma_t", "pressure")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _make(p, box, col):
    md = HipMD(p, box, [sp["capacity"] for sp in col.species])
    md.upload(col)
    md.forces(want=False)
    return md


def _snapshot(md, col):
    c = col.copy()
    md.download(c)
    return c


def _assert_same(a, b, tag):
    assert a.ncells() == b.ncells(), f"{tag}: cell counts {a.ncells()} vs {b.ncells()}"
    for f in FIELDS:
        x, y = a.flat(f), b.flat(f)
        if not np.array_equal(x, y):
            d = np.max(np.abs(np.asarray(x, float) - np.asarray(y, float)))
            raise AssertionError(f"{tag}: field {f} differs (max {d:.3e})")


def _run_pair(builders, nsteps, tune=None):
    """Step each colony alone and all of them as one batch; return both end states."""
    alone, batched = [], []
    for p, box, col in builders:
        md = _make(p, box, col)
        if tune:
            md.set_tuning(*tune[:2])
            md.set_kernel(tune[2])
        ts = md.step(nsteps, 0)
        assert ts == nsteps
        alone.append(_snapshot(md, col))
        md.close()
    members = [_make(p, box, col) for p, box, col in builders]
    if tune:
        for m in members:
            m.set_tuning(*tune[:2])
            m.set_kernel(tune[2])
    batch = HipBatch(members)
    ts = batch.step(nsteps, [0] * len(members))
    assert ts == [nsteps] * len(members)
    for m, (p, box, col) in zip(members, builders):
        batched.append(_snapshot(m, col))
    batch.close()
    for m in members:
        m.close()
    return alone, batched


def test_batch_plain_2d_fp64_matches_single_colonies(gpu):
    builders = []
    for k, (n, L, kT, mot) in enumerate(((60, 16.0, 0.01, 0.1), (150, 24.0, 0.002, 0.4),
                                         (300, 32.0, 0.05, 0.0), (500, 40.0, 0.02, 0.25))):
        builders.append(configs.plain_2d_sc(n_cells=n, L=L, seed=10 + k, kT=kT, motility=mot))
    alone, batched = _run_pair(builders, 58)  # 50-step graph + 8 eager steps
    for k, (a, b) in enumerate(zip(alone, batched)):
        _assert_same(a, b, f"colony {k}")
    say("plain 2D FP64: 4 colonies bit-identical after 58 steps")


def test_batch_plain_2d_fp32_lean_kernel(gpu):
    builders = [configs.plain_2d_sc(n_cells=n, L=L, seed=20 + k, precision=abi.PREC_FP32)
                for k, (n, L) in enumerate(((100, 20.0), (700, 48.0), (40, 12.0)))]
    alone, batched = _run_pair(builders, 23, tune=(2, 128, 0))
    for k, (a, b) in enumerate(zip(alone, batched)):
        _assert_same(a, b, f"colony {k}")


def test_batch_growing_channels_with_division(gpu):
    builders = []
    for k, (tau, kT, length) in enumerate(((0.0015, 0.001, 64.0), (0.0025, 0.01, 96.0), (0.002, 0.005, 128.0))):
        builders.append(configs.channel_2d_hz(length=length, seed=30 + k, kT=kT, tau=tau, capacity=512))
    alone, batched = _run_pair(builders, 240)
    for k, (a, b) in enumerate(zip(alone, batched)):
        _assert_same(a, b, f"colony {k}")
        say(f"channel colony {k}: {a.total_cells()} cells after 240 steps")
    assert max(a.total_cells() for a in alone) > 1  # division happened


def test_batch_theta_cycle_division(gpu):
    builders = [configs.theta_2d_hz(n_cells=n, seed=40 + n) for n in (8, 16, 24)]
    alone, batched = _run_pair(builders, 64)
    for k, (a, b) in enumerate(zip(alone, batched)):
        _assert_same(a, b, f"colony {k}")


def test_batch_autotune_keeps_colonies_consistent(gpu):
    builders = [configs.plain_2d_sc(n_cells=n, L=L, seed=50 + k, precision=abi.PREC_FP32)
                for k, (n, L) in enumerate(((200, 24.0), (400, 32.0)))]
    members = [_make(p, box, col) for p, box, col in builders]
    batch = HipBatch(members)
    best = batch.autotune(reps=2)
    say(f"batch autotune: {best}")
    assert best["bins_per_cutoff"] in (1, 2, 3) and best["block"] in (64, 128, 256)
    ts = batch.step(12, [0, 0])
    got = [_snapshot(m, col) for m, (p, box, col) in zip(members, builders)]
    batch.close()
    for m in members:
        m.close()
    # the same colonies alone with the chosen tuning
    for k, (p, box, col) in enumerate(builders):
        md = _make(p, box, col)
        md.set_tuning(best["bins_per_cutoff"], best["block"])
        md.set_kernel(best["variant"])
        md.forces(want=False)
        md.step(12, 0)
        _assert_same(_snapshot(md, col), got[k], f"colony {k}")
        md.close()
    assert ts == [12, 12]


@pytest.mark.parametrize("model", ["growth", "theta"])
def test_batched_divisions_match_the_oracle(gpu, model):
    """Colonies with division stepped in one batch against the oracle stepped alone: the same
    division events (cell counts and birth steps) and positions within 1e-9 sigma (FP64)."""
    from oracle import oracle as O

    if model == "growth":
        builders = [configs.channel_2d_hz(length=length, seed=60 + k, kT=kT, tau=tau, capacity=512)
                    for k, (tau, kT, length) in enumerate(((0.0015, 0.001, 64.0), (0.002, 0.005, 96.0)))]
        nsteps = 240
    else:
        builders = [configs.theta_2d_hz(n_cells=n, seed=70 + n) for n in (12, 20)]
        nsteps = 64
    members = [_make(p, box, col) for p, box, col in builders]
    batch = HipBatch(members)
    assert batch.step(nsteps, [0] * len(members)) == [nsteps] * len(members)
    got = [_snapshot(m, col) for m, (p, box, col) in zip(members, builders)]
    batch.close()
    for m in members:
        m.close()
    sys.path.insert(0, __file__.rsplit("/", 1)[0])
    from test_gpu_parity import _max_disp_err

    for k, (p, box, col) in enumerate(builders):
        ref = col.copy()
        orc = O.Oracle(p, box, traversal=1)
        orc.forces(ref, 0)
        orc.step(ref, nsteps, 0)
        orc.close()
        a, b = got[k], ref
        assert a.ncells() == b.ncells(), (k, a.ncells(), b.ncells())
        n = a.ncells()[0]
        assert np.array_equal(a.species[0]["t0"][:n], b.species[0]["t0"][:n])
        err = _max_disp_err(b, a, box)
        say(f"{model} colony {k}: {n} cells ({col.ncells()[0]} at the start), {err:.2e} sigma from the oracle")
        assert err <= 1e-9
    assert any(g.ncells()[0] > c.ncells()[0] for g, (_, _, c) in zip(got, builders))  # divisions happened


def test_batch_rejects_mixed_launch_shapes(gpu):
    a = configs.plain_2d_sc(n_cells=40, L=12.0, seed=1)
    b = configs.plain_2d_sc(n_cells=40, L=12.0, seed=2, precision=abi.PREC_FP32)
    ma, mb = _make(*a), _make(*b)
    with pytest.raises(abi.CilError):
        HipBatch([ma, mb])
    ma.close()
    mb.close()
