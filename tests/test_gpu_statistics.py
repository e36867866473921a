"""Long-run statistics of the seeded run's configuration against the reference's own runs.

References:
  - the reference's shipped run of this configuration (output/z0.001t3.16_v2.h5, seed 1, one
    cell at t = 0, snapshots every 0.1 time units to t = 10; fixture
    tests/golden/z0.001t3.16_v2_frames.npz), analysed here with the reference tool's
    definitions (tests/colony_stats.py);
  - the reference tool's own output for two further runs (quantities.npz and
    quantities_v3.2.npz, two builds; fixture tests/golden/quantities_ref.npz for t <= 10).  Their
    first divisions differ from the shipped run's, so they are other realisations (another seed or
    set-up); they are printed beside the comparison, not asserted.

Here the configuration (read by the library's configuration reader) is run on the GPU from one
cell, five colonies with seeds 1-5 in one batch, with counter-based draws, to t = 10.  Asserted
against the shipped run, with bars from the spread of the five seeds:
  - cell counts at t = 2, 4, 6, 8, 10 and the growth rate over the crowded phase t = 7-10;
  - the cell-area distribution at t = 8 and 10 (L1 distance of the density histograms);
  - the mean cell-cycle time of the divisions between t = 1 and 10.
"""
import os
import sys
import time

import numpy as np
import pytest

from cil_cuda_amd import abi, config
from cil_cuda_amd.colony import HipBatch, HipMD

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
from colony_stats import AREA_EDGES, area_hist, division_times, load_frames  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))
SEEDED = os.path.join(HERE, "golden", "configs", "z0.001t3.16_v2.RunParams.json")
SHIPPED = os.path.join(HERE, "golden", "z0.001t3.16_v2_frames.npz")
QUANT = os.path.join(HERE, "golden", "quantities_ref.npz")
SEEDS = (1, 2, 3, 4, 5)
T_END = 10.0

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def run_colonies():
    members, cols = [], []
    for seed in SEEDS:
        p, box, run = config.read_config(SEEDED)
        p.seed = seed
        col = config.initial_colony(SEEDED, p, [4096])
        p.rng = abi.RNG_COUNTER
        md = HipMD(p, box, [4096])
        md.upload(col)
        md.forces(want=False)
        members.append(md)
        cols.append(col)
    dt, gts, tau = p.dt, run.gts, p.species[0].division_tau.mu
    nframes = int(round(T_END / (gts * dt)))
    batch = HipBatch(members)
    frames = [[] for _ in SEEDS]
    ts = [0] * len(SEEDS)

    def snap():
        for c, md in enumerate(members):
            s = md.download(cols[c].copy()).species[0]
            n = s["ncells"]
            frames[c].append((s["theta"][:n].copy(), s["rho"][:n].copy(), s["t0"][:n].copy(), s["sigma"][:n, 0].copy()))

    snap()
    t0 = time.perf_counter()
    for k in range(1, nframes + 1):
        ts = batch.step(gts, ts)
        snap()
        if k % 10 == 0:
            say(f"statistics: t = {k * gts * dt:.1f}: cells {[len(f[-1][0]) for f in frames]} "
                f"({time.perf_counter() - t0:.1f} s)")
    batch.close()
    for md in members:
        md.close()
    return frames, dt, gts, tau


@pytest.fixture(scope="module")
def runs():
    return run_colonies()


def _band(vals, ref):
    """ref inside [min, max] of vals widened by their spread."""
    lo, hi = float(np.min(vals)), float(np.max(vals))
    w = hi - lo
    return lo - w <= ref <= hi + w


def test_cell_number_growth(runs):
    frames, dt, gts, tau = runs
    ship = load_frames(SHIPPED)
    q = np.load(QUANT, allow_pickle=False)
    counts = np.array([[len(f[0]) for f in fr] for fr in frames])
    per = gts * dt
    ok = True
    for t in (2, 4, 6, 8, 10):
        i = int(round(t / per))
        ref = len(ship[i][0])
        ok = ok and _band(counts[:, i], ref)
        say(f"statistics: t = {t}: cells {counts[:, i].tolist()}; shipped run {ref}; other reference runs "
            f"{q['cellNumber'][:, i].tolist()}")
    i7, i10 = int(round(7 / per)), int(round(10 / per))
    mine = np.log(counts[:, i10] / counts[:, i7]) / 3.0
    ref_rate = float(np.log(len(ship[i10][0]) / len(ship[i7][0])) / 3.0)
    other = np.log(q["cellNumber"][:, i10] / q["cellNumber"][:, i7]) / 3.0
    say(f"statistics: growth rate t = 7..10: {np.round(mine, 4).tolist()}; shipped run {ref_rate:.4f}; "
        f"other reference runs {np.round(other, 4).tolist()}")
    assert ok and _band(mine, ref_rate)


def test_cell_area_distribution(runs):
    frames, dt, gts, tau = runs
    ship = load_frames(SHIPPED)
    q = np.load(QUANT, allow_pickle=False)
    w = np.diff(AREA_EDGES)
    centres = (AREA_EDGES[1:] + AREA_EDGES[:-1]) / 2
    for t in (8, 10):  # earlier, most cells are larger than the bins' 15.5
        i = int(round(t / (gts * dt)))
        mine = np.array([area_hist(fr[i][3]) for fr in frames])
        ref = area_hist(ship[i][3])
        l1 = lambda a, b: float(np.sum(np.abs(a - b) * w))  # noqa: E731  (0 equal, 2 disjoint)
        d_seeds = [l1(m, ref) for m in mine]
        d_pairs = [l1(a, b) for k, a in enumerate(mine) for b in mine[k + 1:]]
        other = np.nan_to_num(q["cellAreaDistributions"][:, i])  # an empty histogram is NaN there
        say(f"statistics: area histogram t = {t}: L1 to the shipped run per seed {np.round(d_seeds, 3).tolist()}, "
            f"between seeds max {max(d_pairs):.3f}; other reference runs to the shipped run "
            f"{[round(l1(o, ref), 3) for o in other]}; mean area {np.round([np.sum(m * centres * w) for m in mine], 3).tolist()} "
            f"vs {float(np.sum(ref * centres * w)):.3f}")
        assert np.median(d_seeds) <= max(d_pairs) + 0.05


def test_division_time_distribution(runs):
    frames, dt, gts, tau = runs
    ship = load_frames(SHIPPED)
    q = np.load(QUANT, allow_pickle=False)
    mine = [division_times(fr, dt, gts, tau)[0] for fr in frames]
    ref = division_times(ship, dt, gts, tau)[0]
    means = np.array([m.mean() for m in mine])
    say(f"statistics: mean cell-cycle time (divisions t = 1..10): per seed {np.round(means, 4).tolist()} over "
        f"{[m.size for m in mine]} divisions; shipped run {ref.mean():.4f} over {ref.size}; other reference runs "
        f"{[round(float(q[f'divisionTimes_{k}'].mean()), 4) for k in range(2)]}")
    assert _band(means, float(ref.mean()))
