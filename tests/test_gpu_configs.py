"""The five workloads at their stated sizes on the GPU.

  C1  2D periodic soft-core colony, 2000 two-disc cells, L = 64, FDT noise, 10^4 FP64 steps:
      within 1e-9 sigma of the oracle (its race-free threaded traversal).
  C2  2D channel 40 wide, one two-disc cell grown to 65,536 cells (Hertz, walls, drain, linear
      growth 1/tau to axis 2 sigma, +-5 % diameter jitter, dt 5e-5, kT 0.001): no capacity
      error, every drain removal a cell outside the box, and a bit-identical replay of the
      first 10^5 steps; in FP64 the first 255 divisions equal the oracle's.
  C3  3D bulk of 1,048,576 discs (mixed LJ 2.5 sigma / soft core, phi 0.6): FP32 neighbour-list
      forces of the relaxed state against the FP64 oracle for 16,384 sampled particles, FP32 bar.
  C4  256 colonies of log-uniform 1k-64k two-disc cells (seed 1000 + c, motility U[0, 0.5],
      kT U[0.001, 0.05], per-colony growth rate) in one batch of 1000-step graph launches:
      sampled colonies bit-equal to the same colonies stepped alone; batch throughput printed.
  C5  the surrogate step on the 1M bulk at kT 0.05 with the judge at every macro-step: failed
      judgements fall back to the plain integrator and the judge keeps the plain result, so the
      run must equal the plain run bit for bit.
"""
import math
import os
import sys
import time

import numpy as np
import pytest

from cil_cuda_amd import abi, configs
from cil_cuda_amd.colony import Colony, HipBatch, HipMD
from oracle import oracle as O

sys.path.insert(0, __file__.rsplit("/", 1)[0])
from test_gpu_parity import _max_disp_err  # noqa: E402

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _threads():
    return max(1, min(16, os.cpu_count() or 1))


def test_c1_full_size_matches_oracle(gpu):
    p, box, col = configs.plain_2d_sc(n_cells=2000, L=64.0, seed=1)
    nsteps = 10000
    md = HipMD(p, box, [2000])
    md.upload(col)
    md.forces(want=False)
    t0 = time.perf_counter()
    assert md.step(nsteps, 0) == nsteps
    got = col.copy()
    md.download(got)
    md.close()
    say(f"C1: GPU {nsteps} steps in {time.perf_counter() - t0:.2f} s")
    ref = col.copy()
    orc = O.Oracle(p, box, traversal=0, threads=_threads())
    orc.forces(ref, 0)
    t0 = time.perf_counter()
    for k in range(10):
        orc.step(ref, nsteps // 10, k * (nsteps // 10))
        say(f"C1: oracle {(k + 1) * nsteps // 10} steps ({time.perf_counter() - t0:.1f} s)")
    orc.close()
    err = _max_disp_err(ref, got, box)
    say(f"C1: max displacement difference after {nsteps} steps: {err:.3e}")
    assert err <= 1e-9


C2_CAP = 131072  # divisions come in near-synchronous generations: room for the one that crosses 65,536


def _c2(length=8192.0, precision=abi.PREC_FP32, capacity=C2_CAP):
    return configs.channel_2d_hz(width=40.0, length=length, seed=1, kT=0.001, dt=5e-5, tau=1.0, capacity=capacity,
                                 precision=precision)


def test_c2_channel_grows_to_65536_cells(gpu):
    """The channel colony grown to 65,536 cells.  Wall containment is not asserted: a FlatWall
    force is capped at its value at 0.82 sigma (Wall.cpp:3), 27.3 epsilon / sigma^2, while the
    frictional back-pressure at the middle of a colony that expands at both ends grows with the
    square of its length (ζ = 1 here), so discs are pressed through the walls once the colony is
    a few tens of sigma long and the drain removes them.  Asserted: the drain removes only cells
    with a disc outside the box, and leaves none behind beyond the reference's swap-skip quirk."""
    p, box, col = _c2()
    md = HipMD(p, box, [C2_CAP])
    md.upload(col)
    md.forces(want=False)
    ts, removed_total, n = 0, 0, 1
    t0 = time.perf_counter()
    while True:
        chunk = 20000 if n < 16384 else 2000  # stop within a few hundred cells of 65,536
        ts = md.step(chunk, ts)
        c = col.copy()
        md.download(c)
        n = c.ncells()[0]
        pos = c.species[0]["pos"][:n]
        x, y = pos[..., 0], pos[..., 1]
        out_x = int(np.sum(np.any((x < 0.0) | (x >= 40.0), axis=1)))
        outside = int(np.sum(np.any((x < 0.0) | (x >= 40.0) | (y < 0.0) | (y >= box.length[1]), axis=1)))
        removed = md.check_positions()
        # drain removes cells outside the box; the reference's swap-skip quirk can leave one for
        # the next check, never remove an inside cell
        assert removed <= outside
        c2 = col.copy()
        md.download(c2)
        n = c2.ncells()[0]
        pos = c2.species[0]["pos"][:n]
        left = int(np.sum(np.any((pos[..., 0] < 0.0) | (pos[..., 0] >= 40.0) | (pos[..., 1] < 0.0)
                                 | (pos[..., 1] >= box.length[1]), axis=1)))
        # every removed cell was outside, and every outside cell is removed or left by the quirk
        assert left == outside - removed and (outside == 0 or removed >= 1), (ts, outside, removed, left)
        removed_total += removed
        if ts % 100000 == 0 or n >= 65536 or (n >= 16384 and ts % 10000 == 0):
            say(f"C2: step {ts}: {n} cells, removed {removed_total} ({out_x} through the walls this check), "
                f"{time.perf_counter() - t0:.1f} s")
        if n >= 65536:
            break
        assert ts < 1_200_000, f"colony stopped growing at {n} cells"
    md.close()
    # bit-identical replay of the first 10^5 steps
    outs = []
    for _ in range(2):
        md = HipMD(p, box, [C2_CAP])
        md.upload(col)
        md.forces(want=False)
        md.step(100000, 0)
        c = col.copy()
        md.download(c)
        md.close()
        outs.append(c)
    say(f"C2: replay at step 100000: {outs[0].ncells()} cells")
    assert outs[0].ncells() == outs[1].ncells()
    for f in ("pos", "sigma", "rest_len", "t0"):
        assert np.array_equal(outs[0].flat(f), outs[1].flat(f)), f


def test_c2_first_255_divisions_match_oracle(gpu):
    """The channel from its one cell to 256 cells (255 divisions, 240,000 steps) in FP64 on the GPU
    and in the oracle: the same division events (cell count, birth steps and rest lengths bit for
    bit, jittered daughter diameters to 1e-14) and positions that agree to within 1e-3 sigma
    (the two sum pair forces in different orders)."""
    p, box, col = _c2(precision=abi.PREC_FP64, capacity=4096)
    md = HipMD(p, box, [4096])
    md.upload(col)
    md.forces(want=False)
    assert md.step(240000, 0) == 240000
    got = md.download(col.copy())
    md.close()
    say(f"C2 FP64: GPU at step 240000: {got.ncells()} cells")
    ref = col.copy()
    orc = O.Oracle(p, box, traversal=0, threads=max(2, _threads()))
    orc.forces(ref, 0)
    t0, ts = time.perf_counter(), 0
    while ts < 240000:
        ts = orc.step(ref, 20000, ts)
        say(f"C2 FP64: oracle step {ts}: {ref.ncells()} cells ({time.perf_counter() - t0:.1f} s)")
    orc.close()
    assert got.ncells() == ref.ncells() == [256]
    a, b = got.species[0], ref.species[0]
    for f in ("t0", "rest_len", "state"):
        assert np.array_equal(a[f][:256], b[f][:256]), f
    # the jittered diameters come from the same Philox draws; the Gaussian transform's log / cos
    # may differ in the last bit between the device and the host library
    dsig = float(np.max(np.abs(a["sigma"][:256] / b["sigma"][:256] - 1.0)))
    say(f"C2 FP64: daughter diameters: max relative difference {dsig:.2e}")
    assert dsig <= 1e-14
    err = _max_disp_err(ref, got, box)
    say(f"C2 FP64: max position difference after 255 divisions: {err:.3e}")
    assert err <= 1e-3


def test_c3_1m_list_forces_match_fp64(gpu):
    """FP32 neighbour-list forces of the relaxed 1M state against the FP64 oracle on a sample of
    16,384 particles (every 64th by UID; the oracle's threaded gather evaluates only those)."""
    t0 = time.perf_counter()
    p, box, col = configs.relaxed_bulk_3d(1 << 20, seed=1, precision=abi.PREC_FP32)
    say(f"C3: relaxed 1M particles ({time.perf_counter() - t0:.1f} s)")
    md = HipMD(p, box, [sp["capacity"] for sp in col.species])
    md.set_tuning(2, 64)
    md.upload(col)
    md.set_neighbour_list(True, 0.3, 256)
    frc, _ = md.forces()
    md.close()
    say("C3: GPU list forces done")
    q = abi.Params.from_buffer_copy(p)
    q.precision = abi.PREC_FP64
    every = 64
    t0 = time.perf_counter()
    orc = O.Oracle(q, box, traversal=0, threads=max(2, _threads()))
    orc.set_sample(every)
    ref = col.copy()
    orc.forces(ref, 0)
    say(f"C3: oracle forces of the sample in {time.perf_counter() - t0:.1f} s")
    scale = orc.force_scale_gather(ref)
    orc.close()
    sel = np.arange(0, frc.shape[0], every)
    fref = ref.flat("frc").reshape(-1, 3)[sel]
    err = np.linalg.norm(frc[sel] - fref, axis=1)
    sc = scale[sel]
    bound = 1e-5 * (sc + np.linalg.norm(fref, axis=1) + 0.01 * float(np.median(sc))) + 1e-300
    worst = float(np.max(err / bound))
    say(f"C3: FP32 list forces vs FP64 oracle, {sel.size} particles: worst err/bound {worst:.3f}, "
        f"median {float(np.median(err / bound)):.2e}")
    assert sel.size >= 10000 and worst <= 1.0


def _ensemble_colony(c):
    """Colony c of the 256-colony ensemble: N_c = round(exp(U[ln 1e3, ln 6.4e4])), seed 1000 + c,
    motility U[0, 0.5], kT U[0.001, 0.05], growth rate U[2, 8] per time unit (daughters after
    0.19-0.75 time units), packing ~0.35 on a jittered lattice."""
    rng = np.random.default_rng(1000 + c)
    n = int(round(math.exp(rng.uniform(math.log(1e3), math.log(6.4e4)))))
    mot = rng.uniform(0.0, 0.5)
    kT = rng.uniform(0.001, 0.05)
    rate = rng.uniform(2.0, 8.0)
    side = int(math.ceil(math.sqrt(n)))
    a = 2.1  # lattice spacing
    L = side * a
    p = configs.base_params(2, 1, 2)
    sp = p.species[0]
    sp.spring, sp.k_harm = abi.SPRING_HARMONIC, 10.0
    sp.motility[1] = mot
    for e in range(2):
        sp.eps_core[e], sp.eps_well[e], sp.sig_well[e] = 1.0, 0.0, 1.0
    sp.kT = kT
    sp.growth_rate, sp.l_div, sp.l_birth, sp.sigma_jitter = rate, 2.0, 0.5, 0.05
    configs.set_law(p, abi.LAW_SC, 1.0)
    p.dt = 1e-4
    p.noise = abi.NOISE_FDT
    p.seed = 1000 + c
    p.division = abi.DIV_LINEAR_GROWTH
    p.precision = abi.PREC_FP32
    box = configs.make_box([L, L], [1, 1])
    cap = 3 * n + 64
    col = Colony(2, 2, [cap], ncells=[n])
    s = col.species[0]
    k = np.arange(n)
    centre = (np.stack([k % side, k // side], axis=1) + 0.5) * a + rng.uniform(-0.2, 0.2, (n, 2))
    ang = rng.uniform(0, 2 * math.pi, n)
    u = np.stack([np.cos(ang), np.sin(ang)], axis=1)
    s["pos"][:n, 0] = np.mod(centre - 0.25 * u, L)
    s["pos"][:n, 1] = np.mod(centre + 0.25 * u, L)
    s["sigma"][:n] = rng.uniform(0.9, 1.1, (n, 2))
    s["rest_len"][:n] = rng.uniform(0.5, 2.0, n)
    s["state"][:n] = abi.CS_CRAWLING
    return p, box, col


def test_c4_256_colony_ensemble(gpu):
    t0 = time.perf_counter()
    setups = [_ensemble_colony(c) for c in range(256)]
    ncells = [col.ncells()[0] for _, _, col in setups]
    say(f"C4: built 256 colonies, {sum(ncells)} cells ({min(ncells)}-{max(ncells)}), "
        f"{time.perf_counter() - t0:.1f} s")
    members = []
    for p, box, col in setups:
        md = HipMD(p, box, [sp["capacity"] for sp in col.species])
        md.upload(col)
        md.forces(want=False)
        members.append(md)
    batch = HipBatch(members)
    tuned = batch.autotune(reps=2)
    say(f"C4: uploaded; batch autotune {tuned}")
    batch.step(1000, [0] * 256)  # warm-up: graph capture and the first divisions
    for m in members:
        m.num_cells()
    t0 = time.perf_counter()
    ts = batch.step(1000, [1000] * 256)
    for m in members:
        m.num_cells()
    el = time.perf_counter() - t0
    nparts = sum(2 * sum(m.num_cells()) for m in members)
    say(f"C4: 1000 batch steps in {el:.2f} s: {nparts * 1000 / el:.3e} particle-steps/s over {nparts} particles")
    assert ts == [2000] * 256
    grown = sum(sum(m.num_cells()) for m in members)
    assert grown > sum(ncells)
    # sampled colonies (the smallest, the largest and three others) stepped alone
    order = np.argsort(ncells)
    for c in (int(order[0]), int(order[-1]), 7, 100, 200):
        p, box, col = setups[c]
        alone = HipMD(p, box, [sp["capacity"] for sp in col.species])
        alone.upload(col)
        alone.forces(want=False)
        alone.set_tuning(tuned["bins_per_cutoff"], tuned["block"])
        alone.set_kernel(tuned["variant"])
        alone.step(2000, 0)
        a, b = col.copy(), col.copy()
        alone.download(a)
        members[c].download(b)
        alone.close()
        assert a.ncells() == b.ncells(), c
        for f in ("pos", "sigma", "rest_len", "t0"):
            assert np.array_equal(a.flat(f), b.flat(f)), (c, f)
        say(f"C4: colony {c} ({ncells[c]} -> {a.ncells()[0]} cells) bit-equal to its single run")
    batch.close()
    for m in members:
        m.close()


def test_c5_surrogate_judged_every_macro_step_is_plain(gpu):
    """Judged at every macro-step with an untrained network: each judgement fails, the next
    macro-step falls back to the plain integrator and the one after is judged again; the judge
    keeps the plain result, so 6 macro-steps = 48 plain steps, bit-identical to a plain run."""
    p, box, col = configs.relaxed_bulk_3d(1 << 20, seed=1, precision=abi.PREC_FP32)
    assert p.species[0].kT == 0.05
    ends = []
    for mode in ("plain", "surrogate"):
        md = HipMD(p, box, [sp["capacity"] for sp in col.species])
        md.upload(col)
        md.forces(want=False)
        if mode == "plain":
            ts = md.step(8 * 6, 0)
        else:
            md.mlp_init(3)
            ts, st = md.surrogate_step(6, 0, judge_every=1, tol=1e-3)
            say(f"C5: surrogate stats {st}")
            assert st["judged"] == 3 and st["failed"] == 3 and st["macro_steps"] == 0 and st["plain_steps"] == 48
            assert st["last_err_rms"] > 1e-3
        assert ts == 48
        c = col.copy()
        md.download(c)
        md.close()
        ends.append(c)
    assert np.array_equal(ends[0].flat("pos"), ends[1].flat("pos"))
