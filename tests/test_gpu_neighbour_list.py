"""GPU parity of the neighbour-list mode (FP32) through the C ABI.

Bars (the FP32 force bar of test_gpu_parity: |F - F_ref| <= 1e-5 (S_p + |F_ref| + 1 % of the
median S_p) against the FP64 oracle):
  - forces with the list built on the spot, for every law, 3D periodic and open boxes, one and
    two discs per cell, including a box small enough for single-bin minimum-image dimensions;
  - forces of steps that reuse a stale list (particles up to skin / 2 outside their
    build-time bins) against the FP64 oracle;
  - trajectories: over 300 noisy steps with a small skin (frequent rebuilds) the list mode,
    the FP32 stencil kernel and an FP64 run agree within 1e-4 sigma (FP32 rounding, amplified
    by the discontinuous LJ cut-off), and the forces of the end state still meet the bar;
  - division and drain: growing channels give the same division events and cell counts as
    the stencil kernel, positions within 1e-5 sigma;
  - two-level lists: bit-identical to one-level lists of the same skin.
"""
import sys

import numpy as np
import pytest

from cil_cuda_amd import abi, configs
from cil_cuda_amd.colony import HipMD, _check
from oracle import oracle as O

sys.path.insert(0, __file__.rsplit("/", 1)[0])
from test_gpu_parity import _max_disp_err, _random_colony  # noqa: E402

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _force_bar(p, box, col, frc, tol=1e-5):
    orc = O.Oracle(p, box, traversal=1)
    ref = col.copy()
    orc.forces(ref, 0)
    scale = orc.force_scale(ref)
    fref = ref.flat("frc").reshape(-1, col.dim)
    err = np.linalg.norm(frc - fref, axis=1)
    bound = tol * (scale + np.linalg.norm(fref, axis=1) + 0.01 * float(np.median(scale))) + 1e-300
    return float(np.max(err / bound))


def _list_forces(p, box, col, skin=0.3, max_nb=256):
    md = HipMD(p, box, [sp["capacity"] for sp in col.species])
    md.upload(col)
    md.set_neighbour_list(True, skin, max_nb)
    frc, _ = md.forces()
    md.close()
    return frc


@pytest.mark.parametrize("law", [abi.LAW_LJ, abi.LAW_SC, abi.LAW_HZ])
@pytest.mark.parametrize("periodic", [1, 0])
@pytest.mark.parametrize("nelem", [1, 2])
def test_list_forces_match_fp64(gpu, law, periodic, nelem):
    say("list forces law", law, "periodic", periodic, "nelem", nelem)
    p, box, col = _random_colony(3, nelem, 800, 9.0, periodic, law, seed=31 + law)
    p.precision = abi.PREC_FP32
    worst = _force_bar(p, box, col, _list_forces(p, box, col))
    assert worst <= 1.0, worst


def test_list_forces_small_periodic_box(gpu):
    # a box of fewer than 2 * reach + 1 bins per dimension: single-bin minimum-image path
    p, box, col = _random_colony(3, 1, 120, 4.5, 1, abi.LAW_LJ, seed=5)
    p.precision = abi.PREC_FP32
    worst = _force_bar(p, box, col, _list_forces(p, box, col))
    assert worst <= 1.0, worst


def _relaxed(n, seed):
    p, box, col = configs.relaxed_bulk_3d(n, seed=seed, precision=abi.PREC_FP32)
    return p, box, col


def test_list_trajectory_matches_stencil_with_rebuilds(gpu):
    """300 noisy steps with skin 0.02 (a rebuild every few steps): the list mode, the FP32
    stencil kernel and an FP64 run stay within 1e-4 sigma of each other.  FP32 paths differ by
    rounding (~1e-8 sigma per step); the unshifted LJ cut-off turns a pair that crosses rc one
    step earlier in one run into a ~5e-6 sigma jump, so FP32 trajectories agree only to that
    order (the 1e-9 sigma trajectory bar is the FP64 mode's).  The end-state forces still meet
    the force bar."""
    p, box, col = _relaxed(8192, 4)
    out = {}
    for mode in ("fp64", "stencil32", "list32"):
        q = abi.Params.from_buffer_copy(p)
        q.precision = abi.PREC_FP64 if mode == "fp64" else abi.PREC_FP32
        md = HipMD(q, box, [sp["capacity"] for sp in col.species])
        md.upload(col)
        if mode == "list32":
            md.set_neighbour_list(True, 0.02, 256)
        md.forces(want=False)
        assert md.step(300, 0) == 300
        c = col.copy()
        md.download(c)
        md.close()
        out[mode] = c
    e_stencil = _max_disp_err(out["fp64"], out["stencil32"], box)
    e_list = _max_disp_err(out["fp64"], out["list32"], box)
    e_pair = _max_disp_err(out["stencil32"], out["list32"], box)
    say(f"after 300 steps, max displacement vs FP64: stencil {e_stencil:.3e}, list {e_list:.3e}; "
        f"list vs stencil {e_pair:.3e}")
    assert max(e_stencil, e_list, e_pair) <= 1e-4
    end = out["list32"]
    worst = _force_bar(p, box, end, _list_forces(p, box, end))
    assert worst <= 1.0, worst


@pytest.mark.parametrize("skin,nsteps", [(0.3, 20), (0.02, 37)])
def test_stale_list_forces_match_fp64(gpu, skin, nsteps):
    """Forces of the last step's list pass (no fresh build: particles up to skin / 2 outside
    their build-time bins, coordinates refreshed in place) against the FP64 oracle on the
    downloaded state; download materialises those lazy forces without a new force pass."""
    p, box, col = _relaxed(8192, 4)
    md = HipMD(p, box, [sp["capacity"] for sp in col.species])
    md.upload(col)
    md.set_neighbour_list(True, skin, 256)
    md.forces(want=False)
    md.step(nsteps, 0)
    c = col.copy()
    md.download(c)
    md.close()
    worst = _force_bar(p, box, c, c.flat("frc").reshape(-1, 3))
    say(f"stale list (skin {skin}, {nsteps} steps): worst {worst:.3f}")
    assert worst <= 1.0, worst


def test_list_graph_and_eager_paths_agree(gpu):
    # 64-step graph replays and single eager steps must give the same states
    p, box, col = _relaxed(4096, 6)
    res = []
    for chunks in ((64, 64), (1,) * 128):
        md = HipMD(p, box, [sp["capacity"] for sp in col.species])
        md.upload(col)
        md.set_neighbour_list(True, 0.05, 256)
        md.forces(want=False)
        ts = 0
        for n in chunks:
            ts = md.step(n, ts)
        c = col.copy()
        md.download(c)
        md.close()
        res.append(c.flat("pos"))
    assert np.array_equal(res[0], res[1])


def test_list_division_and_drain_match_stencil(gpu):
    p, box, col = configs.channel_2d_hz(width=12.0, length=30.0, capacity=512, dt=1e-3, tau=0.05,
                                        precision=abi.PREC_FP32)
    out = []
    for use_list in (False, True):
        md = HipMD(p, box, [sp["capacity"] for sp in col.species])
        md.upload(col)
        if use_list:
            md.set_neighbour_list(True, 0.1, 128)
        md.forces(want=False)
        ts = 0
        for _ in range(6):
            ts = md.step(50, ts)
            md.check_positions()
        c = col.copy()
        md.download(c)
        md.close()
        out.append(c)
    say("channel cells", out[0].ncells(), out[1].ncells())
    assert out[0].ncells() == out[1].ncells()
    assert out[0].ncells()[0] > 8
    assert _max_disp_err(out[0], out[1], box) <= 1e-5


def test_list_overflow_is_reported(gpu):
    # 8 entries per particle cannot hold a dense bulk's lists: the build stores at most 8 and
    # raises the list error, which the step returns as a capacity error with a message
    p, box, col = _relaxed(4096, 6)
    md = HipMD(p, box, [sp["capacity"] for sp in col.species])
    md.upload(col)
    md.set_neighbour_list(True, 0.3, 8)
    with pytest.raises(abi.CilError) as ei:
        md.forces(want=False)
        md.step(1, 0)
    say("list overflow:", ei.value)
    assert "max_neighbours" in str(ei.value)
    md.close()


def test_list_capacity_in_groups_of_four(gpu):
    # max_neighbours counts list slots in whole groups of four: 157 is taken as 160.  A capacity
    # that holds every list gives the same lists, hence bit-identical steps, as a far larger one
    # (two source types here, each starting on a group boundary); rebuilds every few steps.
    p, box, col = _relaxed(4096, 6)
    out = []
    for m in (157, 160, 1024):
        md = HipMD(p, box, [sp["capacity"] for sp in col.species])
        md.upload(col)
        md.set_neighbour_list(True, 0.05, m)
        md.forces(want=False)
        assert md.step(40, 0) == 40
        c = col.copy()
        md.download(c)
        md.close()
        out.append(c.flat("pos"))
        say("capacity", m, "done")
    assert np.array_equal(out[0], out[1])
    assert np.array_equal(out[1], out[2])


def test_list_mode_rejects_fp64(gpu):
    p, box, col = configs.plain_2d_sc(n_cells=50, L=12.0, seed=3)
    md = HipMD(p, box, [sp["capacity"] for sp in col.species])
    with pytest.raises(abi.CilError):
        md.set_neighbour_list(True)
    md.close()


def _run_lists(p, box, col, nsteps, skin, inner):
    md = HipMD(p, box, [sp["capacity"] for sp in col.species])
    md.upload(col)
    md.forces(want=False)
    md.set_neighbour_list(True, skin, 256, inner_skin=inner)
    ts = md.step(nsteps, 0)
    frc, _ = md.forces()
    out = col.copy()
    md.download(out)
    md.close()
    assert ts == nsteps
    return out.flat("pos").reshape(-1, 3), frc


@pytest.mark.parametrize("inner", [0.02, 0.15, 0.45])
def test_two_level_lists_bit_identical_to_one_level(gpu, inner):
    """Two-level lists (outer skin 0.5) against one-level lists of the same skin: the inner list
    holds the outer list's entries within cut-off + inner skin in the same order, so every step
    sums the same pairs in the same order.  Over 1,500 noisy steps of a 32,768-particle bulk
    (kT 0.05), with inner refreshes from almost every step (inner skin 0.02) to rarely (0.45)
    and outer rebuilds in between, positions and the end forces are bit-identical."""
    p, box, col = configs.relaxed_bulk_3d(32768, seed=5, precision=abi.PREC_FP32)
    x1, f1 = _run_lists(p, box, col, 1500, 0.5, 0.0)
    x2, f2 = _run_lists(p, box, col, 1500, 0.5, inner)
    x0 = col.flat("pos").reshape(-1, 3)
    L = np.array(box.length[:3])
    d = x1 - x0
    d -= L * np.round(d / L)
    say(f"two-level lists, inner skin {inner}: largest move {np.abs(d).max():.3f} sigma; identical "
        f"{np.array_equal(x1, x2)} {np.array_equal(f1, f2)}")
    assert np.abs(d).max() > 0.3  # beyond half the outer skin: the outer list was rebuilt on the way
    assert np.array_equal(x1, x2)
    assert np.array_equal(f1, f2)


def test_two_level_lists_2d_two_disc_cells(gpu):
    """The same bit-identity in 2D with two-disc cells (harmonic bond, motility, soft core):
    2,000 cells, 3,000 steps of dt 1e-3, one-level lists of skin 0.4 against two-level lists of
    0.4 / 0.1; discs move beyond half the outer skin, so both lists are rebuilt on the way."""
    p, box, col = configs.plain_2d_sc(n_cells=2000, L=64.0, seed=3, dt=1e-3, precision=abi.PREC_FP32)
    out = []
    for inner in (0.0, 0.1):
        md = HipMD(p, box, [sp["capacity"] for sp in col.species])
        md.upload(col)
        md.forces(want=False)
        md.set_neighbour_list(True, 0.4, 256, inner_skin=inner)
        assert md.step(3000, 0) == 3000
        c = col.copy()
        md.download(c)
        md.close()
        out.append(c.flat("pos").reshape(-1, 2))
    d = out[0] - col.flat("pos").reshape(-1, 2)
    d -= 64.0 * np.round(d / 64.0)
    move = float(np.sqrt((d * d).sum(1)).max())
    say(f"two-level lists, 2D two-disc cells: largest move {move:.3f} sigma; identical {np.array_equal(out[0], out[1])}")
    assert move > 0.2
    assert np.array_equal(out[0], out[1])


def test_two_level_lists_reject_unsupported_setups(gpu):
    p, box, col = configs.relaxed_bulk_3d(4096, seed=6, precision=abi.PREC_FP32)
    md = HipMD(p, box, [sp["capacity"] for sp in col.species])
    md.upload(col)
    md.forces(want=False)
    with pytest.raises(abi.CilError):  # no list
        _check_inner(md, 0.1)
    md.set_neighbour_list(True, 0.3, 256)
    with pytest.raises(abi.CilError):  # inner skin not below the skin
        _check_inner(md, 0.3)
    _check_inner(md, 0.1)
    md.mlp_init(1)
    with pytest.raises(abi.CilError):  # the surrogate needs one-level lists
        md.surrogate_step(1, 0, judge_every=0, tol=1.0)
    _check_inner(md, 0.0)  # back to one level
    ts, _ = md.surrogate_step(1, 0, judge_every=0, tol=1.0)
    assert ts == 8
    md.close()
    q, box2, col2 = _theta_growing()
    m2 = HipMD(q, box2, [sp["capacity"] for sp in col2.species])
    m2.upload(col2)
    m2.set_neighbour_list(True, 0.3, 256)
    with pytest.raises(abi.CilError):  # division
        _check_inner(m2, 0.1)
    m2.close()


def _check_inner(md, skin_in):
    _check(md.L, md.ctx, md.L.cil_hip_set_neighbour_list_inner(md.ctx, float(skin_in)), "inner")


def _theta_growing(n_cells=120, L=24.0, seed=7):
    """Single-disc Hertz colony in the theta cell cycle with long cycles (1.5 time units =
    1500 steps) and phases in [0, 0.55): about 700 steps of diameter growth with no division
    and no motility, so nothing but the growth can make a list stale."""
    p, box, col = configs.theta_2d_hz(n_cells=n_cells, L=L, seed=seed, dt=1e-3, capacity=1024,
                                      precision=abi.PREC_FP32)
    sp = p.species[0]
    sp.motility[0] = 0.0
    sp.division_tau.mu = 1.5
    sp.division_tau.sigma = 0.1
    rng = np.random.default_rng(seed)
    s = col.species[0]
    s["theta"][:n_cells] = rng.uniform(0.0, 0.55, n_cells)
    s["ang_vel"][:n_cells] = 1.0 / 1500.0
    return p, box, col


def test_list_with_theta_cycle_growth_matches_stencil(gpu):
    """Neighbour lists under theta-cycle diameter growth: the cut-off of every pair grows with
    the diameters, so the growth since the build is charged against the skin with the
    displacement.  With a skin of 0.05 sigma the cut-offs outgrow it within ~100 steps while
    no particle moves; a list that only checked displacement would miss the new contacts and
    the discs would overlap without force.  Over 1000 steps (divisions from step ~700 on) the
    list run must track the FP32 stencil run: same cell counts, positions within 1e-4 sigma,
    and the end-state forces must hold every contact (within 1e-3 of the force scale of the
    FP64 oracle)."""
    p, box, col = _theta_growing()
    out = []
    for use_list in (False, True):
        md = HipMD(p, box, [sp["capacity"] for sp in col.species])
        md.upload(col)
        if use_list:
            md.set_neighbour_list(True, 0.05, 64)
        md.forces(want=False)
        ts = 0
        for _ in range(10):
            ts = md.step(100, ts)
        c = col.copy()
        md.download(c)
        md.close()
        out.append(c)
    n0, n1 = out[0].ncells()[0], out[1].ncells()[0]
    say("theta + list: cells", n0, n1)
    assert n0 == n1 and n0 > 120
    s0, s1 = out[0].species[0], out[1].species[0]
    dth = float(np.max(np.abs(s0["theta"][:n0] - s1["theta"][:n1])))
    assert dth <= 1e-6, dth
    err = _max_disp_err(out[0], out[1], box)
    dsig = float(np.max(np.abs(s0["sigma"][:n0] - s1["sigma"][:n1])))
    say(f"list vs stencil: positions {err:.3e}, diameters {dsig:.3e}; "
        f"diameters grew to {float(np.max(s1['sigma'][:n1])):.3f}")
    assert err <= 1e-4 and dsig <= 1e-6
    # end-state forces against the FP64 oracle: a contact missing from the list would be an error
    # of the order of its own term; FP32 rounding of shallow Hertz overlaps between discs of
    # diameter ~1.6 (h = sigma_bar - r rounded to ~1e-7) stays below 1e-3 of the force scale
    q = abi.Params.from_buffer_copy(p)
    q.division = abi.DIV_NONE
    worst = _force_bar(q, box, out[1], out[1].flat("frc").reshape(-1, 2), tol=1e-3)
    say(f"end-state list forces vs FP64 oracle (bar 1e-3): worst {worst:.3f}")
    assert worst <= 1.0, worst


def test_two_list_contexts_step_concurrently(gpu):
    """Two list-mode contexts (each with its own stream and persistent sort kernel) stepped from
    two host threads at once give the states they give when stepped one after the other."""
    import threading

    setups = [_relaxed(32768, s) for s in (11, 12)]

    def make(k):
        p, box, col = setups[k]
        md = HipMD(p, box, [sp["capacity"] for sp in col.species])
        md.upload(col)
        md.set_neighbour_list(True, 0.05, 256)
        md.forces(want=False)
        return md

    def run(md, out, k):
        ts = 0
        for _ in range(10):
            ts = md.step(50, ts)
        out[k] = md.download(setups[k][2].copy()).flat("pos")

    seq, par = {}, {}
    for k in (0, 1):
        md = make(k)
        run(md, seq, k)
        md.close()
    mds = [make(0), make(1)]
    th = [threading.Thread(target=run, args=(mds[k], par, k)) for k in (0, 1)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    for md in mds:
        md.close()
    assert set(par) == {0, 1}
    for k in (0, 1):
        assert np.array_equal(seq[k], par[k]), k


@pytest.mark.parametrize("skin,inner", [(0.3, 0.0), (0.6, 0.2)])
def test_list_autotune_keeps_the_state(gpu, skin, inner):
    """The list-mode autotuner times rebuild, plain (and, with two-level lists, refresh) passes
    at 1-3 bins per cut-off at the current positions: the state is not advanced, the forces it
    leaves are those of a fresh context with the chosen bins (bit for bit), and stepping
    continues."""
    p, box, col = _relaxed(8192, 4)
    md = HipMD(p, box, [sp["capacity"] for sp in col.species])
    md.upload(col)
    md.set_neighbour_list(True, skin, 256, inner_skin=inner)
    md.forces(want=False)
    before = md.download(col.copy())
    best = md.autotune(reps=2)
    say(f"list autotune: {best}")
    assert best["bins_per_cutoff"] in (1, 2, 3) and best["force_pass_us"] > 0.0
    after = md.download(col.copy())
    assert np.array_equal(before.flat("pos"), after.flat("pos"))
    fresh = HipMD(p, box, [sp["capacity"] for sp in col.species])
    fresh.set_tuning(best["bins_per_cutoff"], 64)
    fresh.upload(before)
    fresh.set_neighbour_list(True, skin, 256, inner_skin=inner)
    fresh.forces(want=False)
    ref = fresh.download(col.copy())
    fresh.close()
    assert np.array_equal(after.flat("frc"), ref.flat("frc"))
    say(f"list autotune: forces against FP64, worst {_force_bar(p, box, after, after.flat('frc').reshape(-1, 3)):.3f}")
    assert md.step(50, 0) == 50
    md.close()
