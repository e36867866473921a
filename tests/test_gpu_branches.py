"""GPU parity of the step's less common branches, through the C ABI against the CPU oracle.

Bars as in test_gpu_parity: forces within 1e-12 (FP64) / 1e-5 (FP32) of the force scale,
trajectories within 1e-9 sigma (FP64) over their length, division events bit-exact.
Branches covered here, each switched on in a run of its own:
  - pressure-dependent friction zeta_p (p / p0)^beta in the mobility (CIL.cu:921);
  - force-shifted LJ and soft-core laws (OPT::FORCE_SHIFTED, MD.cuh:100-104, 200-202), in the
    FP64 kernel, the FP32 stencil kernel and the FP32 neighbour-list segments;
  - active radial / angular noise of two-disc cells (CIL.cu:803-814, MD.cu:573-578), drawn
    from counters and replayed from the reference's dSFMT stream;
  - the CIL motility flag (MD.cu:569-572: unscaled axis);
  - the Dividing state of the two-disc bond (MD.cu:547,560-568) and the two-disc division
    cycle calculDivisions (CIL.cu:369-455) with successful and failed divisions and swelling
    (updateDiameters, CIL.cu:970-990);
  - the stochastic Heun integrator (CIL.cu:1028-1049), alone and with the division cycle, in
    FP64 against the oracle and in the FP32 neighbour-list mode against the FP32 stencil mode;
  - batches of division-cycle colonies against the same colonies stepped alone.
"""
import math
import sys

import numpy as np
import pytest

from cil_cuda_amd import abi, configs
from cil_cuda_amd.colony import HipBatch, HipMD
from oracle import oracle as O

sys.path.insert(0, __file__.rsplit("/", 1)[0])
from test_gpu_parity import _force_parity, _max_disp_err, _random_colony, _trajectory  # noqa: E402

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _events_equal(ref, got):
    assert ref.ncells() == got.ncells(), (ref.ncells(), got.ncells())
    for a, b in zip(ref.species, got.species):
        n = a["ncells"]
        for f in ("state", "t0", "tswitch"):
            assert np.array_equal(a[f][:n], b[f][:n]), f
        assert np.allclose(a["sigma"][:n], b["sigma"][:n], rtol=1e-12, atol=1e-13)


# ------------------------------------------------------------------------------------------
# pressure friction
# ------------------------------------------------------------------------------------------
def test_pressure_friction_trajectory(gpu):
    say("trajectory with pressure friction")
    p, box, col = _random_colony(2, 1, 300, 16.0, 1, abi.LAW_HZ, seed=41, nspec=1)
    sp = p.species[0]
    # no adhesion: a negative pressure would make (p / p0)^beta undefined, in the reference too
    sp.eps_core[0], sp.sig_well[0], sp.eps_well[0], sp.motility[0] = 200.0, 0.4, 0.0, 0.3
    sp.zeta_p[0], sp.zeta_p_beta[0], sp.cycle_p0 = 2.0, 1.5, 3.0
    col.species[0]["vel"][:, 0] = [1.0, 0.0]
    p.pressure = 1
    p.dt = 2e-4
    ref, got = _trajectory(p, box, col, 500)
    pr = ref.species[0]["pressure"][: ref.ncells()[0]]
    assert np.count_nonzero(pr > 0) > 20  # the term is active for many cells
    err = _max_disp_err(ref, got, box)
    say("max displacement error", err)
    assert err <= 1e-9


# ------------------------------------------------------------------------------------------
# force shift
# ------------------------------------------------------------------------------------------
@pytest.mark.parametrize("law", [abi.LAW_LJ, abi.LAW_SC])
@pytest.mark.parametrize("mode", ["fp64", "fp32", "list32"])
def test_force_shifted_laws(gpu, law, mode):
    """Shifted laws against the FP64 oracle.  FP64: the 1e-12 bar.  FP32: the shift subtracts
    f0(rc) rc / r from f0(r); near the cut-off both are of the order of f0(rc) while their
    difference vanishes, so the FP32 rounding of a pair term scales with |f0(r) r| + |f0(rc) rc|,
    not with the shifted term.  The FP32 bar's force scale is therefore the sum of the unshifted
    and the shifted scales, which bounds sum |f0(rc) rc| from above."""
    say("force shifted law", law, mode)
    p, box, col = _random_colony(3, 1, 600, 9.0, 1, law, seed=51 + law)
    p.force_shifted = 1
    if law == abi.LAW_SC:
        for s in range(2):
            p.species[s].eps_well[0] = 0.5  # a well, so the shift is not zero
    if mode == "fp64":
        worst, *_ = _force_parity(p, box, col, 1e-12)
    else:
        p.precision = abi.PREC_FP32
        orc = O.Oracle(p, box, traversal=1)
        ref = col.copy()
        orc.forces(ref, 0)
        q = abi.Params.from_buffer_copy(p)
        q.force_shifted = 0
        scale = orc.force_scale(ref) + O.Oracle(q, box, traversal=1).force_scale(ref)
        md = HipMD(p, box, [sp["capacity"] for sp in col.species])
        md.upload(col)
        if mode == "list32":
            md.set_neighbour_list(True, 0.2, 512)
        frc, _ = md.forces()
        md.close()
        fref = ref.flat("frc").reshape(-1, 3)
        err = np.linalg.norm(frc - fref, axis=1)
        bound = 1e-5 * (scale + np.linalg.norm(fref, axis=1) + 0.01 * float(np.median(scale))) + 1e-300
        worst = float(np.max(err / bound))
    say("worst", worst)
    assert worst <= 1.0, worst


def test_force_shifted_trajectory_fp64(gpu):
    p, box, col = _random_colony(3, 1, 500, 9.0, 1, abi.LAW_SC, seed=53)
    p.force_shifted = 1
    for s in range(2):
        p.species[s].eps_well[0] = 0.5
    p.dt = 1e-4
    ref, got = _trajectory(p, box, col, 300)
    assert _max_disp_err(ref, got, box) <= 1e-9


# ------------------------------------------------------------------------------------------
# active noise and the CIL flag
# ------------------------------------------------------------------------------------------
def _fene_colony(seed, cil=0):
    p, box, col = configs.plain_2d_sc(n_cells=150, L=18.0, seed=seed, noise=abi.NOISE_NONE)
    sp = p.species[0]
    sp.spring = abi.SPRING_FENE
    sp.kappa, sp.kappa_push, sp.spring_ratio, sp.r2max, sp.rss = 1.0, 5.0, 0.3, 2.0, 0.5
    sp.motility[0], sp.motility[1] = 0.0, 0.1
    sp.cil = cil
    return p, box, col


def test_active_noise_counter_draws(gpu):
    say("active noise, counter draws")
    p, box, col = _fene_colony(11)
    p.active_noise = 1
    p.species[0].d_noise_radial, p.species[0].d_noise_angular = 0.3, 0.2
    ref, got = _trajectory(p, box, col, 500)
    err = _max_disp_err(ref, got, box)
    say("max displacement error", err)
    assert err <= 1e-9


def test_active_noise_dsfmt_replay(gpu):
    """Passive and active draws from the reference's dSFMT stream, in its order: every passive
    component, then every radial, then every angular draw (CIL.cu:793-818)."""
    say("active noise, dsfmt replay")
    p, box, col = _fene_colony(12)
    n = col.ncells()[0]
    sp = p.species[0]
    p.noise = abi.NOISE_CELL_WIENER
    sp.d_noise = math.sqrt(2 * 0.005) / math.sqrt(p.dt)
    p.active_noise = 1
    sp.d_noise_radial, sp.d_noise_angular = 0.3, 0.2
    p.rng = abi.RNG_EXTERNAL
    nsteps = 400
    per = n * 2 + 2 * n
    z = O.dsfmt_fill(7, 1, per * (nsteps + 1)).reshape(nsteps + 1, per)
    ref, got = _trajectory(p, box, col, nsteps, ext=z)
    err = _max_disp_err(ref, got, box)
    say("max displacement error", err)
    assert err <= 1e-9


def test_cil_flag_trajectory(gpu):
    p, box, col = _fene_colony(13, cil=1)
    ref, got = _trajectory(p, box, col, 500)
    assert _max_disp_err(ref, got, box) <= 1e-9


# ------------------------------------------------------------------------------------------
# two-disc division cycle, Dividing-state bond, swelling; Euler and Heun
# ------------------------------------------------------------------------------------------
@pytest.mark.parametrize("integrator", [abi.INT_EULER, abi.INT_HEUN])
def test_binary_division_cycle(gpu, integrator):
    say("binary division cycle, integrator", integrator)
    p, box, col = configs.binary_2d_fene(integrator=integrator)
    ref, got = _trajectory(p, box, col, 600)
    n0 = col.ncells()[0]
    say("cells", n0, "->", ref.ncells()[0], "dividing", int((ref.species[0]["state"][: ref.ncells()[0]] == 0).sum()))
    assert ref.ncells()[0] > n0 + 50
    _events_equal(ref, got)
    err = _max_disp_err(ref, got, box)
    say("max displacement error", err)
    assert err <= 1e-9


@pytest.mark.parametrize("integrator", [abi.INT_EULER, abi.INT_HEUN])
def test_binary_division_failures(gpu, integrator):
    """Rmin out of reach: every Dividing cell fails at its switch time (binaryCellDivisionFail),
    goes back to crawling and starts again; no cell is added."""
    p, box, col = configs.binary_2d_fene(integrator=integrator, n_cells=100)
    p.species[0].division_r2min = 100.0
    p.species[0].division_tau.mu = 0.2
    ref, got = _trajectory(p, box, col, 500)
    assert ref.ncells() == col.ncells()
    t0 = ref.species[0]["tswitch"][:100]
    assert np.all(t0 > 500)  # every cell has switched at least twice
    _events_equal(ref, got)
    assert _max_disp_err(ref, got, box) <= 1e-9


def test_heun_plain_trajectory(gpu):
    say("heun trajectory, reference FENE cells with noise")
    p, box, col = _fene_colony(14)
    p.integrator = abi.INT_HEUN
    p.noise = abi.NOISE_CELL_WIENER
    p.species[0].d_noise = math.sqrt(2 * 0.005) / math.sqrt(p.dt)
    ref, got = _trajectory(p, box, col, 1000)
    err = _max_disp_err(ref, got, box)
    say("max displacement error", err)
    assert err <= 1e-9


def test_heun_list_mode_matches_stencil(gpu):
    """FP32 Heun with the division cycle: the neighbour-list run (two force passes per step,
    alternate rebuild flags) against the stencil run: same events, positions within 1e-5."""
    p, box, col = configs.binary_2d_fene(integrator=abi.INT_HEUN, precision=abi.PREC_FP32)
    out = []
    for use_list in (False, True):
        md = HipMD(p, box, [sp["capacity"] for sp in col.species])
        md.upload(col)
        if use_list:
            md.set_neighbour_list(True, 0.1, 64)
        md.forces(want=False)
        ts = 0
        for _ in range(6):
            ts = md.step(100, ts)
        c = col.copy()
        md.download(c)
        md.close()
        out.append(c)
    say("heun list: cells", out[0].ncells(), out[1].ncells())
    _events_equal(out[0], out[1])
    err = _max_disp_err(out[0], out[1], box)
    say("list vs stencil", err)
    assert err <= 1e-5


def test_heun_rejects_fdt_and_external_draws(gpu):
    p, box, col = configs.binary_2d_fene(integrator=abi.INT_HEUN)
    for field, val in (("noise", abi.NOISE_FDT), ("rng", abi.RNG_EXTERNAL)):
        q = abi.Params.from_buffer_copy(p)
        setattr(q, field, val)
        with pytest.raises(abi.CilError):
            HipMD(q, box, [64])


def test_batch_binary_division_matches_single(gpu):
    say("batch of division-cycle colonies")
    members, singles = [], []
    setups = [configs.binary_2d_fene(n_cells=n, L=L, seed=s) for n, L, s in ((60, 14.0, 1), (150, 20.0, 2), (20, 9.0, 3))]
    for p, box, col in setups:
        for lst in (members, singles):
            md = HipMD(p, box, [sp["capacity"] for sp in col.species])
            md.upload(col)
            md.forces(want=False)
            lst.append(md)
    batch = HipBatch(members)
    ts = batch.step(300, [0] * len(members))
    for md in singles:
        md.step(300, 0)
    for (p, box, col), a, b in zip(setups, members, singles):
        ca, cb = col.copy(), col.copy()
        a.download(ca)
        b.download(cb)
        assert ca.ncells() == cb.ncells() and ca.ncells()[0] > col.ncells()[0]
        for f in ("pos", "sigma", "state", "tswitch", "t0"):
            assert np.array_equal(ca.species[0][f][: ca.ncells()[0]], cb.species[0][f][: cb.ncells()[0]]), f
    batch.close()
    for md in members + singles:
        md.close()
    assert ts == [300] * len(members)
