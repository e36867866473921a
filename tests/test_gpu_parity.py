"""GPU parity: the HIP path through the C ABI against the CPU oracle on identical inputs.

Bars (written here, per the method's contract):
  - interacting pair sets and division / removal events: bit-exact;
  - per-particle forces: |F_gpu - F_ref| <= tol * (S_p + |F_ref|), tol = 1e-12 (FP64 mode)
    or 1e-5 (FP32 mode), S_p = sum over the particle's pair terms |f0 r_ij|;
  - trajectories (noise-free, counter-noise, dSFMT replay): max displacement difference
    <= 1e-9 sigma over 10^3 steps.
The oracle uses its all-pairs traversal (canonical i<j orientation, CIL.cu:232-263).
"""
import math
import sys

import numpy as np
import pytest

from cil_cuda_amd import abi, configs
from cil_cuda_amd.colony import Colony, HipMD
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _minimg(d, box, dim):
    d = d.copy()
    for k in range(dim):
        if box.periodic[k]:
            L = box.length[k]
            d[..., k] -= L * np.round(d[..., k] / L)
    return d


def _random_colony(dim, nelem, n, L, periodic, law, seed, nspec=2):
    rng = np.random.default_rng(seed)
    p = configs.base_params(dim, nspec, nelem)
    for s in range(nspec):
        for e in range(nelem):
            p.species[s].sig_well[e] = 0.5 + 0.2 * s
            p.species[s].eps_core[e] = 10.0
            p.species[s].eps_well[e] = 1.0
    configs.set_law(p, law, 2.0 if law == abi.LAW_LJ else 1.0)
    box = configs.make_box([L] * dim, [periodic] * dim)
    per = n // nspec
    from cil_cuda_amd.colony import Colony

    col = Colony(dim, nelem, [per] * nspec, ncells=[per] * nspec)
    for sp in col.species:
        c = rng.uniform(0, L, (per, dim))
        sp["pos"][:, 0] = c
        if nelem == 2:
            u = rng.normal(size=(per, dim))
            u /= np.linalg.norm(u, axis=1, keepdims=True)
            sp["pos"][:, 1] = np.mod(c + 0.6 * u, L) if periodic else c + 0.6 * u
        sp["sigma"][:] = rng.uniform(0.8, 1.25, (per, nelem))
        sp["rest_len"][:] = 0.6
        sp["state"][:] = abi.CS_CRAWLING
    return p, box, col


@pytest.mark.parametrize("mode", ["fp64", "fp32", "list32"])
@pytest.mark.parametrize("n,L", [(12000, 14.0), (300, 3.4)])
def test_forces_crowded_buckets(gpu, mode, n, L):
    """Buckets beyond the bin sort's per-lane network (more than 8 particles: wave-wide sort in
    registers) and beyond one wave (more than 64: sort in LDS): bins of one cut-off in a dense
    box, and a box of a single bin.  Forces against the FP64 oracle.  Soft core: these random
    colonies overlap, and the LJ r^-13 wall would turn FP32 coordinate rounding at near-zero
    distances into errors beyond the FP32 bar, which is not what this test is about."""
    p, box, col = _random_colony(3, 1, n, L, 1, abi.LAW_SC, seed=17)
    p.precision = abi.PREC_FP64 if mode == "fp64" else abi.PREC_FP32
    tol = 1e-12 if mode == "fp64" else 1e-5
    orc = O.Oracle(p, box, traversal=1)
    ref = col.copy()
    orc.forces(ref, 0)
    scale = orc.force_scale(ref)
    md = HipMD(p, box, [sp["capacity"] for sp in col.species])
    md.set_tuning(1, 64)
    md.upload(col)
    if mode == "list32":
        md.set_neighbour_list(True, 0.2, 1024)
    nb, _ = md.bin_info()
    frc, _ = md.forces()
    md.close()
    per_bucket = n / (2 * int(np.prod(nb)))
    say(f"crowded {mode}: n {n} bins {nb} -> {per_bucket:.0f} particles per bucket")
    assert per_bucket > 8
    fref = ref.flat("frc").reshape(-1, 3)
    err = np.linalg.norm(frc - fref, axis=1)
    floor = 0.01 * float(np.median(scale)) if mode != "fp64" else 0.0
    worst = float(np.max(err / (tol * (scale + np.linalg.norm(fref, axis=1) + floor) + 1e-300)))
    assert worst <= 1.0, worst


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("periodic", [1, 0])
@pytest.mark.parametrize("nelem", [1, 2])
@pytest.mark.parametrize("law", [abi.LAW_LJ, abi.LAW_SC, abi.LAW_HZ])
def test_pairs_bit_exact(gpu, dim, periodic, nelem, law):
    n = 400 if dim == 2 else 600
    L = 18.0 if dim == 2 else 9.0
    say(f"pairs dim={dim} pbc={periodic} nelem={nelem} law={law}")
    p, box, col = _random_colony(dim, nelem, n, L, periodic, law, seed=100 + dim * 10 + nelem + law)
    ref = O.Oracle(p, box, traversal=1).pairs(col, mode=1)
    md = HipMD(p, box, [sp["capacity"] for sp in col.species])
    md.upload(col)
    got = md.pairs()
    md.close()
    assert len(ref) > 0
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("mode", ["fp64", "fp32", "list32"])
def test_empty_and_ragged_species(gpu, mode):
    """Ragged input: species 0 empty (its region allocated), species 1 populated.  Forces meet
    the mode's bar against the oracle, 40 steps match the oracle's trajectory (FP64: 1e-9 sigma;
    FP32 modes: 1e-5 sigma), and a colony with every species empty steps without error and
    stays empty."""
    p, box, col = _random_colony(3, 1, 600, 8.0, 1, abi.LAW_SC, seed=23)
    p.precision = abi.PREC_FP64 if mode == "fp64" else abi.PREC_FP32
    col.species[0]["ncells"] = 0
    assert col.ncells() == [0, 300]
    tol = 1e-12 if mode == "fp64" else 1e-5
    orc = O.Oracle(p, box, traversal=1)
    ref = col.copy()
    orc.forces(ref, 0)
    scale = orc.force_scale(ref)
    md = HipMD(p, box, [sp["capacity"] for sp in col.species])
    md.upload(col)
    if mode == "list32":
        md.set_neighbour_list(True, 0.2, 256)
    frc, _ = md.forces()
    fref = ref.flat("frc").reshape(-1, 3)
    assert frc.shape == fref.shape == (300, 3)
    err = np.linalg.norm(frc - fref, axis=1)
    floor = 0.01 * float(np.median(scale)) if mode != "fp64" else 0.0
    worst = float(np.max(err / (tol * (scale + np.linalg.norm(fref, axis=1) + floor) + 1e-300)))
    say(f"ragged {mode}: worst force err/bound {worst:.3f}")
    assert worst <= 1.0, worst
    assert orc.step(ref, 40, 0) == 40
    assert md.step(40, 0) == 40
    got = col.copy()
    md.download(got)
    md.close()
    assert got.ncells() == ref.ncells() == [0, 300]
    d = _max_disp_err(ref, got, box)
    say(f"ragged {mode}: 40-step max displacement difference {d:.3e}")
    assert d <= (1e-9 if mode == "fp64" else 1e-5)

    for sp in col.species:
        sp["ncells"] = 0
    md = HipMD(p, box, [sp["capacity"] for sp in col.species])
    md.upload(col)
    if mode == "list32":
        md.set_neighbour_list(True, 0.2, 256)
    md.forces(want=False)
    assert md.step(10, 0) == 10
    got = col.copy()
    md.download(got)
    md.close()
    assert got.ncells() == [0, 0]


def _force_parity(p, box, col, tol, ts=0):
    """Worst err/bound over particles.  In FP32 mode the bound has an absolute floor of 1% of
    the median force scale: a grazing contact (S_p -> 0) carries FP32's absolute rounding
    of r - sigma_bar (~1e-7 x the law's force scale), which no relative bound can hold."""
    orc = O.Oracle(p, box, traversal=1)
    ref = col.copy()
    orc.forces(ref, ts)
    scale = orc.force_scale(ref)
    md = HipMD(p, box, [sp["capacity"] for sp in col.species])
    md.upload(col)
    frc, pr = md.forces()
    md.close()
    fref = ref.flat("frc").reshape(-1, col.dim)
    err = np.linalg.norm(frc - fref, axis=1)
    floor = 0.01 * float(np.median(scale)) if p.precision == abi.PREC_FP32 else 0.0
    bound = tol * (scale + np.linalg.norm(fref, axis=1) + floor) + 1e-300
    worst = np.max(err / bound)
    return worst, frc, fref, pr, ref.flat("pressure")


@pytest.mark.parametrize("precision,tol", [(abi.PREC_FP64, 1e-12), (abi.PREC_FP32, 1e-5)])
def test_forces_plain_2d(gpu, precision, tol):
    say("forces plain_2d precision", precision)
    p, box, col = configs.plain_2d_sc(n_cells=300, L=24.0, seed=2, precision=precision)
    worst, *_ = _force_parity(p, box, col, tol)
    assert worst <= 1.0, worst


@pytest.mark.parametrize("law", [abi.LAW_LJ, abi.LAW_SC, abi.LAW_HZ, abi.LAW_EA])
@pytest.mark.parametrize("dim", [2, 3])
def test_forces_laws_fp64(gpu, law, dim):
    say("forces law", law, "dim", dim)
    p, box, col = _random_colony(dim, 1, 300 if dim == 2 else 500, 16.0 if dim == 2 else 8.0, 1, law, seed=7 + law)
    p.pressure = 1
    p.species[1].cross_adhesion = 0.5  # asymmetric adhesion scaling between species
    worst, frc, fref, pr, prr = _force_parity(p, box, col, 1e-12)
    assert worst <= 1.0, worst
    if law in (abi.LAW_HZ, abi.LAW_EA):
        assert np.allclose(pr, prr, rtol=1e-11, atol=1e-11 * np.max(np.abs(prr)))


@pytest.mark.parametrize("law", [abi.LAW_LJ, abi.LAW_SC, abi.LAW_HZ])
def test_forces_laws_fp32(gpu, law):
    say("forces fp32 law", law)
    p, box, col = _random_colony(3, 1, 800, 9.0, 1, law, seed=17 + law)
    p.precision = abi.PREC_FP32
    worst, *_ = _force_parity(p, box, col, 1e-5)
    assert worst <= 1.0, worst


def _trajectory(p, box, col, nsteps, ext=None):
    orc = O.Oracle(p, box, traversal=1)
    ref = col.copy()
    md = HipMD(p, box, [sp["capacity"] for sp in col.species])
    md.upload(col)
    if ext is not None:
        orc.set_external_noise(ext)
        md.set_external_noise(ext)
    orc.forces(ref, 0)
    md.forces(want=False)
    ts_r = orc.step(ref, nsteps, 0)
    ts_g = md.step(nsteps, 0)
    assert ts_r == ts_g == nsteps
    got = col.copy()
    md.download(got)
    md.close()
    return ref, got


def _max_disp_err(a, b, box):
    assert a.ncells() == b.ncells()
    pa = a.flat("pos").reshape(-1, a.dim)
    pb = b.flat("pos").reshape(-1, b.dim)
    return float(np.max(np.linalg.norm(_minimg(pa - pb, box, a.dim), axis=1)))


@pytest.mark.parametrize("noise", [abi.NOISE_NONE, abi.NOISE_FDT])
def test_trajectory_plain_2d_1000_steps(gpu, noise):
    say("trajectory noise", noise)
    p, box, col = configs.plain_2d_sc(n_cells=200, L=20.0, seed=4, noise=noise)
    ref, got = _trajectory(p, box, col, 1000)
    err = _max_disp_err(ref, got, box)
    say("max displacement error", err)
    assert err <= 1e-9


def test_trajectory_dsfmt_replay(gpu):
    # reference noise: one Wiener vector per cell drawn from the reference's dSFMT stream
    say("trajectory dsfmt replay")
    p, box, col = configs.plain_2d_sc(n_cells=150, L=18.0, seed=5, noise=abi.NOISE_CELL_WIENER)
    p.species[0].d_noise = math.sqrt(2 * 0.01) / math.sqrt(p.dt)
    p.rng = abi.RNG_EXTERNAL
    nsteps = 1000
    per = 150 * 2
    z = O.dsfmt_fill(1, 1, per * (nsteps + 1)).reshape(nsteps + 1, per)
    ref, got = _trajectory(p, box, col, nsteps, ext=z)
    err = _max_disp_err(ref, got, box)
    say("max displacement error", err)
    assert err <= 1e-9


@pytest.mark.parametrize("quad_drag,kappa_push,kappa,e0", [(0, 5.0, 1.0, 1.0), (1, 0.2, 0.1, 0.1)])
def test_trajectory_reference_fene_cells(gpu, quad_drag, kappa_push, kappa, e0):
    # reference two-disc cells: FENE + push spring in r^2, rss motility, quadratic drag
    # (explicit Euler with the quadratic drag is only stable for weak forces)
    say("trajectory fene quad_drag", quad_drag)
    p, box, col = configs.plain_2d_sc(n_cells=150, L=18.0, seed=6, noise=abi.NOISE_NONE)
    sp = p.species[0]
    sp.spring = abi.SPRING_FENE
    sp.kappa = kappa
    sp.kappa_push = kappa_push
    sp.spring_ratio = 0.3
    sp.r2max = 2.0
    sp.rss = 0.5
    sp.quad_drag = quad_drag
    sp.motility[0] = 0.05
    sp.eps_core[0] = sp.eps_core[1] = e0
    ref, got = _trajectory(p, box, col, 500)
    err = _max_disp_err(ref, got, box)
    say("max displacement error", err)
    assert err <= 1e-9


def test_theta_cycle_divisions_bit_exact(gpu):
    say("theta divisions")
    p, box, col = configs.theta_2d_hz(n_cells=12, L=30.0, seed=3)
    ref, got = _trajectory(p, box, col, 200)
    assert ref.ncells() == got.ncells()
    assert ref.ncells()[0] > 12, "no division happened"
    a, b = ref.species[0], got.species[0]
    n = a["ncells"]
    for f in ("state", "t0", "tswitch", "tpressure"):
        assert np.array_equal(a[f][:n], b[f][:n]), f
    for f in ("theta", "rho", "ang_vel", "sigma_t"):
        assert np.allclose(a[f][:n], b[f][:n], rtol=1e-11, atol=1e-12), f
    assert _max_disp_err(ref, got, box) <= 1e-9


def test_theta_cycle_non_dividing_species(gpu):
    """Two species in the theta cell cycle, the second one non-dividing (CellParams::DIV false,
    divisionFunc NULL): only species 0 runs the cycle and divides (CIL.cu:496); species 1 keeps
    its count and its cycle variables, its pressure timer still counts (CIL.cu:492-493).
    Events are bit-exact against the oracle."""
    say("theta divisions, mixed dividing / non-dividing species")
    p, box, c1 = configs.theta_2d_hz(n_cells=24, L=30.0, seed=5)
    p.nspecies = 2
    abi_sp = p.species
    ctypes_copy = abi.Species.from_buffer_copy(abi_sp[0])
    abi_sp[1] = ctypes_copy
    abi_sp[1].no_division = 1
    configs.set_law(p, abi.LAW_HZ, 1.0)
    col = Colony(2, 1, [256, 256], ncells=[12, 12])
    for s_ in range(2):
        for f in ("pos", "sigma", "sigma_t", "state", "rho", "theta", "ang_vel"):
            col.species[s_][f][:12] = c1.species[0][f][12 * s_:12 * (s_ + 1)]
    th1 = col.species[1]["theta"][:12].copy()
    ref, got = _trajectory(p, box, col, 200)
    assert ref.ncells() == got.ncells()
    assert ref.ncells()[0] > 12 and ref.ncells()[1] == 12, ref.ncells()
    assert np.array_equal(got.species[1]["theta"][:12], th1)
    for s_ in range(2):
        a, b = ref.species[s_], got.species[s_]
        n = a["ncells"]
        for f in ("state", "t0", "tswitch", "tpressure"):
            assert np.array_equal(a[f][:n], b[f][:n]), (s_, f)
        for f in ("theta", "rho", "ang_vel", "sigma_t"):
            assert np.allclose(a[f][:n], b[f][:n], rtol=1e-11, atol=1e-12), (s_, f)
    assert _max_disp_err(ref, got, box) <= 1e-9


def test_linear_growth_divisions_and_drain(gpu):
    say("linear growth + drain")
    p, box, col = configs.channel_2d_hz(width=12.0, length=30.0, capacity=512, dt=1e-3, tau=0.05)
    ref, got = _trajectory(p, box, col, 300)
    assert ref.ncells() == got.ncells()
    assert ref.ncells()[0] > 8
    n = ref.ncells()[0]
    assert np.array_equal(ref.species[0]["t0"][:n], got.species[0]["t0"][:n])
    assert np.allclose(ref.species[0]["rest_len"][:n], got.species[0]["rest_len"][:n], atol=1e-12)
    assert _max_disp_err(ref, got, box) <= 1e-9


def test_division_beyond_capacity_is_reported(gpu):
    """A full species region: a division that would overflow it is not applied, the step
    returns a capacity error with a message, and the cell count stays at the capacity with
    finite positions."""
    say("division beyond capacity")
    p, box, col = configs.channel_2d_hz(width=12.0, length=30.0, capacity=4, dt=1e-3, tau=0.05)
    md = HipMD(p, box, [4])
    md.upload(col)
    md.forces(want=False)
    ts = 0
    with pytest.raises(abi.CilError) as ei:
        for _ in range(40):
            ts = md.step(25, ts)
    say("capacity error:", ei.value)
    assert "capacity" in str(ei.value)
    got = col.copy()
    md.download(got)
    md.close()
    assert got.ncells() == [4]
    assert np.all(np.isfinite(got.flat("pos")))


def test_drain_removal_order(gpu):
    say("drain removal")
    p, box, col = configs.channel_2d_hz(width=12.0, length=30.0, capacity=64, dt=1e-3)
    p.division = abi.DIV_NONE
    rng = np.random.default_rng(9)
    s = col.species[0]
    n = 40
    s["ncells"] = n
    c = np.stack([rng.uniform(1, 11, n), rng.uniform(-3, 33, n)], axis=1)
    s["pos"][:n, 0] = c
    s["pos"][:n, 1] = c + [0.0, 0.5]
    s["sigma"][:n] = 1.0
    s["rest_len"][:n] = 0.5
    s["tswitch"][:n] = np.arange(n)  # tag cells to follow the moves
    orc = O.Oracle(p, box, traversal=1)
    ref = col.copy()
    orc.forces(ref, 0)
    removed_ref = orc.check_positions(ref)
    md = HipMD(p, box, [64])
    md.upload(col)
    md.forces(want=False)
    removed = md.check_positions()
    got = col.copy()
    md.download(got)
    md.close()
    assert removed == removed_ref > 0
    m = ref.species[0]["ncells"]
    assert got.species[0]["ncells"] == m
    assert np.array_equal(got.species[0]["tswitch"][:m], ref.species[0]["tswitch"][:m])


def test_bulk3d_fp32_against_fp64(gpu):
    say("bulk 3d fp32 vs fp64")
    p, box, col = configs.bulk_3d_mixed(n=8192, seed=2, precision=abi.PREC_FP64)
    # push-off relaxation on the GPU, then compare force modes on the relaxed state
    q = configs.relaxation_params(p)
    md = HipMD(q, box, [sp["capacity"] for sp in col.species])
    md.upload(col)
    md.forces(want=False)
    md.step(100, 0)
    md.download(col)
    md.close()
    worst, *_ = _force_parity(p, box, col, 1e-12)
    assert worst <= 1.0
    p.precision = abi.PREC_FP32
    worst32, frc, fref, _, _ = _force_parity(p, box, col, 1e-5)
    orc = O.Oracle(p, box, traversal=1)
    scale = orc.force_scale(col)
    err = np.linalg.norm(frc - fref, axis=1)
    rel = err / (scale + np.linalg.norm(fref, axis=1))
    k = int(np.argmax(rel))
    say("fp32 worst", worst32, "at particle", k, "S_p", scale[k], "|F|", np.linalg.norm(fref[k]), "err", err[k],
        "median S", float(np.median(scale)), "99.9% rel", float(np.quantile(rel, 0.999)),
        "max err/median S", float(err.max() / np.median(scale)))
    assert worst32 <= 1.0


def test_capped_relaxation_trajectory(gpu):
    # relaxation of random positions: production laws, displacement capped per step
    say("capped relaxation")
    p, box, col = configs.bulk_3d_mixed(n=3000, seed=8, precision=abi.PREC_FP64)
    q = configs.relaxation_params(p)
    ref, got = _trajectory(q, box, col, 30)
    err = _max_disp_err(ref, got, box)
    say("max displacement error", err)
    assert err <= 1e-9
