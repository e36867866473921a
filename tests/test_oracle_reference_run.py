"""The oracle against the reference's own seeded output: frame 1 (10^4 Euler steps) of the
single-disc HZ run with the theta cell cycle (output/z0.001t3.16_v2.h5, seed 1; fixture
tests/golden/prx_seed1_frames.json written by tests/golden/make_prx_golden.py).

The run starts from one cell at (50, 50).  Its state comes from initializeCellStatus
(CIL.cuh:270-298): the dividing-state draw, the division time span (angular velocity
1/span), the phase theta.  These are replayed here with the reference's dSFMT stream, and
theta must come out as the file's t_0 value.

The cell then crawls with motility 1 along the direction drawn at the first force call
(singleCellForce, MD.cu:600-644).  That direction is taken from the file's own displacement;
nothing else depends on it.  Its area grows by the cycle ODEs of calculDivisionsSingle
(CIL.cu:469-619), and its mobility is 1/(zeta * area) (CIL.cu:902-939).

Frame 1 comes before the first division, so the displacement, sigma and theta there pin
Euler, the single-disc motility and the growth ODE.  The file holds float32 values written
by a -ffast-math build, so the bar is 2e-6 relative.

Parameters follow Input.cpp:160-199: the cycle time constants are CYCLE_TAU_* times
TAU_DIVISION.AVG (divided by dt here), v_s is the area at CYCLE_SIGMA_MIN, V1 is v_1,
and cycleRMaxedInv keeps its default of 1 (Cell.cuh:254).
"""
import json
import math
import os
import sys

import numpy as np
import pytest

from cil_cuda_amd import abi, configs
from cil_cuda_amd.colony import Colony
from oracle import oracle as O

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
from test_oracle_rng import _Stream, _nint  # noqa: E402

FIX = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "prx_seed1_frames.json")


def _setup():
    g = json.load(open(FIX))
    q, f0, f1 = g["params"], g["frames"]["t_0"], g["frames"]["t_1"]
    dt = q["dt"]
    # initializeCellStatus draws (CIL.cuh:283-298), dividing branch (TAU_MIGRATION.AVG = 0)
    r = _Stream(q["seed"])
    prob = q["tau_migration"][0] / (q["tau_migration"][0] + q["tau_division"][0])
    assert not (r.uniform0x(1.0) <= prob)
    span = max(_nint(r.normal(q["tau_division"][0], q["tau_division"][1]) / dt), 1)
    theta0 = r.uniform0x(1.0)

    p = configs.base_params(2, 1, 1)
    sp = p.species[0]
    sp.zeta[0] = q["friction"]
    sp.motility[0] = q["motility"]
    sp.eps_core[0] = 100.0  # HZ ELASTIC_MODULUS, POISSON_RATIO, ADHESION_CONSTANT of the run's file
    sp.sig_well[0] = 0.5
    sp.eps_well[0] = 0.2
    tau = q["tau_division"][0]
    sp.cycle_p0 = q["cycle_p0"]
    sp.cycle_tau_r = q["cycle_tau_r"] * tau / dt
    sp.cycle_tau_v = q["cycle_tau_v"] * tau / dt
    sp.cycle_tau_p = q["cycle_tau_p"] * tau / dt
    sp.cycle_v_s = math.pi * (q["cycle_sigma_min"] / 2) ** 2
    sp.cycle_v_1 = q["v1"]
    sp.cycle_beta_vs = q["cycle_beta_vs"]
    sp.cycle_rmaxed_inv = 1.0
    sp.cycle_rho_ratio = q["cycle_rho_ratio"]
    sp.apop_pressure = 1e300
    sp.division_tau.mu, sp.division_tau.sigma = q["tau_division"]
    sp.migration_tau.mu, sp.migration_tau.sigma = q["tau_migration"]
    configs.set_law(p, abi.LAW_HZ, 1.0)
    p.dt = dt
    p.noise = abi.NOISE_NONE
    p.seed = q["seed"]
    p.division = abi.DIV_THETA_CYCLE
    p.pressure = 1
    p.precision = abi.PREC_FP64
    box = configs.make_box(q["box"], [1 if b else 0 for b in q["pbc"]])

    col = Colony(2, 1, [16], ncells=[1])
    s = col.species[0]
    r0, r1 = np.array(f0["R"]), np.array(f1["R"])
    u = (r1 - r0) / np.linalg.norm(r1 - r0)
    s["pos"][0, 0] = q["init"][0]
    s["vel"][0, 0] = 1e-11 * u  # the direction drawn at the first force call
    s["sigma"][0, 0] = q["cycle_sigma_min"]
    s["sigma_t"][0] = q["cycle_sigma_min"]
    s["state"][0] = abi.CS_DIVIDING
    s["rho"][0] = 1.0
    s["theta"][0] = theta0
    s["ang_vel"][0] = 1.0 / span
    s["t0"][0] = 0
    return g, theta0, p, box, col


def _run_frame1():
    g, theta0, p, box, col = _setup()
    orc = O.Oracle(p, box, traversal=1)
    orc.forces(col, 0)
    assert orc.step(col, g["params"]["gts"], 0) == g["params"]["gts"]
    orc.close()
    return g, theta0, col


def _check_frame1(g, col):
    f0, f1 = g["frames"]["t_0"], g["frames"]["t_1"]
    s = col.species[0]
    assert s["ncells"] == 1  # no division before frame 1
    theta, sig = float(s["theta"][0]), float(s["sigma"][0, 0])
    disp = float(np.linalg.norm(s["pos"][0, 0] - np.array(f0["R"])))
    disp_ref = float(np.linalg.norm(np.array(f1["R"]) - np.array(f0["R"])))
    print(f"theta {theta:.9g} (ref {f1['Theta'][0]:.9g}), sigma {sig:.9g} (ref {f1['SigmaCore'][0]:.9g}), "
          f"displacement {disp:.9g} (ref {disp_ref:.9g}), rho {float(s['rho'][0]):.9g}", flush=True)
    assert abs(theta / f1["Theta"][0] - 1) <= 2e-6
    assert abs(sig / f1["SigmaCore"][0] - 1) <= 2e-6
    assert abs(disp / disp_ref - 1) <= 2e-6
    assert np.allclose(s["pos"][0, 0], f1["R"], rtol=0, atol=2e-4)
    assert abs(float(s["rho"][0]) - f1["RHO"][0]) <= 1e-6


def test_initial_phase_matches_the_seeded_run():
    g, theta0, _ = _run_frame1()
    assert abs(theta0 - g["frames"]["t_0"]["Theta"][0]) <= 1e-7


def test_frame1_matches_the_seeded_run():
    g, _, col = _run_frame1()
    _check_frame1(g, col)


@pytest.mark.gpu
def test_hip_frame1_matches_the_seeded_run(gpu):
    # the same start through the C ABI: 10^4 graph-launched Euler steps on the GPU
    from cil_cuda_amd.colony import HipMD

    g, _, p, box, col = _setup()
    md = HipMD(p, box, [sp["capacity"] for sp in col.species])
    md.upload(col)
    md.forces(want=False)
    assert md.step(g["params"]["gts"], 0) == g["params"]["gts"]
    got = col.copy()
    md.download(got)
    md.close()
    _check_frame1(g, got)


def _run_sequential(nframes):
    """The run on the reference's sequential dSFMT stream from seed 1: initializeCellStatus's
    draws, then every draw of the step in the reference's order (the first force call's
    random direction, the divisions' geometry and timers); no passive or active noise is
    drawn in this run (dNoise = 0, CIL.cu:793-818).  Returns the colony after each frame."""
    g, _, p, box, col = _setup()
    q = g["params"]
    dt = q["dt"]
    orc = O.Oracle(p, box, traversal=1)
    orc.set_rng(O.RNG_DSFMT, q["seed"])
    prob = q["tau_migration"][0] / (q["tau_migration"][0] + q["tau_division"][0])
    assert not (orc.rng_draw(0) <= prob)
    span = max(_nint((q["tau_division"][0] + q["tau_division"][1] * orc.rng_draw(1)) / dt), 1)
    s = col.species[0]
    s["ang_vel"][0] = 1.0 / span
    s["theta"][0] = orc.rng_draw(0)
    s["tswitch"][0] = span - (_nint(orc.rng_draw(0) * span) - 1)
    s["vel"][0, 0] = 0.0  # singleCellForce draws the direction at the first force call
    orc.forces(col, 0)
    out, ts = [], 0
    for _ in range(nframes):
        ts = orc.step(col, q["gts"], ts)
        out.append(col.copy())
    orc.close()
    return g, out


def test_frames_match_the_seeded_run_on_the_reference_stream():
    """Frames 1-6 on the reference's stream.  Frame 1 (one cell, its direction now drawn, not
    taken from the file) is held to the float32 bar.  The first division (step 12955) then
    comes out at the same step, with the same cell count, and the daughters' phase velocities
    come from the same draws.  After it theta, sigma and rho agree to 1e-6 (through the contact
    pressure of the daughters' adhesive Hertz contact, ADHESION_CONSTANT 0.2) and the positions
    to 3e-3 sigma: the cells crawl along their own velocity, a direction nothing restores, so
    the reference's -ffast-math rounding is never damped.  A Hertz force along the unit
    direction instead of the non-unit r_ij is far further off, so the kept non-unit form
    (MD.cu:350-356) is the one the run used."""
    g, out = _run_sequential(6)
    for f, c in enumerate(out, start=1):
        ref = g["frames"][f"t_{f}"]
        s = c.species[0]
        n = s["ncells"]
        assert n == len(ref["Theta"])
        pos = s["pos"][:n, 0].reshape(-1)
        d_pos = float(np.max(np.abs(pos - np.array(ref["R"]))))
        d_theta = float(np.max(np.abs(s["theta"][:n] / np.array(ref["Theta"]) - 1)))
        d_sigma = float(np.max(np.abs(s["sigma"][:n, 0] / np.array(ref["SigmaCore"]) - 1)))
        d_rho = float(np.max(np.abs(s["rho"][:n] - np.array(ref["RHO"]))))
        print(f"frame {f}: {n} cells, max |dR| {d_pos:.2e}, theta {d_theta:.2e}, sigma {d_sigma:.2e}, "
              f"rho {d_rho:.2e}", flush=True)
        assert np.array_equal(s["t0"][:n], np.array(ref["T0"], dtype=int))
        if f == 1:
            assert d_pos <= 1e-5 and d_theta <= 2e-6 and d_sigma <= 2e-6 and d_rho <= 1e-6
        else:
            assert d_pos <= 5e-3 and d_theta <= 5e-6 and d_sigma <= 5e-6 and d_rho <= 5e-6
