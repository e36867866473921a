"""The reference's JSON configurations read by the library (cil_hip_config_read / _init), and the
reference's random stream as the library draws it (CIL_RNG_DSFMT).

Pins, each independent of the code under test:
  - the library's dSFMT-19937 against the raw output words of the reference's own mt/dSFMT.c
    (tests/golden/dsfmt_ref.npz) and its Leva normals against the oracle's restatement;
  - the parameters read from the seeded run's RunParams file against the values written in the
    file (extracted independently into tests/golden/prx_seed1_frames.json), with the
    reference's derivations: cycle time constants = CYCLE_TAU_* x TAU_DIVISION.AVG / dt,
    v_s = area at CYCLE_SIGMA_MIN (Input.cpp:160-199, CIL.cuh:218-225);
  - the initial colony from the file alone: initializeCellStatus on the stream gives the
    reference's own switch times (InitConf.dat: 47049 for seed 1 / dt 1e-5, 139509 for seed 408
    / dt 2.5e-6) and the run's t_0 phase;
  - the seeded run (output/z0.001t3.16_v2.h5, frames 1-6, fixture tests/golden/prx_seed1_frames.json)
    reproduced from its configuration alone: by the oracle on the CPU and by the GPU step on the
    device's copy of the stream.  Frame 1 (one cell) to the file's float32 precision; the first
    division at the file's step with the file's cell count; frames 2-6 theta and rho within 5e-6,
    positions within 5e-3 sigma (a -ffast-math build, and a cell direction that only follows
    its own velocity, so direction errors are never damped).
"""
import json
import math
import os

import numpy as np
import pytest

from cil_cuda_amd import abi, config
from oracle import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
CFG = os.path.join(HERE, "golden", "configs")
SEEDED = os.path.join(CFG, "z0.001t3.16_v2.RunParams.json")
FRAMES = os.path.join(HERE, "golden", "prx_seed1_frames.json")


@pytest.mark.parametrize("seed", [1, 408, 4357, 0xDEADBEEF])
def test_library_dsfmt_matches_reference_words(seed):
    ref = np.load(os.path.join(HERE, "golden", "dsfmt_ref.npz"))[f"seed_{seed}"]
    mine = config.dsfmt_fill(seed, 0, len(ref)) + 1.0  # close_open + 1 == close1_open2 exactly
    assert np.array_equal(mine.view(np.uint64), ref)
    # the skip: continuing after k outputs gives the stream's outputs k, k+1, ...
    tail = config.dsfmt_fill(seed, 0, 100, skip=1000) + 1.0
    assert np.array_equal(tail.view(np.uint64), ref[1000:1100])


def test_library_normals_and_open_draws_match_oracle():
    for kind in (1, 2):
        assert np.array_equal(config.dsfmt_fill(11, kind, 5000), O.dsfmt_fill(11, kind, 5000))


def test_all_shipped_configurations_read():
    ok = {}
    for f in sorted(os.listdir(CFG)):
        try:
            p, box, run = config.read_config(os.path.join(CFG, f))
            ok[f] = (p.dim, run.gts)
        except config.ConfigError as e:
            ok[f] = str(e)
    # the reference's own reader rejects the two files that lack P_FRICTION (JsonParser throws)
    assert "P_FRICTION" in ok["test.3d.json"] and "P_FRICTION" in ok["dev2d.json"]
    for f in ("PRX_MDCK_2D.json", "Match_Puliafito_2D.json", "Ecoli_channel_2D.json", "cil.RunParams.json",
              "z0.001t3.16_v2.RunParams.json"):
        assert ok[f][0] == 2, (f, ok[f])


def test_seeded_run_parameters():
    g = json.load(open(FRAMES))["params"]
    p, box, run = config.read_config(SEEDED)
    sp = p.species[0]
    dt = g["dt"]
    assert p.dt == dt and run.gts == g["gts"] and run.seed == g["seed"]
    assert list(box.length)[:2] == g["box"] and list(box.periodic)[:2] == [1, 1]
    assert sp.zeta[0] == g["friction"] and sp.motility[0] == g["motility"]
    tau = g["tau_division"][0]
    assert sp.cycle_tau_r == pytest.approx(g["cycle_tau_r"] * tau / dt, rel=1e-15)
    assert sp.cycle_tau_v == pytest.approx(g["cycle_tau_v"] * tau / dt, rel=1e-15)
    assert sp.cycle_v_s == pytest.approx(math.pi * (g["cycle_sigma_min"] / 2) ** 2, rel=1e-15)
    assert sp.cycle_p0 == g["cycle_p0"] and sp.cycle_v_1 == g["v1"] and sp.cycle_rho_ratio == g["cycle_rho_ratio"]
    assert (sp.division_tau.mu, sp.division_tau.sigma) == tuple(g["tau_division"])
    assert p.law[0][0] == abi.LAW_HZ and p.division == abi.DIV_THETA_CYCLE and p.noise == abi.NOISE_NONE
    # HZ block: modulus 100, Poisson 0.5, adhesion 0.2, diameter CYCLE_SIGMA_MIN
    assert (sp.eps_core[0], sp.sig_well[0], sp.eps_well[0]) == (100.0, 0.5, 0.2)
    assert sp.sigma[0] == g["cycle_sigma_min"]


@pytest.mark.parametrize("name,expected", [("z0.001t3.16_v2.RunParams.json", 47049), ("cil.RunParams.json", 139509)])
def test_initial_switch_time_known_answers(name, expected):
    p, box, run = config.read_config(os.path.join(CFG, name))
    col = config.initial_colony(os.path.join(CFG, name), p)
    s = col.species[0]
    assert col.ncells() == [1]
    assert s["tswitch"][0] == expected and s["state"][0] == abi.CS_DIVIDING
    assert p.rng == abi.RNG_DSFMT and p.rng_offset == 5  # uniform, normal (2 words), uniform, uniform
    if name.startswith("z0.001"):
        t0 = json.load(open(FRAMES))["frames"]["t_0"]
        assert abs(s["theta"][0] - t0["Theta"][0]) <= 1e-7
        assert list(s["pos"][0, 0]) == t0["R"]


def _frames_from_config(run_gpu, nframes=6):
    p, box, run = config.read_config(SEEDED)
    col = config.initial_colony(SEEDED, p, [64])
    if run_gpu:
        from cil_cuda_amd.colony import HipMD

        md = HipMD(p, box, [64])
        md.upload(col)
        md.forces(want=False)
        step = lambda ts: md.step(run.gts, ts)  # noqa: E731
        snap = lambda: md.download(col.copy())  # noqa: E731
    else:
        orc = O.Oracle(p, box, traversal=1)
        orc.forces(col, 0)
        step = lambda ts: orc.step(col, run.gts, ts)  # noqa: E731
        snap = lambda: col.copy()  # noqa: E731
    out, ts = [], 0
    for _ in range(nframes):
        ts = step(ts)
        out.append(snap())
    return out


def _check_frames(out):
    g = json.load(open(FRAMES))
    for f, c in enumerate(out, start=1):
        ref = g["frames"][f"t_{f}"]
        s = c.species[0]
        n = s["ncells"]
        assert n == len(ref["Theta"]), (f, n)
        assert np.array_equal(s["t0"][:n], np.array(ref["T0"], dtype=int))
        d_pos = float(np.max(np.abs(s["pos"][:n, 0].reshape(-1) - np.array(ref["R"]))))
        d_theta = float(np.max(np.abs(s["theta"][:n] / np.array(ref["Theta"]) - 1)))
        d_sigma = float(np.max(np.abs(s["sigma"][:n, 0] / np.array(ref["SigmaCore"]) - 1)))
        d_rho = float(np.max(np.abs(s["rho"][:n] - np.array(ref["RHO"]))))
        print(f"frame {f}: {n} cells, |dR| {d_pos:.2e}, theta {d_theta:.2e}, sigma {d_sigma:.2e}, rho {d_rho:.2e}",
              flush=True)
        if f == 1:
            assert d_pos <= 1e-5 and d_theta <= 2e-6 and d_sigma <= 2e-6 and d_rho <= 1e-6
        else:
            assert d_pos <= 5e-3 and d_theta <= 5e-6 and d_sigma <= 5e-6 and d_rho <= 5e-6


def test_seeded_run_from_its_configuration_oracle():
    _check_frames(_frames_from_config(False))


@pytest.mark.gpu
def test_seeded_run_from_its_configuration_gpu(gpu):
    """The GPU step with the reference's stream on the device: frames 1-6 of the file, and the
    oracle's frames within 1e-9 sigma (the same draws in the same order)."""
    got = _frames_from_config(True)
    _check_frames(got)
    ref = _frames_from_config(False)
    for a, b in zip(ref, got):
        assert a.ncells() == b.ncells()
        n = a.ncells()[0]
        assert np.array_equal(a.species[0]["tswitch"][:n], b.species[0]["tswitch"][:n])
        assert float(np.max(np.abs(a.species[0]["pos"][:n] - b.species[0]["pos"][:n]))) <= 1e-9
