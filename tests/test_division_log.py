"""Division records (the reference's posCellsDiv / cellT0Div behind the HDF5 /div_app group,
Colony.cpp:319-344; CIL.cu:396-399, 542-545): per division, the new cell's element positions
right after the division and the step.

CPU: the oracle's records, taken one step at a time, equal the appended cells' positions and
birth steps at the end of that step (the force pass moves nothing), for the theta cycle and
linear growth.  GPU: the device records equal the oracle's (steps exactly, positions within
1e-9 sigma in FP64), and the HDF5 frames of a run that keeps DIV_APP_RECORD hold them as
/div_app/DIV/d_<id>."""
import os
import re
import subprocess

import numpy as np
import pytest

from cil_cuda_amd import abi, config, configs
from oracle import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
H5DUMP = "/opt/conda/bin/h5dump"


def _builders():
    return [("theta", configs.theta_2d_hz(n_cells=12, seed=81), 80),
            ("growth", configs.channel_2d_hz(width=12.0, length=64.0, seed=82, kT=0.001, tau=0.0015, capacity=512), 240)]


@pytest.mark.parametrize("k", [0, 1])
def test_oracle_records_are_the_new_cells(k):
    name, (p, box, col), nsteps = _builders()[k]
    orc = O.Oracle(p, box, traversal=1)
    orc.set_division_log(True)
    c = col.copy()
    orc.forces(c, 0)
    ts, seen = 0, 0
    nE, D = col.nelem, col.dim
    while ts < nsteps:
        n0 = c.ncells()[0]
        ts = orc.step(c, 1, ts)
        n1 = c.ncells()[0]
        pos, steps = orc.division_log(0, D, nE)
        assert len(steps) == n1 - n0
        if n1 > n0:
            s = c.species[0]
            assert np.array_equal(pos, s["pos"][n0:n1])
            assert np.all(steps == ts) and np.all(s["t0"][n0:n1] == ts)
            seen += n1 - n0
    orc.close()
    assert seen > 0, name


def _h5_dataset(path, name):
    txt = subprocess.run([H5DUMP, "-m", "%.9g", "-d", name, str(path)], capture_output=True, text=True,
                         check=True).stdout
    body = txt.split("DATA {", 1)[1].rsplit("}", 2)[0] if "DATA {" in txt else ""
    vals = []
    for line in body.splitlines():
        line = line.split("):", 1)[-1] if "):" in line else line
        vals += [float(x) for x in re.findall(r"[-+0-9.eE]+", line)]
    return np.array(vals)


@pytest.mark.gpu
@pytest.mark.parametrize("k", [0, 1])
def test_device_records_match_the_oracle(gpu, k):
    from cil_cuda_amd.colony import HipMD

    name, (p, box, col), nsteps = _builders()[k]
    md = HipMD(p, box, [sp["capacity"] for sp in col.species])
    md.upload(col)
    md.set_division_log(True)
    md.forces(want=False)
    md.step(nsteps // 2, 0)
    pos_a, ts_a = md.division_log(0)
    md.step(nsteps - nsteps // 2, nsteps // 2)
    pos_b, ts_b = md.division_log(0)
    md.close()
    orc = O.Oracle(p, box, traversal=1)
    orc.set_division_log(True)
    c = col.copy()
    orc.forces(c, 0)
    orc.step(c, nsteps, 0)
    rpos, rts = orc.division_log(0, col.dim, col.nelem)
    orc.close()
    gpos, gts = np.concatenate([pos_a, pos_b]), np.concatenate([ts_a, ts_b])
    print(f"{name}: {len(gts)} divisions recorded ({len(ts_a)} before the read in the middle)", flush=True)
    assert len(gts) == len(rts) > 0 and np.array_equal(gts, rts)
    assert float(np.max(np.abs(gpos - rpos))) <= 1e-9


@pytest.mark.gpu
@pytest.mark.skipif(not os.path.exists(H5DUMP), reason="h5dump is not installed")
def test_hdf5_frames_hold_the_division_records(gpu, tmp_path):
    """The seeded run's configuration with DIV_APP_RECORD on, six frames on the GPU (the
    reference's stream on the device): each frame's /div_app/DIV/d_<id> holds the divisions of
    the oracle's run over that frame (T exactly, R to float precision)."""
    from cil_cuda_amd.colony import HipMD
    from cil_cuda_amd.output import H5Writer

    src = os.path.join(HERE, "golden", "configs", "z0.001t3.16_v2.RunParams.json")
    p, box, run = config.read_config(src)
    run.div_app_record = 1
    col = config.initial_colony(src, p, [64])
    md = HipMD(p, box, [64])
    md.upload(col)
    md.set_division_log(True)
    md.forces(want=False)
    orc = O.Oracle(p, box, traversal=1)
    orc.set_division_log(True)
    ref = col.copy()
    orc.forces(ref, 0)
    out = tmp_path / "run.h5"
    ts = tso = 0
    expect = []
    with H5Writer(out, p, box, run) as f:
        for fr in range(1, 7):
            ts = md.step(run.gts, ts)
            f.frame_ctx(md, fr, ts, run)
            tso = orc.step(ref, run.gts, tso)
            expect.append(orc.division_log(0, 2, 1))
    md.close()
    orc.close()
    total = 0
    for fr, (rpos, rts) in enumerate(expect, start=1):
        T = _h5_dataset(out, f"/div_app/DIV/d_{fr}/T")
        R = _h5_dataset(out, f"/div_app/DIV/d_{fr}/R").reshape(-1, 1, 2)
        print(f"frame {fr}: {len(rts)} divisions", flush=True)
        assert np.array_equal(T, rts)
        if len(rts):
            assert float(np.max(np.abs(R - rpos.astype(np.float32)))) <= 1e-4
        total += len(rts)
    assert total > 0
