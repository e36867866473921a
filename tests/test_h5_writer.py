"""The library's HDF5 writer (cil_hip_h5_*) against the reference's own HDF5 output, compared as
the installed `h5dump` prints them (types, shapes, attributes, values):

  - the reference's `z0.001t3.16_v2.Last.h5` (tests/golden; written by outputHDF5Frame at the
    end of the seeded run): the colony it holds, written again by the library as a bare file's
    root frame, prints identically;
  - the /params group of the reference's seeded run file (tests/golden/z0.001t3.16_v2.params.txt,
    from output/z0.001t3.16_v2.h5 by tests/golden/make_h5_golden.py): the library's file
    created from that run's configuration prints identically;
  - snapshot groups /snapshots/t_<id> carry the attribute ts and the species' arrays, one
    subgroup per species when there are several.
"""
import os
import re
import subprocess

import numpy as np
import pytest

from cil_cuda_amd import abi, config, configs
from cil_cuda_amd.colony import Colony
from cil_cuda_amd.output import H5Writer

HERE = os.path.dirname(os.path.abspath(__file__))
LAST = os.path.join(HERE, "golden", "z0.001t3.16_v2.Last.h5")
PARAMS = os.path.join(HERE, "golden", "z0.001t3.16_v2.params.txt")
SEEDED = os.path.join(HERE, "golden", "configs", "z0.001t3.16_v2.RunParams.json")
H5DUMP = "/opt/conda/bin/h5dump"

pytestmark = pytest.mark.skipif(not os.path.exists(H5DUMP), reason="h5dump is not installed")


def dump(path, *args):
    """h5dump's text without its first line (the file name)."""
    txt = subprocess.run([H5DUMP, *args, str(path)], capture_output=True, text=True, check=True).stdout
    return txt.split("\n", 1)[1]


def dataset(path, name):
    txt = subprocess.run([H5DUMP, "-m", "%.9g", "-d", name, str(path)], capture_output=True, text=True,
                         check=True).stdout
    body = txt.split("DATA {", 1)[1].rsplit("}", 2)[0]
    vals = []
    for line in body.splitlines():
        line = line.split("):", 1)[-1] if "):" in line else line
        vals += [float(x) for x in re.findall(r"[-+0-9.eE]+", line)]
    return np.array(vals)


def test_reference_last_file_written_again(tmp_path):
    r = dataset(LAST, "/R").reshape(-1, 2)
    n = r.shape[0]
    col = Colony(2, 1, [n], ncells=[n])
    s = col.species[0]
    s["pos"][:, 0] = r
    s["vel"][:, 0] = dataset(LAST, "/Vel").reshape(-1, 2)
    s["frc"][:, 0] = dataset(LAST, "/F").reshape(-1, 2)
    s["sigma"][:, 0] = dataset(LAST, "/SigmaCore")
    s["rho"][:] = dataset(LAST, "/RHO")
    s["theta"][:] = dataset(LAST, "/Theta")
    s["pressure"][:] = dataset(LAST, "/PR")
    s["state"][:] = dataset(LAST, "/State").astype(int)
    s["t0"][:] = dataset(LAST, "/T0").astype(int)
    cpu = float(dataset(LAST, "/CpuTime")[0])
    p, box, run = config.read_config(SEEDED)
    out = tmp_path / "last.h5"
    with H5Writer(out) as f:
        f.frame(p, col, -1, -1, run, cpu_time=cpu)
    assert dump(out) == dump(LAST)


def test_run_parameters_group(tmp_path):
    p, box, run = config.read_config(SEEDED)
    out = tmp_path / "run.h5"
    with H5Writer(out, p, box, run):
        pass
    assert dump(out, "-g", "/params") == open(PARAMS).read()


def test_snapshot_groups(tmp_path):
    p, box, col = configs.channel_2d_hz(width=12.0, length=30.0, seed=2, capacity=16)
    q = configs.base_params(2, 2, 2)
    c2 = Colony(2, 2, [8, 8], ncells=[3, 5])
    rng = np.random.default_rng(1)
    for sp in c2.species:
        n = sp["ncells"]
        sp["pos"][:n] = rng.uniform(0, 12, (n, 2, 2))
        sp["sigma"][:n] = rng.uniform(0.8, 1.2, (n, 2))
        sp["t0"][:n] = rng.integers(0, 100, n)
    out = tmp_path / "snap.h5"
    with H5Writer(out, q, configs.make_box([12.0, 30.0], [1, 1])) as f:
        f.frame(q, c2, 0, 0)
        f.frame(q, c2, 1, 250)
    hdr = dump(out, "-H")
    for g in ("/snapshots/t_0/C1", "/snapshots/t_1/C2"):
        assert f'GROUP "{g.rsplit("/", 1)[1]}"' in hdr
    txt = dump(out, "-a", "/snapshots/t_1/ts")
    assert "(0): 250" in txt
    r = dataset(out, "/snapshots/t_1/C2/R").reshape(5, 2, 2)
    assert np.array_equal(r.astype(np.float32), c2.species[1]["pos"][:5].astype(np.float32))
    t0 = dataset(out, "/snapshots/t_0/C1/T0")
    assert np.array_equal(t0, c2.species[0]["t0"][:3])


def test_unwritable_path_is_an_error(tmp_path):
    with pytest.raises(abi.CilError):
        H5Writer(tmp_path / "no_such_dir" / "x.h5")


@pytest.mark.gpu
def test_frame_from_device_equals_frame_from_host(gpu, tmp_path):
    """cil_hip_h5_write_frame_ctx (the device colony) prints as the host writer's frame of the
    downloaded colony, after the first division of a two-disc channel."""
    from cil_cuda_amd.colony import HipMD

    p, box, col = configs.channel_2d_hz(width=12.0, length=40.0, seed=3, dt=1e-3, capacity=64)
    md = HipMD(p, box, [64])
    md.upload(col)
    md.forces(want=False)
    ts = md.step(1700, 0)
    a, b = tmp_path / "dev.h5", tmp_path / "host.h5"
    with H5Writer(a, p, box) as f:
        f.frame_ctx(md, 0, ts)
    host = md.download(col.copy())
    md.close()
    assert host.ncells()[0] > 1
    with H5Writer(b, p, box) as f:
        f.frame(p, host, 0, ts)
    assert dump(a) == dump(b)
