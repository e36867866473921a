"""Write z0.001t3.16_v2_frames.npz: the per-cell arrays the statistics compare (Theta, RHO, T0,
SigmaCore) of all 101 snapshots (t = 0, 0.1, ..., 10) of the reference's seeded run file
output/z0.001t3.16_v2.h5 (the run of output/RunParams/z0.001t3.16_v2.RunParams.json), read
with the installed `h5dump` tool.  Arrays are concatenated over snapshots; `offsets[i]` is
where snapshot i starts.

Run in a container that holds the reference tree; the file it writes is the committed fixture.
"""
import os
import re
import subprocess

import numpy as np

REF = "/root/reference"
H5 = os.path.join(REF, "output", "z0.001t3.16_v2.h5")
H5DUMP = "/opt/conda/bin/h5dump"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "z0.001t3.16_v2_frames.npz")


def dataset(path):
    txt = subprocess.run([H5DUMP, "-m", "%.9g", "-d", path, H5], capture_output=True, text=True, check=True).stdout
    body = txt.split("DATA {", 1)[1].rsplit("}", 2)[0]
    vals = []
    for line in body.splitlines():
        line = line.split("):", 1)[-1] if "):" in line else line
        vals += [float(x) for x in re.findall(r"[-+0-9.eE]+", line)]
    return vals


def main():
    cols = {"Theta": [], "RHO": [], "T0": [], "SigmaCore": []}
    offsets = [0]
    for t in range(101):
        n = None
        for k in cols:
            v = dataset(f"/snapshots/t_{t}/{k}")
            n = len(v) if n is None else n
            assert len(v) == n
            cols[k] += v
        offsets.append(offsets[-1] + n)
    np.savez_compressed(OUT, offsets=np.array(offsets), theta=np.array(cols["Theta"], np.float32),
                        rho=np.array(cols["RHO"], np.float32), t0=np.array(cols["T0"], np.int64),
                        sigma=np.array(cols["SigmaCore"], np.float32))


if __name__ == "__main__":
    main()
