"""Write prx_seed1_frames.json: the first seven snapshots of the reference's seeded single-disc
HZ run (output/z0.001t3.16_v2.h5, config output/RunParams/z0.001t3.16_v2.RunParams.json,
seed 1) and the run parameters the oracle check needs.

Run in a container that holds the reference tree (it reads /root/reference with the
installed `h5dump` tool); the JSON it writes is the committed fixture.
"""
import json
import os
import re
import subprocess

REF = "/root/reference"
H5 = os.path.join(REF, "output", "z0.001t3.16_v2.h5")
RUNPARAMS = os.path.join(REF, "output", "RunParams", "z0.001t3.16_v2.RunParams.json")
H5DUMP = "/opt/conda/bin/h5dump"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "prx_seed1_frames.json")


def dataset(path):
    txt = subprocess.run([H5DUMP, "-m", "%.9g", "-d", path, H5], capture_output=True, text=True, check=True).stdout
    body = txt.split("DATA {", 1)[1].rsplit("}", 2)[0]
    vals = []
    for line in body.splitlines():
        line = line.split("):", 1)[-1] if "):" in line else line
        vals += [float(x) for x in re.findall(r"[-+0-9.eE]+", line)]
    return vals


def main():
    frames = {}
    for t in range(7):
        frames[f"t_{t}"] = {k: dataset(f"/snapshots/t_{t}/{k}")
                            for k in ("R", "SigmaCore", "Theta", "RHO", "State", "T0", "PR")}
    rp = json.load(open(RUNPARAMS))
    md = rp["MD"]
    c1 = rp["SPECIES"]["C1"]
    div = c1["DIVISION"]
    params = {
        "seed": rp["RUNTIME"]["RNG"]["SEED"],
        "dt": md["DT"]["MANUAL"],
        "gts": md["GTS"],
        "box": md["BOX"],
        "pbc": md["PBC"],
        "friction": c1["FRICTION"][0],
        "motility": c1["MOTILITY"][0],
        "tau_division": [div["TAU_DIVISION"]["AVG"], div["TAU_DIVISION"]["SIGMA"]],
        "tau_migration": [div["TAU_MIGRATION"]["AVG"], div["TAU_MIGRATION"]["SIGMA"]],
        "cycle_tau_r": div["CYCLE_TAU_R"],
        "cycle_tau_v": div["CYCLE_TAU_V"],
        "cycle_tau_p": div["CYCLE_TAU_P"],
        "cycle_sigma_min": div["CYCLE_SIGMA_MIN"],
        "cycle_p0": div["CYCLE_P0"],
        "cycle_beta_vs": div["CYCLE_BETA_VS"],
        "cycle_rho_ratio": div["CYCLE_RHO_RATIO"],
        "v1": div["V1"],
        "init": rp["INIT"]["C1"]["JSON"],
    }
    json.dump({"source": "output/z0.001t3.16_v2.h5 (snapshots t_0 to t_6) and its RunParams.json",
               "params": params, "frames": frames}, open(OUT, "w"), indent=1)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
