"""Write quantities_ref.npz: the part of the reference's analysed seeded run (quantities.npz and
quantities_v3.2.npz, two builds of the reference, written by cilJintaoTools.analyse_folder) that
the statistics test compares, for t <= 10: cell counts, cell-area histograms and the division
times of divisions up to t = 10.

Run in a container that holds the reference tree; the file it writes is the committed fixture.
"""
import os

import numpy as np

REF = "/root/reference"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "quantities_ref.npz")
T_END = 10.0


def main():
    runs = [np.load(os.path.join(REF, f), allow_pickle=False) for f in ("quantities.npz", "quantities_v3.2.npz")]
    n = int(np.sum(runs[0]["t"] <= T_END + 1e-9))
    out = {
        "t": runs[0]["t"][:n],
        "cellNumber": np.stack([r["cellNumber"][:n] for r in runs]),
        "cellAreaBinEdges": runs[0]["cellAreaBinEdges"],
        "cellAreaDistributions": np.stack([np.asarray(r["cellAreaDistributions"])[:n] for r in runs]),
        "divisionBinEdges": runs[0]["divisionBinEdges"],
    }
    for k, r in enumerate(runs):
        keep = r["divisionColonyAge"] <= T_END
        out[f"divisionTimes_{k}"] = r["divisionTimes"][keep]
    np.savez_compressed(OUT, **out)


if __name__ == "__main__":
    main()
