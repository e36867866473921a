"""The observables of the reference's analysis tool (cilJintaoTools.analyse_folder,
cilJintaoTools.py:76-252) restated in numpy for snapshot series of single-disc colonies.

A snapshot is (theta, rho, t0, sigma) per cell; snapshots are `gts` steps of `dt` apart.
"""
import math

import numpy as np

AREA_EDGES = np.arange(0.5, 16, 0.5)       # cilJintaoTools.py:191
DIVISION_EDGES = np.arange(0, 10, 0.125)   # cilJintaoTools.py:246


def area_hist(sigma, edges=AREA_EDGES):
    """Density histogram of the cell areas pi (sigma / 2)^2 (cilJintaoTools.py:112, 191-199)."""
    h, _ = np.histogram(math.pi * (np.asarray(sigma) / 2) ** 2, bins=edges, density=True)
    return h


def division_times(frames, dt, gts, tau_div):
    """cilJintaoTools.py:222-243: for each snapshot from the tenth on, the divisions since the
    previous one (the cell count increase) are the cells whose theta fell; each one's age at
    division is its time since birth (T0) at the previous snapshot plus an Euler step of theta
    to 1, at most 0.1; in units of tau_div.  Returns (times, colony age at division)."""
    out, age = [], []
    for idx in range(10, len(frames)):
        th1, rho1, t01, _ = frames[idx - 1]
        th2 = frames[idx][0]
        nd = len(th2) - len(th1)
        if nd <= 0:
            continue
        ev = th2[: len(th1)] < th1
        ext = np.minimum((1.0 - th1[ev]) / rho1[ev], 0.1)
        out.extend((ext + ((idx - 1) * gts - t01[ev]) * dt / tau_div).tolist())
        age.extend((ext + (idx - 1) * gts * dt).tolist())
    return np.array(out), np.array(age)


def load_frames(path):
    """Snapshot series of tests/golden/make_run_golden.py."""
    d = np.load(path, allow_pickle=False)
    o = d["offsets"]
    return [(d["theta"][o[i]:o[i + 1]].astype(np.float64), d["rho"][o[i]:o[i + 1]].astype(np.float64),
             d["t0"][o[i]:o[i + 1]], d["sigma"][o[i]:o[i + 1]].astype(np.float64)) for i in range(len(o) - 1)]
