"""The restated analysis (tests/colony_stats.py) on the reference's own seeded run file
(tests/golden/z0.001t3.16_v2_frames.npz): the properties its definitions fix."""
import os

import numpy as np

from colony_stats import AREA_EDGES, DIVISION_EDGES, area_hist, division_times, load_frames

HERE = os.path.dirname(os.path.abspath(__file__))


def test_reference_run_observables():
    frames = load_frames(os.path.join(HERE, "golden", "z0.001t3.16_v2_frames.npz"))
    counts = [len(f[0]) for f in frames]
    assert counts[0] == 1 and counts[100] == 934 and all(b >= a for a, b in zip(counts, counts[1:]))
    times, age = division_times(frames, 1e-5, 10000, 1.0)
    # every division from the tenth snapshot on is one recorded time (no cell is removed)
    assert times.size == counts[100] - counts[9]
    assert np.all(times > 0) and np.all((age >= 0.9) & (age <= 10.0))
    # a daughter is born at most one interval before its mother's next division is seen
    assert np.all(times >= 0.0)
    for i in (50, 100):
        h = area_hist(frames[i][3])
        assert abs(float(np.sum(h * np.diff(AREA_EDGES))) - 1.0) < 1e-12
    h, _ = np.histogram(times, bins=DIVISION_EDGES, density=True)
    assert abs(float(np.sum(h * np.diff(DIVISION_EDGES))) - 1.0) < 1e-12
