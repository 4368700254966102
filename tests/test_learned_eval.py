"""The equal-L1 estimate of the learned-correction evaluation (host logic,
no GPU): the plain resolution and wall time at the learned run's density L1,
interpolated in log-log between the bracketing measured runs."""
import math

import numpy as np

from hyperflow_amd import learned as L


def test_equal_l1_point_hits_measured_points_and_power_laws():
    ns = [1024, 1152, 1280, 1536]
    l1 = [2.0 * n ** -1.5 for n in ns]
    ts = [3e-9 * n ** 3 for n in ns]
    # at a measured point: that point
    n, t = L.equal_l1_point(ns, l1, ts, l1[2])
    assert math.isclose(n, 1280, rel_tol=1e-12) and math.isclose(t, ts[2], rel_tol=1e-12)
    # between points, exact power laws are reproduced
    target = 2.0 * 1100.0 ** -1.5
    n, t = L.equal_l1_point(ns, l1, ts, target)
    assert math.isclose(n, 1100.0, rel_tol=1e-12)
    assert math.isclose(t, 3e-9 * 1100.0 ** 3, rel_tol=1e-12)


def test_equal_l1_point_outside_the_measured_range_is_nan():
    ns, l1, ts = [1024, 1280], [1e-3, 8e-4], [0.01, 0.02]
    assert all(math.isnan(v) for v in L.equal_l1_point(ns, l1, ts, 2e-3))
    assert all(math.isnan(v) for v in L.equal_l1_point(ns, l1, ts, 5e-4))


def test_summary_falls_back_to_the_fits_outside_the_range():
    rows = []
    for seed in range(2):
        rows.append({"plain": {n: {"l1": 2.0 * n ** -1.5, "seconds": 3e-9 * n ** 3} for n in (1024, 1280)},
                     "learned": {8: {"l1": 2.0 * 900.0 ** -1.5, "seconds": 0.005}}, "fine_seconds": 0.05})
    s = L.summarize(rows, (1024, 1280), (8,))["learned"][8]
    assert math.isclose(s["plain_n_at_equal_l1"], 900.0, rel_tol=1e-9)
    assert math.isclose(s["speedup_at_equal_l1"], s["speedup_at_equal_l1_fit"], rel_tol=1e-12)
    assert np.isfinite(s["speedup_at_equal_l1"])
