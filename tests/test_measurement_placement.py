"""Closed-form pins of the host-side measurement and placement code (CPU):

* windowStats (aero.ts:33-52): the batch-means standard error of iid noise, in expectation over
  many seeded series, is sigma / sqrt(n); a series that is constant within each of the eight batches
  gives the textbook sample-variance form exactly; the window keeps the last `window` time units.
* crossingFrequency (strouhal.ts:5-30): the frequency of a sampled sinusoid from its upward mean
  crossings (exact up to the linear interpolation between samples), the hysteresis that ignores
  noise near the mean, and the at-least-four-crossings rule.
* placement_matrix / grid_wheels (mesh.ts:321-379): scale, ride height, centre, the sense of yaw
  (about +y) and pitch (positive pitch lowers +x), the wheel transform.
"""
import math

import numpy as np
import pytest

from windtunnel_vis_amd.analysis import crossingFrequency, windowStats
from windtunnel_vis_amd.geometry import Mesh, grid_wheels, placement_matrix


def test_window_stats_se_of_iid_noise():
    sigma, n, reps = 0.3, 4000, 400
    rng = np.random.default_rng(7)
    se2, means = [], []
    t = list(range(n))
    for _ in range(reps):
        v = 1.5 + sigma * rng.standard_normal(n)
        w = windowStats(t, v.tolist(), n)
        se2.append(w["se"] ** 2)
        means.append(w["mean"])
        assert w["n"] == n and w["span"] == n - 1
    # batch means of equal batches: E[se^2] = sigma^2 / n (unbiased)
    assert np.mean(se2) == pytest.approx(sigma ** 2 / n, rel=0.1)
    assert np.std(means) == pytest.approx(sigma / math.sqrt(n), rel=0.15)


def test_window_stats_exact_batches_and_window():
    # 80 samples at t = 0..79, value = batch index (10 samples per batch): mean 3.5, batch means 0..7,
    # sample variance of the batch means 6, se = sqrt(6 / 8)
    t = list(range(80))
    v = [float(i // 10) for i in t]
    w = windowStats(t, v, 79)
    assert w["mean"] == pytest.approx(3.5, abs=1e-12)
    assert w["se"] == pytest.approx(math.sqrt(6 / 8), abs=1e-12)
    # a window of 39 time units keeps t = 40..79 (values 4..7 in batches of 5): mean 5.5, batch means
    # 4, 4, 5, 5, 6, 6, 7, 7: variance 8/7 (with ties at the batch edges resolved as floor)
    w2 = windowStats(t, v, 39)
    assert w2["n"] == 40 and w2["span"] == 39
    assert w2["mean"] == pytest.approx(5.5, abs=1e-12)
    # fewer than 16 samples: undefined
    assert math.isnan(windowStats(t[:15], v[:15], 100)["mean"])


@pytest.mark.parametrize("f,dt", [(0.01, 1.0), (1 / 2353.0, 59.0), (0.037, 0.5)])
def test_crossing_frequency_of_a_sinusoid(f, dt):
    n = int(12 / (f * dt))
    t = [k * dt for k in range(n)]
    v = [0.2 + 0.7 * math.sin(2 * math.pi * f * x + 0.3) for x in t]
    r = crossingFrequency(t, v)
    # linear interpolation of a sine across a sample step: error O((2 pi f dt)^2)
    tol = 2 * (2 * math.pi * f * dt) ** 2 + 1e-9
    assert r["f"] == pytest.approx(f, rel=tol)
    assert r["amplitude"] == pytest.approx(0.7, rel=(2 * math.pi * f * dt) ** 2)
    assert r["cycles"] in (10, 11)


def test_crossing_frequency_hysteresis_and_minimum_cycles():
    rng = np.random.default_rng(3)
    f, dt = 0.02, 1.0
    t = np.arange(0, 600, dt)
    clean = np.sin(2 * math.pi * f * t)
    noisy = clean + 0.05 * rng.standard_normal(len(t))  # 5 % noise: many spurious mean crossings
    r = crossingFrequency(t.tolist(), noisy.tolist())
    assert r["f"] == pytest.approx(f, rel=0.01)
    # three upward crossings are not enough (strouhal.ts:26)
    short = np.arange(0, 3.5 / f, dt)
    r3 = crossingFrequency(short.tolist(), np.sin(2 * math.pi * f * short + 0.1).tolist())
    assert math.isnan(r3["f"]) and r3["cycles"] <= 2


def _box_mesh(lo, hi, wheels=None):
    x0, y0, z0 = lo
    x1, y1, z1 = hi
    p = np.array([[x, y, z] for x in (x0, x1) for y in (y0, y1) for z in (z0, z1)], np.float32)
    idx = np.array([[0, 1, 3], [0, 3, 2], [4, 6, 7], [4, 7, 5], [0, 4, 5], [0, 5, 1], [2, 3, 7], [2, 7, 6],
                    [0, 2, 6], [0, 6, 4], [1, 5, 7], [1, 7, 3]], np.uint32)
    return Mesh(p, idx, wheels)


def _apply(m16, p):
    M = np.asarray(m16, np.float64).reshape(4, 4).T  # column-major
    return (M @ np.r_[np.asarray(p, np.float64), 1.0])[:3]


def test_placement_ground_mode_scale_ride_and_centre():
    dims = (300, 100, 80)
    mesh = _box_mesh((-0.5, 0.0, -0.2), (0.5, 0.3, 0.2))
    m, scale = placement_matrix(mesh, dims, mode="ground", length_frac=1 / 3, x_frac=0.34, ride_cells=2.5)
    assert scale == pytest.approx(100.0)
    lo = _apply(m, (-0.5, 0.0, -0.2))
    hi = _apply(m, (0.5, 0.3, 0.2))
    assert lo == pytest.approx([0.34 * 300 - 50, 2.5, 40 - 20], abs=1e-4)
    assert hi == pytest.approx([0.34 * 300 + 50, 2.5 + 30, 40 + 20], abs=1e-4)
    # centre mode: height = diam_frac ny, centre at ny / 2 + 0.37
    m2, s2 = placement_matrix(mesh, dims, mode="center", diam_frac=0.3)
    assert s2 == pytest.approx(0.3 * 100 / 0.3)
    assert _apply(m2, (0.0, 0.15, 0.0)) == pytest.approx([0.34 * 300, 50.37, 40], abs=1e-4)


def test_placement_yaw_and_pitch_sense():
    dims = (300, 100, 80)
    mesh = _box_mesh((-0.5, 0.0, -0.2), (0.5, 0.3, 0.2))
    # yaw +90 about +y: the mesh's +x end goes to -z (x' = cos x + sin z, z' = -sin x + cos z)
    m, s = placement_matrix(mesh, dims, yaw_deg=90.0, ride_cells=0.0)
    nose = _apply(m, (0.5, 0.15, 0.0))
    centre = _apply(m, (0.0, 0.15, 0.0))
    assert nose[2] - centre[2] == pytest.approx(-50.0, abs=1e-4)
    assert nose[0] - centre[0] == pytest.approx(0.0, abs=1e-4)
    # pitch +10: rotation by -10 degrees about z, so +x goes down by sin(10) per unit length
    mp, _ = placement_matrix(mesh, dims, pitch_deg=10.0, ride_cells=0.0)
    a, b = _apply(mp, (0.5, 0.15, 0.0)), _apply(mp, (-0.5, 0.15, 0.0))
    assert a[1] - b[1] == pytest.approx(-100 * math.sin(math.radians(10)), abs=1e-4)
    # on the ground the lowest rotated vertex sits on the floor: the front bottom corner
    assert min(_apply(mp, c)[1] for c in mesh.positions) == pytest.approx(0.0, abs=1e-4)


def test_grid_wheels_transform():
    dims = (300, 100, 80)
    wheels = [dict(c=[0.3, 0.05, 0.15], axis=[0.0, 0.0, 2.0], r=0.05, hw=0.02)]
    mesh = _box_mesh((-0.5, 0.0, -0.2), (0.5, 0.3, 0.2), wheels)
    m, s = placement_matrix(mesh, dims, yaw_deg=30.0, ride_cells=1.0)
    g = grid_wheels(mesh, m)[0]
    assert g["c"] == pytest.approx(list(_apply(m, wheels[0]["c"])), abs=1e-4)
    # the axis is rotated and normalised; the half width scales with the transformed axis length
    # (2 x 100 for the given axis), the radius with the scale
    yaw = math.radians(30.0)
    assert g["axis"] == pytest.approx([math.sin(yaw), 0.0, math.cos(yaw)], abs=1e-6)
    assert g["r"] == pytest.approx(0.05 * 100, rel=1e-6)
    assert g["hw"] == pytest.approx(0.02 * 200, rel=1e-6)
