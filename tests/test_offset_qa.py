"""Offset-field quality control (cull, offoutliers, fitoff, polynomial fits).

Pinned by closed forms and brute force rather than by retyping the
reference: ordinary least squares in closed form for the straight-line fits,
the point-to-line distance |s t - v + c| / sqrt(1 + s^2) for the outlier
distances, exact recovery of affine maps and polynomials from noise-free
data, an exhaustive search over interpolating triples for the L1 optimum, and
the known Householder decomposition of rotation/scale/skew matrices.
"""
import itertools

import numpy as np
import pytest

from isce2_amd.offset_qa import OffsetField, eval_poly, fit_offset_polynomials, fitoff, offoutliers, svd_solve


def _field(n=400, seed=0, outliers=0):
    rng = np.random.default_rng(seed)
    x = rng.uniform(0, 4000, n)
    y = rng.uniform(0, 8000, n)
    dx = 1.5 + 2e-4 * x + rng.normal(0, 0.02, n)
    dy = -0.75 + 1e-4 * y + rng.normal(0, 0.02, n)
    if outliers:
        k = rng.choice(n, outliers, replace=False)
        dx[k] += rng.choice([-1, 1], outliers) * rng.uniform(2, 5, outliers)
    snr = rng.uniform(1, 30, n)
    return OffsetField(x, dx, y, dy, snr)


def test_cull_keeps_snr_at_or_above_threshold():
    f = OffsetField([0, 1, 2, 3], [0] * 4, [0] * 4, [0] * 4, [1.0, 2.0, np.nan, 3.0])
    assert list(f.cull(2.0).x) == [1, 2, 3]           # snr < 2 dropped; NaN is not < 2
    assert len(f.cull(0.0)) == 4


def test_from_dense_grid_positions():
    off = np.zeros((3, 4, 2), np.float32)
    off[..., 0] = np.arange(12).reshape(3, 4)         # down
    off[..., 1] = -np.arange(12).reshape(3, 4)        # across
    f = OffsetField.from_dense(off, np.ones((3, 4)), np.zeros((3, 4, 3)), x_start=52, y_start=40, x_step=32,
                               y_step=16)
    assert f.x[:4].tolist() == [52, 84, 116, 148] and f.y[::4].tolist() == [40, 56, 72]
    assert np.array_equal(f.dy, np.arange(12)) and np.array_equal(f.dx, -np.arange(12))


def test_svd_solve_is_least_squares():
    rng = np.random.default_rng(1)
    a = rng.standard_normal((50, 4))
    b = rng.standard_normal(50)
    assert np.allclose(svd_solve(a, b), np.linalg.solve(a.T @ a, a.T @ b))
    # a rank-deficient column pair: the tiny singular value is dropped (minimum-norm answer)
    a2 = np.concatenate([a, a[:, :1]], axis=1)
    c = svd_solve(a2, b)
    assert np.allclose(a2 @ c, a @ np.linalg.solve(a.T @ a, a.T @ b))


def test_offoutliers_lines_and_distances():
    f = _field(outliers=25)
    r = offoutliers(f, distance=0.5, snr_threshold=2.0)
    g = f.cull(2.0)
    # closed-form ordinary least squares of dx on x and dy on y
    for t, v, s, c in ((g.x, g.dx, r["slope_across"], r["average_offset_across"]),
                       (g.y, g.dy, r["slope_down"], r["average_offset_down"])):
        tm, vm = t.mean(), v.mean()
        s0 = ((t - tm) * (v - vm)).sum() / ((t - tm) ** 2).sum()
        assert np.isclose(s, s0, rtol=1e-9) and np.isclose(c, vm - s0 * tm, rtol=1e-9, atol=1e-12)
    dist = lambda t, v, s, c: np.abs(s * t - v + c) / np.sqrt(1 + s * s)
    keep = (dist(g.x, g.dx, r["slope_across"], r["average_offset_across"]) <= 0.5) & \
           (dist(g.y, g.dy, r["slope_down"], r["average_offset_down"]) <= 0.5)
    assert np.array_equal(r["index"], np.nonzero(keep)[0])
    assert len(r["refined"]) == keep.sum() and keep.sum() < len(g)


def _affine_data(n=300, seed=2, outliers=0, noise=0.0):
    rng = np.random.default_rng(seed)
    x = rng.uniform(0, 5000, n)
    y = rng.uniform(0, 20000, n)
    m = np.array([[1.0001, 2e-5], [-3e-5, 0.99995]])
    t = np.array([12.25, -7.5])
    x2 = m[0, 0] * x + m[0, 1] * y + t[0] + rng.normal(0, noise, n)
    y2 = m[1, 0] * x + m[1, 1] * y + t[1] + rng.normal(0, noise, n)
    if outliers:
        k = rng.choice(n, outliers, replace=False)
        x2[k] += rng.uniform(3, 9, outliers)
        y2[k] -= rng.uniform(3, 9, outliers)
    return OffsetField(x, x2 - x, y, y2 - y, np.full(n, 10.0)), m, t


@pytest.mark.parametrize("l1", [True, False])
def test_fitoff_recovers_affine_exactly(l1):
    f, m, t = _affine_data()
    r = fitoff(f, minpoint=10, l1norm=l1)
    assert r["status"] == "converged" and r["iterations"] == 1 and len(r["refined"]) == len(f)
    assert np.allclose(r["affine"], [m[0, 0], m[0, 1], m[1, 0], m[1, 1], t[0], t[1]], atol=1e-7)
    assert r["average_offset_across"] == r["affine"][4] and r["average_offset_down"] == r["affine"][5]


def test_fitoff_culls_outliers():
    f, m, t = _affine_data(n=400, outliers=30, noise=0.01)
    r = fitoff(f, minpoint=10, l1norm=True)
    assert r["status"] == "converged" and r["rms_x"] < 0.08 and r["rms_y"] < 0.08
    bad = set(np.nonzero(np.abs(f.dx - ((m[0, 0] - 1) * f.x + m[0, 1] * f.y + t[0])) > 1)[0])
    assert not bad & set(r["index"].tolist())
    assert np.allclose(r["affine"][4:], t, atol=0.05)
    # too few points: stops with the min_points status
    assert fitoff(f, minpoint=1000, l1norm=False)["status"] == "min_points"


def test_l1_fit_matches_brute_force_optimum():
    # the L1 optimum of a 3-parameter fit interpolates 3 of the points: enumerate them
    rng = np.random.default_rng(4)
    n = 9
    x, y = rng.uniform(0, 10, n), rng.uniform(0, 10, n)
    x2 = 0.5 * x - 0.2 * y + 1.0 + rng.standard_t(1.5, n)
    f = OffsetField(x, x2 - x, y, np.zeros(n), np.ones(n))
    r = fitoff(f, maxrms=1e9, minpoint=0, l1norm=True)       # one fit, no culling
    a = np.stack([x, y, np.ones(n)], axis=1)
    px = r["affine"][[0, 1, 4]]
    best = np.inf
    for tri in itertools.combinations(range(n), 3):
        s = a[list(tri)]
        if abs(np.linalg.det(s)) < 1e-9:
            continue
        c = np.linalg.solve(s, x2[list(tri)])
        best = min(best, np.abs(a @ c - x2).sum())
    assert np.isclose(np.abs(a @ px - x2).sum(), best, rtol=1e-9)


def test_affine_decomposition_rotation_scale_skew():
    th = np.radians(0.7)
    rot = np.array([[np.cos(th), -np.sin(th)], [np.sin(th), np.cos(th)]])
    up = np.array([[1.002, 0.003], [0.0, 0.997]])           # scales and skew (upper triangular)
    m = rot @ up
    x = np.array([0.0, 1000, 0, 1000, 500])
    y = np.array([0.0, 0, 1000, 1000, 300])
    xy2 = m @ np.stack([x, y]) + np.array([[5.0], [-2.0]])
    r = fitoff(OffsetField(x, xy2[0] - x, y, xy2[1] - y, np.ones(5)), minpoint=0, l1norm=False)
    assert np.isclose(r["rotation_deg"], 0.7, atol=1e-9)
    assert np.allclose(r["scale"], (1.002, 0.997), atol=1e-9)
    assert np.isclose(r["skew"], 0.003 / 1.002, atol=1e-9)


@pytest.mark.parametrize("order,maxo", [(1, True), (1, False), (2, True), (3, True), (2, False)])
def test_polynomial_fits_recover_noise_free_polynomials(order, maxo):
    rng = np.random.default_rng(5)
    n = 500
    x, y = rng.uniform(0, 3000, n), rng.uniform(100, 9000, n)
    az = np.zeros((order + 1, order + 1))
    rg = np.zeros((order + 1, order + 1))
    for i in range(order + 1):
        for j in range(order + 1):
            if (maxo and i + j <= order) or (not maxo and (order > 1 or True)):
                az[i, j] = rng.normal() * 10.0 ** (-3 * (i + j))
                rg[i, j] = rng.normal() * 10.0 ** (-3 * (i + j))
    if order == 1 and maxo:
        az[1, 1] = rg[1, 1] = 0.0
    f = OffsetField(x, eval_poly(rg, y - 100.0, x), y, eval_poly(az, y - 100.0, x), rng.uniform(2, 50, n))
    a, r, azmin = fit_offset_polynomials(f, order, order, maxo)
    assert azmin == y.min()
    assert np.allclose(eval_poly(a, y, x, azmin), f.dy, atol=1e-6)
    assert np.allclose(eval_poly(r, y, x, azmin), f.dx, atol=1e-6)
