"""Pins the oracle's experimental variograms (CPU only).

* Point sets: the printed semivariograms of the reference's solved problem
  5.2 (tests/golden/variogram/, made by tests/golden/make_fixtures.py from
  the problem's data files and results).  Those were computed by the
  reference's Python scan with float32 accumulation, so they agree to
  ~2e-5 relative; a wrong tunnel test, lag rule or factor of two misses by
  far more.
* Grids: a brute force written here over all informed cell pairs of small
  grids (numpy), and a linear trend along x whose semivariogram at lag k is
  a^2 k^2 / 2 exactly.
* Ellipsoid: closed forms of the rotation (with the reference's 3.14159265358).
* Layers: a hand-worked column.
"""
import json
import os

import numpy as np
import pytest

import oracle_binding as O
from cases import Grid

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "variogram")


def load_gslib(path):
    lines = open(path).read().split("\n")
    n = int(lines[1])
    names = [ln.strip() for ln in lines[2:2 + n]]
    rows = np.array([[float(t) for t in ln.split()] for ln in lines[2 + n:] if ln.strip()])
    return {nm: rows[:, i] for i, nm in enumerate(names)}


def golden_cases():
    return json.load(open(os.path.join(GOLD, "problem_5_2.json")))


@pytest.mark.parametrize("name,tol", [("2d_vertical", 1e-6), ("2d_horizontal", 1e-6), ("3d_vertical", 5e-5)])
def test_point_set_matches_problem_5_2(name, tol):
    c = golden_cases()[name]
    d = load_gslib(os.path.join(GOLD, c["file"]))
    n = len(d[c["value"]])
    col = lambda k: d[k] if k else np.zeros(n)
    e = c["ellipsoid"]
    t = O.variogram_template(c["lag_width"], c["lag_separation"], c["tol_distance"], c["num_lags"], 0.0,
                             O.ellipsoid(e["R1"], e["R2"], e["R3"], e["azimuth"], e["dip"], e["rotation"]))
    got = O.calc_variograms_from_point_set(t, col(c["x"]), col(c["y"]), col(c["z"]), col(c["value"]))
    want = np.array(c["expected"])
    assert np.max(np.abs(got - want) / np.maximum(np.abs(want), 1.0)) <= tol, (got, want)


def brute_grid(values, mask, templ_args, ell):
    """All informed pairs of a small grid: offset in the tunnel, lag ranges
    by the closed-interval rule of update_lags, semivariance per lag."""
    lw, ls, tol, nl, first = templ_args
    d1, d2, d3 = (np.array(v.data) for v in (ell.direction1, ell.direction2, ell.direction3))
    X, Y, Z = values.shape
    idx = np.argwhere(mask != 0)
    v = values[mask != 0].astype(np.float64)
    s = np.zeros(nl)
    c = np.zeros(nl)
    for a in range(len(idx)):
        off = idx - idx[a]
        ss1, ss2, ss3 = off @ d1, off @ d2, off @ d3
        td = np.sqrt((ss2 / ell.R2) ** 2 + (ss3 / ell.R3) ** 2)
        ok = (td <= 1.0) & (tol * td <= ss1)
        dist = np.sqrt((off.astype(np.float64) ** 2).sum(axis=1))
        var = (v[a] - v) ** 2
        lo = np.ceil((dist - lw / 2 - first) / ls)
        hi = np.floor((dist + lw / 2 - first) / ls)
        for k in range(nl):
            sel = ok & (lo <= k) & (k <= hi)
            s[k] += var[sel].sum()
            c[k] += sel.sum()
    with np.errstate(invalid="ignore", divide="ignore"):
        return (s / c / 2).astype(np.float32)


@pytest.mark.parametrize("seed,az,dip", [(1, 0, 0), (2, 45, 0), (3, 30, 60)])
def test_grid_matches_bruteforce(seed, az, dip):
    rng = np.random.default_rng(seed)
    g = Grid(11, 9, 6)
    vals = np.asfortranarray(rng.normal(0, 1, g.shape).astype(np.float32))
    mask = np.asfortranarray((rng.random(g.shape) < 0.4).astype(np.uint8))
    ell = O.ellipsoid(1, 3.0, 2.0, az, dip, 0)
    args = (1.5, 1.5, 0.5, 6, 0.0)
    t = O.variogram_template(*args[:4], args[4], ell)
    got = O.calc_variograms(t, vals, mask)
    want = brute_grid(vals, mask, args, ell)
    both = np.isfinite(want)
    assert np.array_equal(np.isfinite(got), both)
    assert np.max(np.abs(got[both] - want[both]) / np.maximum(np.abs(want[both]), 1e-3)) <= 1e-6


def test_grid_linear_trend_closed_form():
    g = Grid(20, 4, 3)
    x = np.arange(g.x, dtype=np.float32)[:, None, None]
    vals = np.asfortranarray(np.broadcast_to(0.5 * x, g.shape).astype(np.float32))
    mask = np.ones(g.shape, np.uint8, order="F")
    # a thin tunnel along +x: only offsets (i, 0, 0), i >= 0
    t = O.variogram_template(0.5, 1.0, 0.0, 8, 0.0, O.ellipsoid(1, 0.5, 0.5, 0, 0, 0))
    got = O.calc_variograms(t, vals, mask)
    k = np.arange(8)
    assert np.array_equal(got, (0.25 * k * k / 2).astype(np.float32))


def test_ellipsoid_closed_forms():
    e = O.ellipsoid(1, 2, 3, 0, 0, 0)
    assert [list(v.data) for v in (e.direction1, e.direction2, e.direction3)] == [[1, 0, 0], [0, 1, 0], [0, 0, 1]]
    c = 3.14159265358 / 180.0
    for az, dip, rot in [(30, 0, 0), (0, 40, 0), (0, 0, 50), (320, 10, 5)]:
        e = O.ellipsoid(1, 1, 1, az, dip, rot)
        A = np.array([[np.cos(az * c), -np.sin(az * c), 0], [np.sin(az * c), np.cos(az * c), 0], [0, 0, 1]])
        B = np.array([[np.cos(dip * c), 0, -np.sin(dip * c)], [0, 1, 0], [np.sin(dip * c), 0, np.cos(dip * c)]])
        Cm = np.array([[1, 0, 0], [0, np.cos(rot * c), -np.sin(rot * c)], [0, np.sin(rot * c), np.cos(rot * c)]])
        M = A @ B @ Cm
        got = np.array([list(v.data) for v in (e.direction1, e.direction2, e.direction3)]).T
        assert np.max(np.abs(got - M)) <= 1e-15
        assert np.max(np.abs(got.T @ got - np.eye(3))) <= 1e-14


def test_stack_layers_column():
    # one column, three layers: 2.5 units, a negative (erosion) of 1 unit,
    # then 2 units; scalez 1, nz 6, blank -1
    layers = [np.full((1, 1, 1), t, np.float32, order="F") for t in (2.5, -1.0, 2.0)]
    res = np.zeros((1, 1, 6), np.float32, order="F")
    O.stack_layers(layers, [7, 8, 9], 6, 1.0, -1, res)
    # layer 7 fills k = 0..2 (ceil(0) .. floor(2.5)); the erosion blanks from
    # ceil(2.5 - 1) = 2 to the top; layer 9 then fills ceil(1.5) = 2 .. floor(3.5)
    assert list(res.ravel()) == [7, 7, 9, 9, -1, -1]
