"""GPU parity of the experimental variograms against the CPU oracle.

The device sums each offset's pairs in a fixed tree order and the host
adds the offsets' sums per lag, where the oracle (like the reference) keeps
one running sum per lag; both are double sums of the same exact squared
differences, so the float results agree to within an ulp or two (tested at
1e-6 relative, NaN for empty lags in both).  Counts of pairs are exact.
"""
import json
import os

import numpy as np
import pytest

import oracle_binding as O
from cases import Grid

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "variogram")


@pytest.fixture(scope="module")
def cv():
    from hpgl_amd.geo_bsd import cvariogram

    return cvariogram


def _close(got, want, tol=1e-6):
    fin = np.isfinite(want)
    assert np.array_equal(np.isfinite(got), fin)
    assert np.max(np.abs(got[fin] - want[fin]) / np.maximum(np.abs(want[fin]), 1e-6), initial=0.0) <= tol


@pytest.mark.parametrize("shape,frac,az,dip,lags", [((40, 36, 12), 0.3, 0, 0, 10), ((64, 48, 20), 0.05, 35, 10, 16),
                                                    ((30, 30, 30), 0.6, 120, 45, 8)])
def test_grid_variogram_equals_oracle(cv, shape, frac, az, dip, lags):
    rng = np.random.default_rng(sum(shape))
    vals = np.asfortranarray(rng.lognormal(0, 0.7, shape).astype(np.float32))
    mask = np.asfortranarray((rng.random(shape) < frac).astype(np.uint8))
    print("grid variogram", shape, frac, az, dip, flush=True)
    ell = cv.Ellipsoid(1, 4.0, 2.5, az, dip, 0)
    t = cv.VariogramSearchTemplate(2.0, 2.0, 1.0, lags, 0.0, ell)
    borders, got = cv.CalcVariograms(t, (vals, mask))
    ot = O.variogram_template(2.0, 2.0, 1.0, lags, 0.0, O.ellipsoid(1, 4.0, 2.5, az, dip, 0))
    want = O.calc_variograms(ot, vals, mask)
    _close(got, want)
    st = cv.last_stats()
    assert st["offsets"] > 0 and st["pairs_tested"] > 0
    assert list(borders) == [2.0 * k for k in range(lags)]


def test_grid_variogram_c_order_strides_and_determinism(cv):
    rng = np.random.default_rng(9)
    shape = (24, 20, 10)
    vals = np.ascontiguousarray(rng.normal(0, 1, shape).astype(np.float32))  # C order: other strides
    mask = np.ascontiguousarray((rng.random(shape) < 0.5).astype(np.uint8))
    ell = cv.Ellipsoid(1, 3.0, 3.0, 60, 0, 0)
    t = cv.VariogramSearchTemplate(1.0, 1.0, 0.8, 12, 0.5, ell)
    _, a = cv.CalcVariograms(t, (vals, mask))
    _, b = cv.CalcVariograms(t, (vals, mask))
    assert np.array_equal(a, b, equal_nan=True)
    ot = O.variogram_template(1.0, 1.0, 0.8, 12, 0.5, O.ellipsoid(1, 3.0, 3.0, 60, 0, 0))
    _close(a, O.calc_variograms(ot, vals, mask))
    # reversed views (negative strides from the first element, as the
    # reference addresses them) give the variogram of their contiguous copies
    rv, rm = vals[::-1, :, ::-1], mask[::-1, :, ::-1]
    _, c = cv.CalcVariograms(t, (rv, rm))
    _, d = cv.CalcVariograms(t, (np.asfortranarray(rv), np.asfortranarray(rm)))
    assert np.array_equal(c, d, equal_nan=True)


def test_grid_variogram_sampling_is_deterministic_and_close(cv):
    rng = np.random.default_rng(4)
    shape = (48, 40, 16)
    vals = np.asfortranarray(rng.normal(0, 1, shape).astype(np.float32))
    mask = np.asfortranarray((rng.random(shape) < 0.5).astype(np.uint8))
    t = cv.VariogramSearchTemplate(2.0, 2.0, 1.0, 8, 0.0, cv.Ellipsoid(1, 3.0, 3.0, 0, 0, 0))
    _, full = cv.CalcVariograms(t, (vals, mask), 100)
    _, s1 = cv.CalcVariograms(t, (vals, mask), 50)
    _, s2 = cv.CalcVariograms(t, (vals, mask), 50)
    assert np.array_equal(s1, s2)
    fin = np.isfinite(full)
    # half of ~1e5 pairs per lag: the mean of squared N(0,2) differences moves by a few percent at most
    assert np.max(np.abs(s1[fin] - full[fin]) / full[fin]) < 0.05


@pytest.mark.parametrize("name", ["2d_vertical", "2d_horizontal", "3d_vertical"])
def test_point_set_equals_oracle_on_problem_5_2(cv, name):
    from test_oracle_variogram import load_gslib

    c = json.load(open(os.path.join(GOLD, "problem_5_2.json")))[name]
    d = load_gslib(os.path.join(GOLD, c["file"]))
    n = len(d[c["value"]])
    col = lambda k: d[k] if k else np.zeros(n)
    e = c["ellipsoid"]
    t = cv.VariogramSearchTemplate(c["lag_width"], c["lag_separation"], c["tol_distance"], c["num_lags"], 0.0,
                                   cv.Ellipsoid(e["R1"], e["R2"], e["R3"], e["azimuth"], e["dip"], e["rotation"]))
    _, got = cv.CalcVariogramsFromPointSet(t, {"X": col(c["x"]), "Y": col(c["y"]), "Z": col(c["z"]),
                                               "Property": col(c["value"])})
    ot = O.variogram_template(c["lag_width"], c["lag_separation"], c["tol_distance"], c["num_lags"], 0.0,
                              O.ellipsoid(e["R1"], e["R2"], e["R3"], e["azimuth"], e["dip"], e["rotation"]))
    want = O.calc_variograms_from_point_set(ot, col(c["x"]), col(c["y"]), col(c["z"]), col(c["value"]))
    _close(got, want)


def test_ellipsoid_and_stack_layers_equal_oracle(cv):
    for args in [(1, 2, 3, 0, 0, 0), (1, 500, 5, 320, 0, 0), (1, 0.1, 10, 0, 90, 0), (2, 3, 4, 17, 33, 71)]:
        a = cv.Ellipsoid(*args).ell
        b = O.ellipsoid(*args)
        for f in ("direction1", "direction2", "direction3"):
            assert list(getattr(a, f).data) == list(getattr(b, f).data)
    rng = np.random.default_rng(2)
    layers = [np.asfortranarray(rng.uniform(-2, 4, (7, 5, 1)).astype(np.float32)) for _ in range(4)]
    r1 = np.zeros((7, 5, 12), np.float32, order="F")
    r2 = r1.copy(order="F")
    cv.CStackLayers(layers, [3, 5, 7, 9], 12, 1.5, -99, r1)
    O.stack_layers(layers, [3, 5, 7, 9], 12, 1.5, -99, r2)
    assert np.array_equal(r1, r2)


def test_errors_are_reported(cv):
    vals = np.zeros((4, 4, 4), np.float32, order="F")
    mask = np.ones((4, 4, 3), np.uint8, order="F")
    t = cv.VariogramSearchTemplate(1.0, 1.0, 1.0, 3, 0.0, cv.Ellipsoid(1, 1, 1, 0, 0, 0))
    with pytest.raises(RuntimeError, match="shapes differ"):
        cv.CalcVariograms(t, (vals, mask))
