"""Pins the CPU oracle against the reference's golden vectors, published
known answers, library routines and closed forms (no GPU needed).

What each test pins is stated in its docstring; tolerances are written in
the asserts.
"""
import json
import os

import numpy as np
import pytest
import scipy.linalg.lapack as lapack
import scipy.special

import oracle_binding as O

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


class CM:
    def __init__(self, type, ranges, angles=(0.0, 0.0, 0.0), sill=1.0, nugget=0.0):
        self.type, self.ranges, self.angles, self.sill, self.nugget = type, ranges, angles, sill, nugget


class G:
    def __init__(self, x, y, z):
        self.x, self.y, self.z = x, y, z


def _p72():
    with open(os.path.join(GOLDEN, "problem_7_2.json")) as f:
        return json.load(f)


def _p72_input(d):
    data = np.full(20, -99, np.float32)
    mask = np.zeros(20, np.uint8)
    for c, v in zip(d["cells"], d["values"]):
        data[c] = v
        mask[c] = 1
    return data, mask


def test_problem_7_2_sk_golden():
    """SK of problem 7.2 (sample-scripts/.../Result/problem_7.2.txt line 1).

    19 of 20 nodes match at print precision.  Node 10 is a tie witness: its
    two candidate neighbours (cells 7 and 13, both 3 cells away) tie in
    covariance and the golden was produced by a build whose std::sort puts
    7 first; this oracle is pinned to GCC-11 libstdc++, which puts 13 first.
    The test checks both readings of node 10 exactly."""
    d = _p72()
    data, mask = _p72_input(d)
    s = d["sk"]
    out, om = O.simple_kriging((data, mask), G(20, 1, 1), s["radiuses"], s["max_neighbours"],
                               CM(s["cov_type"], s["ranges"]))
    exp = np.array(d["expected_sk"], np.float32)
    diff = np.abs(out - exp)
    assert np.all(om == 1)
    assert np.all(np.delete(diff, 10) <= 5e-8), diff
    assert diff[10] > 1e-3
    # node 10 under both tie orders, from SK weights at integer coordinates
    mean = np.float32(np.float64(np.sum(data[mask == 1].astype(np.float64))) / 5)
    for nbrs, expected in (((8, 13), out[10]), ((8, 7), exp[10])):
        rc, w, _ = O.simple_kriging_weights([10, 0, 0], list(nbrs), [0, 0], [0, 0], CM(1, (5, 1, 1)))
        assert rc == 0
        est = np.float32(mean)
        for wi, n in zip(w.astype(np.float64), nbrs):
            est += np.float32(wi * np.float32(data[n] - mean))
        assert abs(float(est) - float(expected)) <= 2e-7, (nbrs, est, expected)


def test_problem_7_2_sgs_golden():
    """Unconditional SGS of problem 7.2 (Result/problem_7.2.txt line 2):
    LCG random path, mt11213b draws, Odeh-Evans inverse, SK on the box
    table, non-parametric CDF back-transform.  All 20 values within 2e-7
    (the golden's residual 1.2e-7 is float rounding of the back-transform
    on the original x87 build)."""
    d = _p72()
    data, mask = _p72_input(d)
    vals = np.unique(data[mask == 1])
    cdf = (vals[:-1].astype(np.float32), (np.arange(1, vals.size) / float(vals.size)).astype(np.float32))
    s = d["sgs"]
    sim, sm = O.sgs_simulation((np.full(20, -99, np.float32), np.zeros(20, np.uint8)), G(20, 1, 1), cdf,
                               s["radiuses"], s["max_neighbours"], CM(s["cov_type"], s["ranges"]), s["seed"])
    assert np.all(sm == 1)
    assert np.abs(sim - np.array(d["expected_sgs"], np.float32)).max() <= 2e-7
    # second SK on the simulated values at the data cells (line 3); node 10 is the same tie witness
    d2 = np.full(20, -99, np.float32)
    d2[mask == 1] = sim[mask == 1]
    out2, _ = O.simple_kriging((d2, mask), G(20, 1, 1), d["sk"]["radiuses"], 2, CM(1, (5, 1, 1)))
    diff = np.abs(out2 - np.array(d["expected_sk2"], np.float32))
    assert np.all(np.delete(diff, 10) <= 3e-7), diff


def test_mt11213b_boost_validation_value():
    """boost::mt11213b's published check: 10000th output from the default seed 5489 is 3809585648."""
    assert int(O.mt11213b(5489, 10000)[-1]) == 3809585648


def _parta():
    with open(os.path.join(GOLDEN, "gstl_parta.json")) as f:
        return json.load(f)


def test_parta_simple_kriging_known_answers():
    """GsTL/GSLIB parta: SK weights (test_kriging.result, printed to 6
    digits from a solver that differs from a double solve by <= 7e-5),
    GsTL SK estimate with mean 9 and variance, GSLIB SK estimate (mean 0)
    2.243 / variance 0.580; and exact agreement with a LAPACK solve."""
    d = _parta()
    pts = np.array(d["points"], float)
    v = np.array(d["values"])
    rc, w, var = O.simple_kriging_weights([41, 29, 0], pts[:, 0], pts[:, 1], np.zeros(8), CM(1, (10, 10, 10)))
    assert rc == 0
    assert np.abs(w - np.array(d["sk_weights"])).max() < 1e-4
    # GsTL's own weights are off by up to 7e-5; times |v - 9| <= 9.7 that is < 1e-3
    assert abs(9 + np.sum(w * (v - 9)) - d["sk_estimate_mean9"]) < 1e-3
    assert abs(var - d["sk_variance"]) < 1e-4
    assert round(float(np.sum(w.astype(np.float64) * v)), 3) == d["gslib_sk_mean0"][0]
    assert round(var, 3) == d["gslib_sk_mean0"][1]
    A = np.exp(-3 * np.linalg.norm(pts[:, None] - pts[None], axis=2) / 10)
    b = np.exp(-3 * np.linalg.norm(pts - [41, 29], axis=1) / 10)
    c, info = lapack.dpotrf(A, lower=0)
    x, info2 = lapack.dpotrs(c, b, lower=0)
    assert info == 0 and info2 == 0
    assert np.abs(w - x).max() < 1e-7  # float32 output of the C ABI


def test_parta_ordinary_kriging_known_answers():
    """OK weights and estimate (GsTL 2.50749, GSLIB 2.507); Lagrange
    multiplier recovered from the system residual matches GsTL's; the
    textbook variance C0 - sum(w b) - lambda matches GsTL 0.582564 /
    GSLIB 0.583; the reference's own variance subtracts the last weight
    instead of lambda (my_kriging_weights.h:246)."""
    d = _parta()
    pts = np.array(d["points"], int)
    v = np.array(d["values"])
    coords = np.c_[pts, np.zeros(8, int)]
    ok, w, var = O.ok_weights_int([41, 29, 0], coords, CM(1, (10, 10, 10)))
    assert ok
    assert abs(np.sum(w) - 1) < 1e-12
    assert np.abs(w - np.array(d["ok_weights"])).max() < 1e-4
    est = float(np.sum(w * v))
    assert abs(est - d["ok_estimate"]) < 1e-3 and round(est, 3) == d["gslib_ok"][0]
    A = np.exp(-3 * np.linalg.norm(pts[:, None] - pts[None], axis=2) / 10)
    b = np.exp(-3 * np.linalg.norm(pts - [41, 29], axis=1) / 10)
    lam = b - A @ w
    assert np.ptp(lam) < 1e-12  # the same multiplier in every row
    lam = lam.mean()
    assert abs(lam - d["ok_lagrange"]) < 1e-4
    textbook = 1 - float(w @ b) - lam
    assert abs(textbook - d["ok_variance"]) < 1e-3 and round(textbook, 3) == d["gslib_ok"][1]
    assert abs(var - (1 - float(w @ b) - w[-1])) < 1e-12
    # bordered Lagrange system solved directly gives the same weights
    M = np.ones((9, 9))
    M[:8, :8] = A
    M[8, 8] = 0
    sol = np.linalg.solve(M, np.r_[b, 1])
    assert np.abs(sol[:8] - w).max() < 1e-12


@pytest.mark.parametrize("n", [1, 2, 3, 5, 7, 12, 24, 33, 64])
def test_cholesky_matches_lapack(n):
    """dpotf2/dpotrs restatement vs scipy's LAPACK dpotrf/dpotrs (same
    algorithm, possibly blocked/vectorised): factor and solution within
    1e-13 relative; a non-SPD matrix reports the same info."""
    rng = np.random.default_rng(n)
    X = rng.standard_normal((n, n + 3))
    A = X @ X.T / n + np.eye(n) * 0.1
    b = rng.standard_normal(n)
    info, x, f = O.cholesky_solve(A, b)
    c, linfo = lapack.dpotrf(A, lower=0)
    xs, _ = lapack.dpotrs(c, b, lower=0)
    assert info == 0 and linfo == 0
    assert np.abs(np.triu(f) - np.triu(c)).max() <= 1e-13 * np.abs(c).max()
    assert np.abs(x - xs).max() <= 1e-11 * np.abs(xs).max()
    B = A.copy()
    k = n // 2
    B[k, k] = -1.0
    info_b, _, _ = O.cholesky_solve(B, b)
    _, linfo_b = lapack.dpotrf(B, lower=0)
    assert info_b == linfo_b and info_b > 0


@pytest.mark.parametrize("size,seed", [(20, 3244759), (97, 5), (1000, 1001), (4096, 7), (99991, 123456),
                                       (4000000, 1064)])
def test_sgs_path_is_full_period(size, seed):
    """The LCG path (path_random_generator.cpp:21-63) starts at seed % size
    and, by Hull-Dobell, visits every cell exactly once per period."""
    p = O.sgs_path(size, seed, size + 1)
    assert p[0] == seed % size
    assert np.array_equal(np.sort(p[:size]), np.arange(size))
    assert p[size] == p[0]


def test_gauss_inverse_and_cdf_accuracy():
    """Odeh-Evans inverse (gaussian_distribution.h:45-63) and the
    Abramowitz-Stegun 26.2.17 CDF (:26-43) against scipy's ndtri/ndtr within
    the published error bounds (1.5e-8 / 7.5e-8, plus rounding)."""
    ps = np.r_[np.linspace(1e-6, 1 - 1e-6, 1001), 1e-10, 0.5]
    xs = np.array([O.gauss_inverse(p) for p in ps])
    assert np.abs(xs - scipy.special.ndtri(ps)).max() < 2e-7  # inverse through a 4th-order rational
    grid = np.linspace(-6, 6, 1201)
    cs = np.array([O.gauss_prob(x) for x in grid])
    assert np.abs(cs - scipy.special.ndtr(grid)).max() < 1e-7


def test_covariance_closed_forms_and_rotation():
    """C(h) = (sill - nugget) f(h/a) for the three models (cov_model.h:144-181),
    the nugget never on the diagonal, anisotropy scaling and a 90 degree
    azimuth that swaps the x and y ranges."""
    for t, f in ((0, lambda x: np.where(x > 1, 0, 1 - 1.5 * x + 0.5 * x ** 3)), (1, lambda x: np.exp(-3 * x)),
                 (2, lambda x: np.exp(-3 * x * x))):
        cm = CM(t, (10, 5, 2), sill=2.0, nugget=0.5)
        assert O.covariance(cm, 0, 0, 0) == 1.5
        for h in (0.5, 3, 9.99, 12):
            assert abs(O.covariance(cm, h, 0, 0) - 1.5 * f(h / 10)) < 1e-14
            assert abs(O.covariance(cm, 0, h / 2, 0) - 1.5 * f(h / 10)) < 1e-14
            assert abs(O.covariance(cm, 0, 0, h / 5) - 1.5 * f(h / 10)) < 1e-14
        rot = CM(t, (10, 5, 2), angles=(90, 0, 0), sill=2.0, nugget=0.5)
        for h in (1, 4, 7):
            assert abs(O.covariance(rot, 0, h, 0) - O.covariance(cm, h, 0, 0)) < 1e-12
            assert abs(O.covariance(rot, h / 2, 0, 0) - O.covariance(cm, 0, h / 2, 0)) < 1e-12
    with pytest.raises(O.OracleError):
        O.ordinary_kriging((np.zeros(8, np.float32), np.ones(8, np.uint8)), G(2, 2, 2), (1, 1, 1), 4, CM(0, (1, 0, 1)))


def test_template_order_and_threshold():
    """Template (covariance_field.h:63-106): exactly the offsets with
    C > C0/100, in non-increasing covariance order; the box-table and
    analytic templates agree within the box."""
    cm = CM(0, (6, 4, 3), angles=(30, 0, 0))
    t_an = O.template(cm, (5, 4, 3), table=0)
    t_tb = O.template(cm, (5, 4, 3), table=1)
    assert np.array_equal(t_an, t_tb)
    keys = np.array([O.covariance(cm, *o) for o in t_an])
    assert np.all(np.diff(keys) <= 0)
    full = [(x, y, z) for z in range(-3, 4) for y in range(-4, 5) for x in range(-5, 6)]
    kept = {o for o in full if O.covariance(cm, *o) > 1.0 / 100}
    assert set(map(tuple, t_an)) == kept
    assert tuple(t_an[0]) == (0, 0, 0)


def _random_case(rng, shape, frac):
    n = int(np.prod(shape))
    mask = np.zeros(n, np.uint8)
    mask[rng.choice(n, max(1, int(frac * n)), replace=False)] = 1
    data = np.where(mask == 1, rng.lognormal(0, 1, n), -99).astype(np.float32)
    return data, mask


def test_cluster_path_lists_are_top_covariances():
    """Cell-binned path (sugarbox_indexed_neighbour_lookup.h:66-123), which
    no shipped fixture reaches: on a tiny dense grid with large radii every
    node takes it; each list must be the max_n highest-covariance informed
    cells of its 27 bins (ties at the cut may go either way), in
    non-increasing covariance order."""
    rng = np.random.default_rng(11)
    shape = (12, 10, 6)
    data, mask = _random_case(rng, shape, 0.05)
    cm = CM(1, (8, 6, 4))
    r = (8, 6, 4)
    tmpl = O.template(cm, r, table=0)
    assert len(tmpl) // 250 >= 3  # cluster limit > 0 so bins are used
    dump = O.Dump(mask.size, 8)
    try:
        O.ordinary_kriging((data, mask), G(*shape), r, 8, cm)
    finally:
        dump.close()
    X, Y, Z = shape
    inf = np.flatnonzero(mask)
    ix, iy, iz = inf % X, inf // X % Y, inf // (X * Y)
    checked = 0
    for node in np.flatnonzero(mask == 0):
        lst = dump.idx[node, :dump.cnt[node]]
        x, y, z = node % X, node // X % Y, node // (X * Y)
        bx, by, bz = x // r[0], y // r[1], z // r[2]
        sel = (np.abs(ix // r[0] - bx) <= 1) & (np.abs(iy // r[1] - by) <= 1) & (np.abs(iz // r[2] - bz) <= 1)
        cand = inf[sel]
        keys = np.array([O.covariance(cm, x - c % X, y - c // X % Y, z - c // (X * Y)) for c in cand])
        keep = keys > 1.0 / 100
        cand, keys = cand[keep], keys[keep]
        order = np.argsort(-keys, kind="stable")
        top = keys[order][:8]
        got = np.array([O.covariance(cm, x - c % X, y - c // X % Y, z - c // (X * Y)) for c in lst])
        assert len(lst) == min(8, len(cand))
        assert np.all(np.diff(got) <= 0)
        assert np.array_equal(np.sort(got), np.sort(top))
        assert set(lst) <= set(cand)
        checked += 1
    assert checked > 600


def test_template_path_lists_match_bruteforce():
    """Template path (sugarbox_neighbour_lookup.h:36-61) against a direct
    scan of the template on a sparse grid where every bin overflows."""
    rng = np.random.default_rng(5)
    shape = (30, 20, 8)
    data, mask = _random_case(rng, shape, 0.1)
    cm = CM(0, (6, 6, 3))
    r = (6, 6, 3)
    tmpl = O.template(cm, r, table=0)
    dump = O.Dump(mask.size, 6)
    try:
        O.ordinary_kriging((data, mask), G(*shape), r, 6, cm)
    finally:
        dump.close()
    X, Y, Z = shape
    for node in np.flatnonzero(mask == 0)[::7]:
        x, y, z = node % X, node // X % Y, node // (X * Y)
        exp = []
        for dx, dy, dz in tmpl:
            px, py, pz = x + dx, y + dy, z + dz
            if 0 <= px < X and 0 <= py < Y and 0 <= pz < Z and mask[px + X * (py + Y * pz)]:
                exp.append(px + X * (py + Y * pz))
                if len(exp) == 6:
                    break
        assert list(dump.idx[node, :dump.cnt[node]]) == exp


def test_failure_policies():
    """Failure handling (cont_kriging.h:85-101): OK leaves a node without
    neighbours untouched; SK writes the mean; duplicate-location-free
    singular systems (zero covariance beyond the spherical range) come out
    as the mean for SK."""
    data = np.full(40, -99, np.float32)
    mask = np.zeros(40, np.uint8)
    data[0], mask[0] = 3.0, 1
    out, om = O.ordinary_kriging((data, mask), G(40, 1, 1), (2, 1, 1), 4, CM(0, (2, 1, 1)))
    assert om[39] == 0 and out[39] == -99 and om[1] == 1
    out, om = O.simple_kriging((data, mask), G(40, 1, 1), (2, 1, 1), 4, CM(0, (2, 1, 1)), mean=1.5)
    assert om[39] == 1 and out[39] == np.float32(1.5)


def test_sgs_batch_equals_serial_calls():
    """The CPU baseline's batch of realisations (one per OpenMP thread) gives
    each realisation exactly as a serial call with that seed."""
    from cases import Cov, Grid, scattered

    g = Grid(30, 20, 10)
    d, m = scattered(g.shape, 0.02, 5, law="normal")
    cov = Cov(0, (8, 8, 3))
    seeds = [1000, 1001, 1002, 1003]
    bd, bm = O.sgs_simulation_batch((d, m), g, (8, 8, 3), 12, cov, seeds)
    for r, s in enumerate(seeds):
        sd, sm = O.sgs_simulation((d, m), g, None, (8, 8, 3), 12, cov, s)
        assert np.array_equal(bd[r], sd.ravel(order="F")) and np.array_equal(bm[r], sm.ravel(order="F"))
