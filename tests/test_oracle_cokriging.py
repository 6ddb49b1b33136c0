"""Pins the oracle's simple colocated cokriging (CPU only).

* gauss_solve (gauss_solver.cpp:5-65): numpy's LAPACK solve on random
  well-conditioned systems; a zero column is singular; on a permutation
  matrix the back substitution's start from B[i] (not B[order[i]]) returns
  B itself, a closed form of the reference's quirk.
* simple_cokriging_markI / markII (simple_cokriging_markI.cpp:19-305): a
  brute-force Python restatement written from the reference's description
  (template neighbours scanned cell by cell with the analytic model, the
  (n+1)-system solved by numpy, float-rounded combine) on tiny grids, on
  the separately pinned template order and covariance model; estimates
  within 1e-6 relative.  Dropping the secondary term, using the wrong
  cross-covariance model or factor, or the wrong failure value changes
  them by far more.
"""
import numpy as np
import pytest

import oracle_binding as O
from cases import Cov, Grid


def test_gauss_solve_against_lapack():
    rng = np.random.default_rng(3)
    for n in (1, 2, 5, 13, 25, 65):
        a = rng.standard_normal((n, n))
        a = a @ a.T + n * np.eye(n)
        b = rng.standard_normal(n)
        ok, x = O.gauss_solve(a, b)
        assert ok
        ref = np.linalg.solve(a, b)
        assert np.max(np.abs(x - ref)) <= 1e-12 * max(1.0, np.max(np.abs(ref)))


def test_gauss_solve_singular_and_permutation_quirk():
    ok, _ = O.gauss_solve([[1.0, 0.0], [2.0, 0.0]], [1.0, 1.0])
    assert not ok
    # pivots out of row order: X[i] starts from B[i], so the permutation
    # system returns B unchanged (the true solution is B reversed)
    ok, x = O.gauss_solve([[0.0, 1.0], [1.0, 0.0]], [2.0, 3.0])
    assert ok and list(x) == [2.0, 3.0]


def _cov_fn(cm):
    cache = {}

    def c(dx, dy, dz):
        k = (dx, dy, dz)
        if k not in cache:
            cache[k] = O.covariance(cm, float(dx), float(dy), float(dz))
        return cache[k]

    return c


def brute_cokriging(data, mask, sdata, smask, grid, radii, max_n, cm, pm, sm, rho, var2=None, cm2=None):
    X, Y, Z = grid.shape
    N = grid.size
    v = np.asarray(data, np.float32).ravel(order="F")
    m = np.asarray(mask, np.uint8).ravel(order="F")
    sv_all = np.asarray(sdata, np.float32).ravel(order="F")
    sm_all = np.asarray(smask, np.uint8).ravel(order="F")
    tmpl = O.template(cm, radii, False)
    c1f = _cov_fn(cm)
    c1 = c1f(0, 0, 0)
    if cm2 is None:
        crossf = c1f
        sec_var = var2
        coef = rho * np.sqrt(var2) / np.sqrt(c1)
    else:
        crossf = _cov_fn(cm2)
        sec_var = crossf(0, 0, 0)
        coef = rho * np.sqrt(sec_var / c1)
    cross0 = coef * crossf(0, 0, 0)
    pm32, sm32 = np.float32(pm), np.float32(sm)
    out = np.zeros(N, np.float32)
    for node in range(N):
        if m[node]:
            out[node] = v[node]
            continue
        sv = sv_all[node] if sm_all[node] else sm32
        x, y, z = node % X, node // X % Y, node // (X * Y)
        nb = []
        for dx, dy, dz in tmpl:
            if len(nb) >= max_n:
                break
            xx, yy, zz = x + dx, y + dy, z + dz
            if 0 <= xx < X and 0 <= yy < Y and 0 <= zz < Z:
                cell = xx + X * (yy + Y * zz)
                if m[cell]:
                    nb.append((cell, xx, yy, zz))
        w = None
        if nb:
            n = len(nb)
            A = np.zeros((n + 1, n + 1))
            b = np.zeros(n + 1)
            for i, p in enumerate(nb):
                for j, q in enumerate(nb):
                    A[i, j] = c1f(q[1] - p[1], q[2] - p[2], q[3] - p[3])
                A[i, n] = A[n, i] = coef * crossf(p[1] - x, p[2] - y, p[3] - z)
                b[i] = c1f(x - p[1], y - p[2], z - p[3])
            A[n, n] = sec_var
            b[n] = cross0
            try:
                w = np.linalg.solve(A, b)
            except np.linalg.LinAlgError:
                w = None
        if w is None:
            out[node] = pm32 + sv - sm32
            continue
        r = pm32
        for i, p in enumerate(nb):
            r = np.float32(float(r) + float(v[p[0]] - pm32) * w[i])
        r = np.float32(float(r) + w[len(nb)] * float(sv - sm32))
        out[node] = r
    return out.reshape(grid.shape, order="F")


def _case(grid, frac, seed):
    rng = np.random.default_rng(seed)
    n = grid.size
    idx = rng.choice(n, max(1, int(frac * n)), replace=False)
    mask = np.zeros(n, np.uint8)
    mask[idx] = 1
    data = np.full(n, -99.0, np.float32)
    data[idx] = rng.lognormal(0, 0.5, idx.size).astype(np.float32)
    smask = (rng.random(n) < 0.7).astype(np.uint8)
    sdata = np.where(smask == 1, rng.normal(2.0, 0.4, n), -99.0).astype(np.float32)
    f = lambda a: a.reshape(grid.shape, order="F")
    return f(data), f(mask), f(sdata), f(smask)


@pytest.mark.parametrize("seed,cov_type", [(1, 1), (2, 0), (3, 2)])
def test_markI_matches_bruteforce(seed, cov_type):
    g = Grid(9, 8, 4)
    data, mask, sdata, smask = _case(g, 0.12, seed)
    cm = Cov(cov_type, (6.0, 4.0, 3.0), (20.0, 0.0, 0.0), 1.0, 0.1 if cov_type == 2 else 0.0)
    od, om = O.simple_cokriging_markI((data, mask), g, (sdata, smask), 1.2, 2.0, 0.16, 0.6, (4, 4, 2), 6, cm)
    bd = brute_cokriging(data, mask, sdata, smask, g, (4, 4, 2), 6, cm, 1.2, 2.0, 0.6, var2=0.16)
    assert np.all(om == 1)
    scale = np.maximum(np.abs(bd), 1.0)
    assert np.max(np.abs(od - bd) / scale) <= 1e-6


@pytest.mark.parametrize("seed", [4, 5])
def test_markII_matches_bruteforce(seed):
    g = Grid(8, 7, 5)
    data, mask, sdata, smask = _case(g, 0.1, seed)
    cm1 = Cov(1, (5.0, 5.0, 3.0))
    cm2 = Cov(0, (7.0, 6.0, 4.0), (0.0, 0.0, 0.0), 0.5, 0.05)
    prim = {"data": (data, mask), "mean": 1.1, "cov_model": cm1}
    sec = {"data": (sdata, smask), "mean": 1.9, "cov_model": cm2}
    od, om = O.simple_cokriging_markII(g, prim, sec, 0.4, (3, 3, 2), 8)
    bd = brute_cokriging(data, mask, sdata, smask, g, (3, 3, 2), 8, cm1, 1.1, 1.9, 0.4, cm2=cm2)
    scale = np.maximum(np.abs(bd), 1.0)
    assert np.max(np.abs(od - bd) / scale) <= 1e-6


def test_failure_value_and_uncorrelated_secondary():
    """No neighbours: primary_mean + secondary - secondary_mean; zero
    correlation with zero secondary variance: the last column is zero, the
    system is singular and the same failure value applies."""
    g = Grid(6, 5, 3)
    data, mask, sdata, smask = _case(g, 0.05, 9)
    cm = Cov(1, (3.0, 3.0, 3.0))
    od, _ = O.simple_cokriging_markI((data, mask), g, (sdata, smask), 0.5, 1.5, 0.0, 0.0, (1, 1, 1), 4, cm)
    want = np.where(smask == 1, np.float32(0.5) + sdata - np.float32(1.5), np.float32(0.5))
    want = np.where(mask == 1, data, want).astype(np.float32)
    assert np.array_equal(od, want)
