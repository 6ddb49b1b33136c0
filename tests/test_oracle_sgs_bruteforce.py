"""Pins the oracle's sequential Gaussian simulation with the ordinary
kriging kind and with a locally varying mean (CPU only).

The problem 7.2 golden pins only simple-kriging SGS.  Here the oracle is
checked against an independent brute-force restatement written from the
reference's description (sequential_simulation.h:51-110,
sequential_gaussian_simulation.cpp:18-128, kriging_interpolation.h:275-313,
my_kriging_weights.h:26-255, combiner.h:10-24, gaussian_distribution.h:47-67,
sample.h:13-16): a Python loop over the path that searches the template
cell by cell among the hard and already simulated cells, solves the SK or
OK system with numpy's LAPACK solver, combines the neighbours in float,
and draws Phi^-1(u) sqrt(var) + mean with the mt11213b draw of the cell
(N(0, 1) when there is no neighbour or the system is not positive
definite).  Its inputs are separately pinned pieces: the template order,
the path LCG and mt11213b (tests/test_oracle.py) and the analytic
covariance.  Values must agree within 1e-5: the solvers differ in rounding
only.  A dropped Lagrange term, the OK variance's subtraction of the last
weight (my_kriging_weights.h:246), a local mean taken at the wrong cell or
a draw consumed out of order changes most values by far more.
"""
import math

import numpy as np
import pytest

import oracle_binding as O
from cases import Cov, Grid, scattered

from test_oracle_sis import _table_cov


def _gauss_inverse(p, mean, var):
    """gaussian_cdf_t::inverse (gaussian_distribution.h:47-67): the Odeh and
    Evans (1974) rational approximation of the normal quantile."""
    ps = (-0.322232431088, -1.0, -0.342242088547, -0.0204231210245, -0.453642210148e-4)
    qs = (0.0993484626060, 0.588581570495, 0.531103462366, 0.103537752850, 0.38560700634e-2)
    p2 = (1 - p) if p > 0.5 else p
    y = math.sqrt(math.log(1 / (p2 * p2)))
    xp = y + ((((ps[4] * y + ps[3]) * y + ps[2]) * y + ps[1]) * y + ps[0]) / \
        ((((qs[4] * y + qs[3]) * y + qs[2]) * y + qs[1]) * y + qs[0])
    if p < 0.5:
        xp = -xp
    sd = math.sqrt(var) if var >= 0 else float("nan")
    return xp * sd + mean


def brute_sgs(data, mask, grid, cov, radii, max_n, seed, kind="sk", lvm=None, sim_mask=None):
    X, Y, Z = grid.shape
    N = grid.size
    vals = np.asarray(data, np.float32).ravel(order="F").copy()
    inf = np.asarray(mask, np.uint8).ravel(order="F").copy()
    sim = None if sim_mask is None else np.asarray(sim_mask, np.uint8).ravel(order="F")
    means = None if lvm is None else np.asarray(lvm, np.float32).ravel(order="F")
    tmpl = O.template(cov, radii, True)
    cv = _table_cov(cov, radii)
    c0 = cv(0, 0, 0)
    path = O.sgs_path(N, seed, N)
    draws = O.mt11213b(seed & 0xFFFFFFFF, N)
    used = 0
    for node in path:
        if inf[node] or (sim is not None and sim[node] != 1):
            continue
        x, y, z = node % X, node // X % Y, node // (X * Y)
        nb = []
        for dx, dy, dz in tmpl:
            if len(nb) >= max_n:
                break
            xx, yy, zz = x + dx, y + dy, z + dz
            if 0 <= xx < X and 0 <= yy < Y and 0 <= zz < Z:
                cell = xx + X * (yy + Y * zz)
                if inf[cell]:
                    nb.append((cell, xx, yy, zz))
        mean, var = 0.0, 1.0
        if nb:
            n = len(nb)
            A = np.array([[cv(q[1] - p[1], q[2] - p[2], q[3] - p[3]) for q in nb] for p in nb])
            b = np.array([cv(x - p[1], y - p[2], z - p[3]) for p in nb])
            try:
                np.linalg.cholesky(A)
                if kind == "sk":
                    w = np.linalg.solve(A, b)
                    v = c0
                    for i in range(n):
                        v -= w[i] * b[i]
                else:
                    sk = np.linalg.solve(A, b)
                    on = np.linalg.solve(A, np.ones(n))
                    mu = (sum(sk) - 1) / sum(on)
                    w = sk - mu * on
                    v = c0
                    for i in range(n):
                        v -= w[i] * b[i]
                    v -= w[n - 1]
                m0 = np.float32(means[node]) if means is not None else np.float32(0.0)
                r = m0
                for i, p in enumerate(nb):
                    mi = np.float32(means[p[0]]) if means is not None else np.float32(0.0)
                    r = np.float32(r + np.float32(w[i] * float(np.float32(vals[p[0]] - mi))))
                mean, var = float(r), float(v)
            except np.linalg.LinAlgError:
                pass
        u = float(draws[used]) / 4294967295.0
        used += 1
        vals[node] = np.float32(_gauss_inverse(u, mean, var))
        inf[node] = 1
    return vals.reshape(grid.shape, order="F"), inf.reshape(grid.shape, order="F")


def _close(a, b, tol=1e-5):
    fa, fb = np.isfinite(a), np.isfinite(b)
    assert np.array_equal(fa, fb), np.argwhere(fa != fb)[:10]
    d = np.abs(a[fa].astype(np.float64) - b[fb].astype(np.float64))
    assert d.max(initial=0.0) <= tol, (d.max(), np.argwhere(np.abs(a - b) > tol)[:10])


CASES = [
    (Cov(0, (4.0, 3.0, 2.5), (0.0, 0.0, 0.0), 1.0, 0.0), 11),
    (Cov(1, (5.0, 4.0, 2.0), (30.0, 0.0, 0.0), 1.0, 0.1), 2024),
    (Cov(2, (4.5, 4.5, 3.0), (0.0, 0.0, 0.0), 1.0, 0.05), 77),
]


@pytest.mark.parametrize("kind", ["sk", "ok"])
@pytest.mark.parametrize("cov,seed", CASES)
def test_sgs_matches_bruteforce(kind, cov, seed):
    g = Grid(9, 7, 4)
    data, mask = scattered(g.shape, 0.15, seed, law="normal")
    radii, max_n = (3, 3, 2), 6
    od, om = O.sgs_simulation((data, mask), g, None, radii, max_n, cov, seed, kriging_type=kind)
    bd, bm = brute_sgs(data, mask, g, cov, radii, max_n, seed, kind=kind)
    assert np.array_equal(om, bm)
    _close(od, bd)
    assert np.array_equal(od[mask == 1], data[mask == 1])
    if kind == "ok" and cov.type == 2:
        # the smooth Gaussian model gives large weights, so C0 - sum w b minus
        # the last weight goes negative: sqrt(var) is NaN and the NaN values
        # spread to later neighbours, in the reference as here
        assert np.isnan(od).sum() > g.size // 4
    else:
        assert np.isfinite(od).sum() > g.size // 2


@pytest.mark.parametrize("cov,seed", CASES)
def test_sgs_lvm_matches_bruteforce(cov, seed):
    g = Grid(8, 6, 4)
    data, mask = scattered(g.shape, 0.2, seed + 1, law="normal")
    rng = np.random.default_rng(seed)
    lvm = np.asfortranarray(rng.uniform(-0.5, 0.5, g.shape).astype(np.float32))
    radii, max_n = (3, 3, 2), 5
    od, om = O.sgs_simulation((data, mask), g, None, radii, max_n, cov, seed, mean=lvm)
    bd, bm = brute_sgs(data, mask, g, cov, radii, max_n, seed, lvm=lvm)
    assert np.array_equal(om, bm)
    _close(od, bd)


def test_sgs_simulation_mask_bruteforce():
    """Cells outside the simulation mask are neither simulated nor given a
    draw: the draws of the simulated cells follow the path order."""
    g = Grid(8, 7, 3)
    cov, seed = CASES[1]
    data, mask = scattered(g.shape, 0.1, 5, law="normal")
    sm = np.asfortranarray((np.random.default_rng(3).random(g.shape) < 0.6).astype(np.uint8))
    od, om = O.sgs_simulation((data, mask), g, None, (3, 3, 2), 6, cov, seed, kriging_type="ok", mask=sm)
    bd, bm = brute_sgs(data, mask, g, cov, (3, 3, 2), 6, seed, kind="ok", sim_mask=sm)
    assert np.array_equal(om, bm)
    _close(od, bd)
