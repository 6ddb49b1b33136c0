"""Pins the oracle's sequential indicator simulation (CPU only).

The oracle's sis_core is checked against an independent brute-force
restatement written here from the reference's description
(sequential_indicator_simulation.cpp:33-124, sample.cpp:16-29,
kriging_interpolation.h:239-272, my_kriging_weights.h:258-362): a Python
loop over the path that searches the template cell by cell, solves each
category's simple-kriging (or correlogram) system with numpy's LAPACK
solver, forms the probabilities and samples with the mt11213b draw of the
simulated cell.  Its inputs are separately pinned pieces: the template
order, the path LCG and mt11213b (tests/test_oracle.py), and the analytic
covariance.  Categories must agree cell for cell; a swapped sign, a
dropped mean term, a wrong draw index or a wrong median-SIS complement
changes many of them.
"""
import numpy as np
import pytest

import oracle_binding as O
from cases import Cov, Grid


def _table_cov(cm, radii):
    """Covariance of the box table: the analytic model inside +-radii, 0 outside."""
    cache = {}

    def c(dx, dy, dz):
        if abs(dx) > radii[0] or abs(dy) > radii[1] or abs(dz) > radii[2]:
            return 0.0
        key = (dx, dy, dz)
        if key not in cache:
            cache[key] = O.covariance(cm, float(dx), float(dy), float(dz))
        return cache[key]

    return c


def _sigma(m):
    m = float(m)
    if m == 0:
        m += 0.00001
    if m == 1:
        m -= 0.00001
    return np.sqrt(m * (1 - m))


def brute_sis(cat, mask, grid, data, seed, marginal_probs, lvm=None, correlogram=False, sim_mask=None):
    X, Y, Z = grid.shape
    N = grid.size
    K = len(data)
    vals = np.asarray(cat, np.uint8).ravel(order="F").copy()
    inf = np.asarray(mask, np.uint8).ravel(order="F").copy()
    sim = None if sim_mask is None else np.asarray(sim_mask, np.uint8).ravel(order="F")
    means = None if lvm is None else [np.asarray(m, np.float32).ravel(order="F") for m in lvm]
    tmpl = [O.template(d["cov_model"], d["radiuses"], True) for d in data]
    covs = [_table_cov(d["cov_model"], d["radiuses"]) for d in data]
    path = O.sgs_path(N, seed, N)
    draws = O.mt11213b(seed & 0xFFFFFFFF, N)
    kriged = 1 if K == 2 else K
    used = 0
    for node in path:
        if inf[node] or (sim is not None and sim[node] != 1):
            continue
        x, y, z = node % X, node // X % Y, node // (X * Y)
        probs = []
        for c in range(kriged):
            mc = float(np.float32(means[c][node] if means is not None else marginal_probs[c]))
            nb = []
            for dx, dy, dz in tmpl[c]:
                if len(nb) >= data[c]["max_neighbours"]:
                    break
                xx, yy, zz = x + dx, y + dy, z + dz
                if 0 <= xx < X and 0 <= yy < Y and 0 <= zz < Z:
                    cell = xx + X * (yy + Y * zz)
                    if inf[cell]:
                        nb.append((cell, xx, yy, zz))
            prob = mc
            if nb:
                n = len(nb)
                cv = covs[c]
                A = np.array([[cv(q[1] - p[1], q[2] - p[2], q[3] - p[3]) for q in nb] for p in nb])
                b = np.array([cv(x - p[1], y - p[2], z - p[3]) for p in nb])
                mi = [float(np.float32(means[c][p[0]])) if means is not None else mc for p in nb]
                if correlogram:
                    s = np.array([_sigma(m) for m in mi])
                    A = A * np.outer(s, s)
                    b = b * s * _sigma(mc)
                try:
                    np.linalg.cholesky(A)
                    w = np.linalg.solve(A, b)
                    r = mc
                    for i, p in enumerate(nb):
                        ind = 1.0 if vals[p[0]] == c else 0.0
                        r += w[i] * float(np.float32(ind) - np.float32(mi[i]))
                    prob = r
                except np.linalg.LinAlgError:
                    pass
            probs.append(np.float32(prob))
            if K == 2:
                probs.append(np.float32(1 - prob))
        u = float(draws[used]) / 4294967295.0
        used += 1
        s = u * sum(float(p) for p in probs)
        acc, pick = 0.0, len(probs) - 1
        for i, p in enumerate(probs):
            acc += float(p)
            if s < acc:
                pick = i
                break
        vals[node] = pick
        inf[node] = 1
    return vals.reshape(grid.shape, order="F"), inf.reshape(grid.shape, order="F")


def _hard(grid, K, frac, seed):
    rng = np.random.default_rng(seed)
    n = grid.size
    idx = rng.choice(n, max(1, int(frac * n)), replace=False)
    mask = np.zeros(n, np.uint8)
    mask[idx] = 1
    cat = np.zeros(n, np.uint8)
    cat[idx] = rng.integers(0, K, idx.size)
    return cat.reshape(grid.shape, order="F"), mask.reshape(grid.shape, order="F")


def _params(K, max_n=6, radii=(3, 3, 2)):
    out = []
    for k in range(K):
        cm = Cov(1 if k % 2 == 0 else 0, (4.0 + k, 3.0, 2.5), (10.0 * k, 0.0, 0.0), 1.0, 0.1 * (k % 2))
        out.append({"cov_model": cm, "radiuses": radii, "max_neighbours": max_n + k})
    return out


@pytest.mark.parametrize("K,seed", [(3, 12345), (3, 7), (4, 2024), (2, 99), (2, 31337)])
def test_sis_matches_bruteforce(K, seed):
    g = Grid(9, 7, 4)
    cat, mask = _hard(g, K, 0.15, seed)
    data = _params(K)
    marg = [0.2, 0.3, 0.5] if K == 3 else ([0.4, 0.6] if K == 2 else [0.1, 0.2, 0.3, 0.4])
    od, om = O.sis_simulation((cat, mask, K), g, data, seed, marg)
    bd, bm = brute_sis(cat, mask, g, data, seed, marg)
    assert np.array_equal(om, bm)
    assert np.array_equal(od, bd), np.argwhere(od != bd)[:10]
    # hard data are kept
    assert np.array_equal(od[mask == 1], cat[mask == 1])


@pytest.mark.parametrize("correlogram", [False, True])
@pytest.mark.parametrize("K", [2, 3])
def test_sis_lvm_matches_bruteforce(K, correlogram):
    g = Grid(8, 6, 4)
    seed = 4242 + K
    cat, mask = _hard(g, K, 0.2, seed)
    rng = np.random.default_rng(seed)
    raw = rng.uniform(0.05, 1.0, (K,) + g.shape)
    lvm = [np.asfortranarray((raw[k] / raw.sum(axis=0)).astype(np.float32)) for k in range(K)]
    lvm[0][0, 0, 0] = 0.0  # the sigma guards at 0 and 1
    lvm[-1][1, 0, 0] = 1.0
    data = _params(K, max_n=5)
    od, om = O.sis_simulation((cat, mask, K), g, data, seed, lvm, use_correlogram=correlogram)
    bd, bm = brute_sis(cat, mask, g, data, seed, None, lvm=lvm, correlogram=correlogram)
    assert np.array_equal(om, bm)
    assert np.array_equal(od, bd), np.argwhere(od != bd)[:10]


def test_sis_simulation_mask_and_marginals_only():
    """No neighbours (max_neighbours 0): every category is its marginal, so
    cell i of the path takes the category whose cumulative marginal first
    exceeds u_i * sum; cells outside the simulation mask keep mask 0."""
    g = Grid(6, 5, 3)
    K = 3
    cat, mask = _hard(g, K, 0.1, 5)
    sim = np.zeros(g.shape, np.uint8, order="F")
    sim[::2, :, :] = 1
    data = _params(K, max_n=0)
    for d in data:
        d["max_neighbours"] = 0
    marg = [0.25, 0.25, 0.5]
    seed = 77
    od, om = O.sis_simulation((cat, mask, K), g, data, seed, marg, mask=sim)
    flat_m = mask.ravel(order="F")
    flat_s = sim.ravel(order="F")
    path = O.sgs_path(g.size, seed, g.size)
    draws = O.mt11213b(seed, g.size)
    p32 = np.array(marg, np.float32).astype(np.float64)
    cum = np.cumsum(p32)
    want = cat.ravel(order="F").copy()
    used = 0
    for node in path:
        if flat_m[node] or flat_s[node] != 1:
            continue
        s = draws[used] / 4294967295.0 * p32.sum()
        used += 1
        want[node] = int(np.argmax(s < cum)) if s < cum[-1] else K - 1
    assert np.array_equal(od.ravel(order="F"), want)
    assert np.array_equal(om.ravel(order="F"), ((flat_m == 1) | (flat_s == 1)).astype(np.uint8))
