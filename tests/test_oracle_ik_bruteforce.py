"""Pins the oracle's indicator kriging and median IK (CPU only).

No reference fixture exists for either.  The oracle's estimates are checked
against a brute-force restatement written from the reference's description
(indicator_kriging.h:58-124, cdf_utils.cpp:9-24, median_ik.cpp:15-93,
kriging_interpolation.h:225-263, combiner.h:10-24): for every cell and
category, the simple-kriging system of the cell's neighbours on the box
covariance table (0 beyond the radii), solved with numpy's LAPACK solver,
the indicator combine with the category's marginal as mean (in float for
IK, in double for median IK), the marginal itself when there is no
neighbour or the system is not positive definite, then the first most
probable category (IK) or prob <= 0.5 -> 0 (median IK).  The neighbour
lists are the oracle's own, taken from its neighbour dump: the search is
pinned separately (template brute force, top-max_n invariant of the
cell-binned search, tests/test_oracle.py).  Categories must agree cell for
cell; a swapped category index, a mean of the wrong category, a dropped
combine term or the IK rule applied to median IK changes many of them.
"""
import numpy as np
import pytest

import oracle_binding as O
from cases import Cov, Grid, categories, scattered

from test_oracle_sis import _table_cov


def _sk_weights(cells, node, grid, cv):
    X, Y, _ = grid.shape

    def loc(c):
        return c % X, c // X % Y, c // (X * Y)

    x, y, z = loc(node)
    pts = [loc(c) for c in cells]
    A = np.array([[cv(q[0] - p[0], q[1] - p[1], q[2] - p[2]) for q in pts] for p in pts])
    b = np.array([cv(x - p[0], y - p[1], z - p[2]) for p in pts])
    try:
        np.linalg.cholesky(A)
    except np.linalg.LinAlgError:
        return None
    return np.linalg.solve(A, b)


def brute_ik(cat, mask, grid, data, marginals, lists):
    """lists[k][node]: the neighbours of node for category k."""
    N = grid.size
    vals = np.asarray(cat, np.uint8).ravel(order="F")
    K = len(data)
    covs = [_table_cov(d["cov_model"], d["radiuses"]) for d in data]
    out = np.zeros(N, np.uint8)
    for node in range(N):
        probs = []
        for k in range(K):
            mk = np.float32(marginals[k])
            prob = mk
            nb = lists[k][node]
            if nb:
                w = _sk_weights(nb, node, grid, covs[k])
                if w is not None:
                    r = mk
                    for i, c in enumerate(nb):
                        ind = np.float32(1.0 if vals[c] == k else 0.0)
                        r = np.float32(r + np.float32(w[i] * float(ind - mk)))
                    prob = r
            probs.append(float(prob))
        best, arg = probs[0], 0
        for i in range(1, K):
            if probs[i] > best:
                best, arg = probs[i], i
        out[node] = arg
    return out.reshape(grid.shape, order="F")


def brute_median_ik(cat, mask, grid, cov, radii, marginals, lists):
    N = grid.size
    vals = np.asarray(cat, np.uint8).ravel(order="F")
    cv = _table_cov(cov, radii)
    m1 = np.float32(marginals[1])
    out = np.zeros(N, np.uint8)
    for node in range(N):
        prob = float(m1)
        nb = lists[node]
        if nb:
            w = _sk_weights(nb, node, grid, cv)
            if w is not None:
                prob = float(m1)
                for i, c in enumerate(nb):
                    ind = np.float32(1.0 if vals[c] == 1 else 0.0)
                    prob += float(w[i] * float(ind - m1))
        out[node] = 0 if prob <= 0.5 else 1
    return out.reshape(grid.shape, order="F")


def _ik_case(seed, K=4):
    g = Grid(14, 12, 6)
    data, mask = scattered(g.shape, 0.1, seed, law="normal")
    qs = tuple((i + 1) / K for i in range(K - 1))
    cat, marg = categories(data, mask, quantiles=qs, k=K)
    return g, cat, mask, marg


def _ik_params(K):
    models = [Cov(0, (6.0, 5.0, 3.0), (0.0, 0.0, 0.0), 1.0, 0.0),
              Cov(1, (5.0, 4.0, 2.5), (35.0, 0.0, 0.0), 1.0, 0.1),
              Cov(2, (4.0, 4.0, 3.0), (0.0, 0.0, 0.0), 1.0, 0.05),
              Cov(0, (7.0, 3.0, 2.0), (60.0, 0.0, 0.0), 1.0, 0.2)]
    return [{"cov_model": models[k % 4], "radiuses": (4 + k % 2, 4, 2), "max_neighbours": 6 + 2 * k}
            for k in range(K)]


@pytest.mark.parametrize("K,seed", [(3, 5), (4, 17), (5, 2027)])
def test_indicator_kriging_matches_bruteforce(K, seed):
    g, cat, mask, marg = _ik_case(seed, K)
    data = _ik_params(K)
    mx = max(d["max_neighbours"] for d in data)
    d = O.Dump(K * g.size, mx)
    try:
        od, om = O.indicator_kriging((cat, mask, K), g, data, marg)
    finally:
        d.close()
    lists = [[tuple(int(c) for c in d.idx[k * g.size + n, :max(d.cnt[k * g.size + n], 0)])
              for n in range(g.size)] for k in range(K)]
    assert all(d.cnt[k * g.size + n] >= 0 for k in range(K) for n in range(g.size))
    bd = brute_ik(cat, mask, g, data, marg, lists)
    assert np.all(om == 1)
    assert np.array_equal(od, bd), np.argwhere(od != bd)[:10]
    # informed cells keep their category (the cell is its own first neighbour)
    assert np.array_equal(od[mask == 1], cat[mask == 1])
    # every category is chosen somewhere
    assert len(np.unique(od)) == K


@pytest.mark.parametrize("seed,marg", [(8, (0.7, 0.3)), (31, (0.45, 0.55))])
def test_median_ik_matches_bruteforce(seed, marg):
    g = Grid(14, 12, 6)
    data, mask = scattered(g.shape, 0.12, seed, law="normal")
    cat = np.asfortranarray((data > np.quantile(data[mask == 1], marg[0])).astype(np.uint8) * mask)
    cov, radii, mx = Cov(1, (6.0, 5.0, 3.0), (20.0, 0.0, 0.0), 1.0, 0.05), (5, 4, 2), 10
    d = O.Dump(g.size, mx)
    try:
        od, om = O.median_ik((cat, mask, 2), g, marg, radii, mx, cov)
    finally:
        d.close()
    lists = [tuple(int(c) for c in d.idx[n, :max(d.cnt[n], 0)]) for n in range(g.size)]
    bd = brute_median_ik(cat, mask, g, cov, radii, marg, lists)
    assert np.all(om == 1)
    assert np.array_equal(od, bd), np.argwhere(od != bd)[:10]
    assert 0 < int(od.sum()) < g.size
