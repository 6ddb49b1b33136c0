"""Pins the oracle's locally varying mean (LVM) kriging (CPU only).

No reference fixture exists for LVM kriging.  The oracle is checked against
a brute-force restatement written from the reference's description
(lvm_kriging.cpp:16-37 with cont_kriging.h:39-121, kriging_interpolation.h:
228-265 and combiner.h:10-24): for every uninformed cell, the simple-kriging
system of its neighbours on the box covariance table (0 beyond the radii),
solved with numpy's LAPACK solver; the estimate starts from the mean at the
cell, mean[cell], and adds w_i (v_i - mean[n_i]) with each neighbour's own
mean (the combiner selects the means at the neighbours' indices), the
difference in float, the product in double, the sum in float; a cell
without neighbours or with a system that is not positive definite gets
mean[cell], with its mask set.  Informed cells keep their value.  The
neighbour lists are the oracle's own (the search is pinned separately).

A planted mistake shows the check is sharp: the same restatement with the
centre's mean used for every neighbour must disagree with the oracle.
"""
import numpy as np
import pytest

import oracle_binding as O
from cases import Cov, Grid, scattered, smooth_field

from test_oracle_ik_bruteforce import _sk_weights
from test_oracle_sis import _table_cov


def brute_lvm(data, mask, mean, grid, cov, radii, lists, centre_mean_everywhere=False):
    N = grid.size
    v = np.asarray(data, np.float32).ravel(order="F")
    m = np.asarray(mask, np.uint8).ravel(order="F")
    mu = np.asarray(mean, np.float32).ravel(order="F")
    cv = _table_cov(cov, radii)
    out = v.copy()
    out_mask = m.copy()
    for node in range(N):
        if m[node]:
            continue
        r = np.float32(mu[node])
        nb = lists[node]
        if nb:
            w = _sk_weights(nb, node, grid, cv)
            if w is not None:
                for i, c in enumerate(nb):
                    mi = mu[node] if centre_mean_everywhere else mu[c]
                    r = np.float32(r + np.float32(w[i] * float(np.float32(v[c] - mi))))
        out[node] = r
        out_mask[node] = 1
    return out.reshape(grid.shape, order="F"), out_mask.reshape(grid.shape, order="F")


def _case(seed, shape=(16, 14, 6), frac=0.08):
    g = Grid(*shape)
    data, mask = scattered(g.shape, frac, seed, law="normal")
    mean = smooth_field(g.shape, 4.0, seed + 1) * 0.8 + 0.3
    return g, data, mask, mean


@pytest.mark.parametrize("seed,cov,radii,mx", [
    (3, Cov(1, (6.0, 5.0, 3.0), (0.0, 0.0, 0.0), 1.0, 0.1), (5, 4, 2), 8),
    (19, Cov(0, (7.0, 4.0, 3.0), (30.0, 0.0, 0.0), 1.0, 0.0), (6, 4, 3), 12),
    (40, Cov(2, (4.0, 4.0, 2.0), (0.0, 0.0, 0.0), 1.0, 0.05), (4, 4, 2), 6),
])
def test_lvm_kriging_matches_bruteforce(seed, cov, radii, mx):
    g, data, mask, mean = _case(seed)
    d = O.Dump(g.size, mx)
    try:
        od, om = O.lvm_kriging((data, mask), g, mean, radii, mx, cov)
    finally:
        d.close()
    lists = [tuple(int(c) for c in d.idx[n, :max(d.cnt[n], 0)]) for n in range(g.size)]
    bd, bm = brute_lvm(data, mask, mean, g, cov, radii, lists)
    assert np.array_equal(om, bm)
    # the brute force solves by LU, the oracle by Cholesky: the weights, and
    # so the float sums, may differ in the last bits
    scale = np.maximum(np.abs(bd), 1.0)
    assert np.max(np.abs(od - bd) / scale) <= 2e-6
    assert np.array_equal(od[mask == 1], data[mask == 1])
    # some cells have no neighbours or a failing system: they take their own mean
    none, sing = O.last_counts()
    lone = [n for n in range(g.size) if mask.ravel(order="F")[n] == 0 and not lists[n]]
    assert none == len(lone)
    for n in lone:
        assert od.ravel(order="F")[n] == np.float32(mean.ravel(order="F")[n])


def test_lvm_planted_mistake_is_detected():
    """The restatement with the centre's mean at every neighbour (the mistake
    a port of combiner.h could make) must not match the oracle."""
    cov, radii, mx = Cov(1, (6.0, 5.0, 3.0), (0.0, 0.0, 0.0), 1.0, 0.1), (5, 4, 2), 8
    g, data, mask, mean = _case(3)
    d = O.Dump(g.size, mx)
    try:
        od, om = O.lvm_kriging((data, mask), g, mean, radii, mx, cov)
    finally:
        d.close()
    lists = [tuple(int(c) for c in d.idx[n, :max(d.cnt[n], 0)]) for n in range(g.size)]
    bad, _ = brute_lvm(data, mask, mean, g, cov, radii, lists, centre_mean_everywhere=True)
    scale = np.maximum(np.abs(od), 1.0)
    assert np.max(np.abs(od - bad) / scale) > 1e-2
