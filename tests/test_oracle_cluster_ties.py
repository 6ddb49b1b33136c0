"""Pins the tie order and the partial overflow of the oracle's cell-binned
search (CPU only).

In the reference (sugarbox_indexed_neighbour_lookup.h:68-123,
clusterizer.cpp:24-52 and :147-184), informed cells are added to bins of
r_x x r_y x r_z cells in ascending cell order, and a bin that would receive
more than limit = template size / 250 cells is marked exceeded.  A target
any of whose 27 bins is exceeded takes the template path; otherwise the
candidates are gathered bin by bin in (k, j, i) order, ascending within a
bin, filtered by covariance > C(0) / 100 and sorted with std::sort by
covariance descending.  For at most 16 elements libstdc++'s std::sort is a
plain insertion sort, which is stable, so there the expected list is the
stable descending sort of the gathered candidates: the order of equal
covariances is the gather order.

The cases put the informed cells on a lattice, so that many candidates of
a target share a covariance (ties), plus a few random cells that overfill
some bins, so that one job mixes targets on both paths.  A gather in
another bin order, or an unstable sort, changes the tied runs; the tests
also check the planted variant with the bin loop in (i, j, k) order fails.
"""
import numpy as np
import pytest

import oracle_binding as O
from cases import Cov, Grid


def lattice_case(shape, spacing, extra, seed):
    X, Y, Z = shape
    mask = np.zeros(shape, np.uint8)
    mask[::spacing[0], ::spacing[1], ::spacing[2]] = 1
    rng = np.random.default_rng(seed)
    free = np.flatnonzero(mask.ravel(order="F") == 0)
    pick = rng.choice(free, extra, replace=False)
    flat = mask.ravel(order="F")
    flat[pick] = 1
    mask = flat.reshape(shape, order="F")
    data = np.where(mask == 1, rng.standard_normal(shape), -99.0).astype(np.float32)
    return np.asfortranarray(data), np.asfortranarray(mask)


def expected_lists(mask, grid, cov, radii, max_n, table, bin_order="kji"):
    """{target: list} for the cluster-path targets whose filtered candidate
    list has at most 16 entries, and the number of template-path targets."""
    X, Y, Z = grid.shape
    rx, ry, rz = radii
    limit = len(O.template(cov, radii, table=table)) // 250
    assert limit >= 1
    bx, by, bz = X // rx + 2, Y // ry + 2, Z // rz + 2
    bins = {}
    exceeded = set()
    flat = mask.ravel(order="F")
    for c in np.flatnonzero(flat):  # ascending cell order, as add_node is called
        x, y, z = c % X, c // X % Y, c // (X * Y)
        b = (x // rx, y // ry, z // rz)
        if b in exceeded:
            continue
        lst = bins.setdefault(b, [])
        if len(lst) >= limit:
            exceeded.add(b)
            del bins[b]
        else:
            lst.append(int(c))
    c0 = cov.sill - cov.nugget

    def key(dx, dy, dz):
        if table and (abs(dx) > rx or abs(dy) > ry or abs(dz) > rz):
            return 0.0
        return O.covariance(cov, float(dx), float(dy), float(dz))

    out = {}
    n_template = 0
    for t in np.flatnonzero(flat == 0):
        x, y, z = t % X, t // X % Y, t // (X * Y)
        cx, cy, cz = x // rx, y // ry, z // rz
        around = [(i, j, k) for k in (cz - 1, cz, cz + 1) for j in (cy - 1, cy, cy + 1) for i in (cx - 1, cx, cx + 1)
                  if 0 <= i < bx and 0 <= j < by and 0 <= k < bz]
        if any(b in exceeded for b in around):
            n_template += 1
            continue
        if bin_order == "ijk":
            around.sort(key=lambda b: (b[0], b[1], b[2]))
        cand = [c for b in around for c in bins.get(b, [])]
        keyed = [(key(x - c % X, y - c // X % Y, z - c // (X * Y)), c) for c in cand]
        keyed = [(k, c) for k, c in keyed if k > c0 / 100]
        if len(keyed) > 16:
            continue
        order = sorted(range(len(keyed)), key=lambda i: -keyed[i][0])  # Python's sort is stable
        out[int(t)] = [keyed[i][1] for i in order][:max_n]
    return out, n_template


def _ties(lst, cov, grid, t):
    X, Y, _ = grid.shape
    x, y, z = t % X, t // X % Y, t // (X * Y)
    keys = [O.covariance(cov, float(x - c % X), float(y - c // X % Y), float(z - c // (X * Y))) for c in lst]
    return any(keys[i] == keys[i + 1] for i in range(len(keys) - 1))


CASES = [
    # (kind, shape, lattice spacing, extra random cells, covariance, radii, max_neighbours)
    ("sk", (36, 30, 12), (5, 5, 3), 40, Cov(0, (9.0, 9.0, 4.0), (0.0, 0.0, 0.0), 1.0, 0.1), (6, 6, 3), 12),
    ("ok", (36, 30, 12), (4, 6, 3), 60, Cov(1, (8.0, 8.0, 3.0), (0.0, 0.0, 0.0), 1.0, 0.0), (6, 6, 3), 8),
    ("sk", (30, 30, 15), (3, 3, 5), 25, Cov(2, (7.0, 7.0, 5.0), (0.0, 0.0, 0.0), 1.0, 0.05), (5, 5, 4), 16),
]


def _run(kind, grid, data, mask, cov, radii, max_n):
    d = O.Dump(grid.size, max_n)
    try:
        if kind == "sk":
            O.simple_kriging((data, mask), grid, radii, max_n, cov, mean=0.0)
        else:
            O.ordinary_kriging((data, mask), grid, radii, max_n, cov)
    finally:
        d.close()
    return d


@pytest.mark.parametrize("case", CASES, ids=lambda c: "%s-%s" % (c[0], c[5]))
def test_cluster_path_tie_order_and_partial_overflow(case):
    kind, shape, spacing, extra, cov, radii, max_n = case
    g = Grid(*shape)
    data, mask = lattice_case(shape, spacing, extra, 7)
    d = _run(kind, g, data, mask, cov, radii, max_n)
    exp, n_template = expected_lists(mask, g, cov, radii, max_n, table=(kind == "sk"))
    assert n_template > 0 and len(exp) > 50, (n_template, len(exp))
    tied = 0
    for t, lst in exp.items():
        got = [int(c) for c in d.idx[t, :d.cnt[t]]]
        assert got == lst, "target %d: %s vs %s" % (t, got, lst)
        tied += _ties(lst, cov, g, t)
    assert tied > 20, "the case must exercise ties (%d tied lists)" % tied


def test_planted_bin_order_is_detected():
    """The same expectation with the bins gathered in (i, j, k) order must
    disagree with the oracle on tied lists."""
    kind, shape, spacing, extra, cov, radii, max_n = CASES[0]
    g = Grid(*shape)
    data, mask = lattice_case(shape, spacing, extra, 7)
    d = _run(kind, g, data, mask, cov, radii, max_n)
    exp, _ = expected_lists(mask, g, cov, radii, max_n, table=True, bin_order="ijk")
    wrong = sum([int(c) for c in d.idx[t, :d.cnt[t]]] != lst for t, lst in exp.items())
    assert wrong > 0
