"""GPU parity: the HIP path through the C ABI against the CPU oracle.

Neighbour lists (members, order, tie-breaks) and masks must match bit for
bit; estimates within 1e-5 relative (scale floor: the data's spread) in the
mixed-precision mode; in FP64 mode, table-covariance jobs (SK, LVM, IK)
must match the oracle bit for bit because the FP64 kernel follows the
reference's dpotf2/dpotrs operation order.
"""
import ctypes as C
import json
import os

import numpy as np
import pytest

import oracle_binding as O
from cases import Cov, Grid, categories, rel_err, scattered, smooth_field

pytestmark = pytest.mark.gpu

TOL = 1e-5


@pytest.fixture(scope="module")
def geo():
    import hpgl_amd.geo_bsd as g

    return g


@pytest.fixture(autouse=True)
def _mixed(geo):
    geo.set_precision("mixed")
    yield
    geo.set_precision("mixed")


def _dump_product(n_cells, max_n):
    from hpgl_amd.geo_bsd import hpgl_wrap

    idx = np.full((n_cells, max_n), -1, np.int32)
    cnt = np.full(n_cells, -1, np.int32)
    hpgl_wrap._hpgl_so.hpgl_amd_set_neighbour_dump(idx.ctypes.data, cnt.ctypes.data, n_cells, max_n)
    return idx, cnt


def _same_lists(a_idx, a_cnt, b_idx, b_cnt):
    assert np.array_equal(a_cnt, b_cnt), np.flatnonzero(a_cnt != b_cnt)[:10]
    rows = np.flatnonzero(a_cnt > 0)
    for r in rows:
        c = a_cnt[r]
        if not np.array_equal(a_idx[r, :c], b_idx[r, :c]):
            raise AssertionError("cell %d: %s vs %s" % (r, a_idx[r, :c], b_idx[r, :c]))


def _ok_case(shape=(60, 50, 20), frac=0.02, seed=1001):
    data, mask = scattered(shape, frac, seed)
    return Grid(*shape), data, mask


def test_ordinary_kriging_template_path(geo):
    """OK, spherical with anisotropy; every node on the template path."""
    g, data, mask = _ok_case()
    cov = Cov(0, (10, 8, 3), sill=1.0, nugget=0.1)
    r, mx = (10, 10, 4), 12
    pidx, pcnt = _dump_product(g.size, mx)
    out = geo.ordinary_kriging((data, mask), g, r, mx, cov)
    st = geo.last_stats()
    d = O.Dump(g.size, mx)
    try:
        od, om = O.ordinary_kriging((data, mask), g, r, mx, cov)
    finally:
        d.close()
    _same_lists(pidx, pcnt, d.idx, d.cnt)
    assert st["nodes_template_path"] > 0
    assert np.array_equal(out.mask, om)
    m = om == 1
    assert rel_err(out.data[m], od[m], 1.0) <= TOL


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_simple_kriging_cluster_path(geo, precision):
    """SK on a dense grid with large radii: the cell-binned path with the
    exact std::sort tie order; FP64 mode is bit-exact."""
    geo.set_precision(precision)
    shape = (48, 40, 16)
    data, mask = scattered(shape, 0.03, 7, law="normal")
    g = Grid(*shape)
    cov = Cov(1, (12, 12, 4), sill=1.0, nugget=0.05)
    r, mx = (12, 12, 4), 24
    pidx, pcnt = _dump_product(g.size, mx)
    out = geo.simple_kriging((data, mask), g, r, mx, cov, mean=0.0)
    st = geo.last_stats()
    d = O.Dump(g.size, mx)
    try:
        od, om = O.simple_kriging((data, mask), g, r, mx, cov, mean=0.0)
    finally:
        d.close()
    assert st["nodes_cluster_path"] > 0
    _same_lists(pidx, pcnt, d.idx, d.cnt)
    assert np.array_equal(out.mask, om)
    if precision == "fp64":
        assert np.array_equal(out.data, od)
    else:
        assert rel_err(out.data, od, 1.0) <= TOL


def _two_density(shape, split_x, dense, sparse, seed):
    """Normal values, informed with probability `dense` for x < split_x and
    `sparse` beyond."""
    rng = np.random.default_rng(seed)
    p = np.full(shape, sparse)
    p[:split_x] = dense
    mask = (rng.random(shape) < p).astype(np.uint8)
    data = np.where(mask == 1, rng.standard_normal(shape), -99.0).astype(np.float32)
    return np.asfortranarray(data), np.asfortranarray(mask)


@pytest.mark.parametrize("angle", [0.0, 30.0])
def test_cluster_path_long_candidate_lists(geo, angle):
    """Dense data in a third of the grid and sparse data elsewhere: the
    list capacity is chosen from the mean density, so the lists of the dense
    part exceed it and are sorted in global scratch; with and without
    rotation of the covariance model."""
    shape = (180, 90, 24)
    data, mask = _two_density(shape, 60, 0.03, 0.001, 11)
    g = Grid(*shape)
    cov = Cov(1, (30, 30, 8), angles=(angle, 0.0, 0.0), sill=1.0, nugget=0.05)
    r, mx = (30, 30, 8), 24
    pidx, pcnt = _dump_product(g.size, mx)
    out = geo.simple_kriging((data, mask), g, r, mx, cov, mean=0.0)
    st = geo.last_stats()
    d = O.Dump(g.size, mx)
    try:
        od, om = O.simple_kriging((data, mask), g, r, mx, cov, mean=0.0)
    finally:
        d.close()
    assert st["nodes_cluster_path"] > 0
    assert st["cluster_lists_spilled"] > 1000
    _same_lists(pidx, pcnt, d.idx, d.cnt)
    assert np.array_equal(out.mask, om)
    assert rel_err(out.data, od, 1.0) <= TOL


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_ordinary_kriging_cluster_path_analytic_keys(geo, precision):
    """OK's cluster path orders by the analytic covariance, including
    candidates beyond the search radius.  Gaussian covariance without a
    nugget gives ill-conditioned systems: the dpotrf decisions (masks) must
    still equal the oracle's in both modes, FP64 values bit for bit."""
    geo.set_precision(precision)
    shape = (40, 36, 12)
    data, mask = scattered(shape, 0.015, 3)  # ~40% of nodes on each path
    g = Grid(*shape)
    cov = Cov(2, (9, 9, 3), angles=(25, 0, 0), sill=1.0, nugget=0.0)
    r, mx = (9, 9, 3), 16
    pidx, pcnt = _dump_product(g.size, mx)
    out = geo.ordinary_kriging((data, mask), g, r, mx, cov)
    st = geo.last_stats()
    d = O.Dump(g.size, mx)
    try:
        od, om = O.ordinary_kriging((data, mask), g, r, mx, cov)
    finally:
        d.close()
    none, sing = O.last_counts()
    assert st["nodes_cluster_path"] > 0 and st["nodes_template_path"] > 0
    _same_lists(pidx, pcnt, d.idx, d.cnt)
    assert st["nodes_singular"] == sing and st["nodes_no_neighbours"] == none
    diff = np.flatnonzero(out.mask.ravel(order="F") != om.ravel(order="F"))
    assert diff.size == 0, "%d masks differ, first cells %s" % (diff.size, diff[:10])
    m = (mask == 0) & (om == 1)
    if precision == "fp64":
        bad = np.flatnonzero(out.data[m] != od[m])
        assert bad.size == 0, "%d FP64 values differ" % bad.size
    else:
        assert rel_err(out.data[m], od[m], 1.0) <= TOL


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_lvm_kriging(geo, precision):
    geo.set_precision(precision)
    shape = (50, 44, 10)
    data, mask = scattered(shape, 0.02, 11, law="normal")
    mean = smooth_field(shape, 8.0, 12) * 0.5
    g = Grid(*shape)
    cov = Cov(1, (10, 10, 3))
    r, mx = (10, 10, 3), 16
    out = geo.lvm_kriging((data, mask), g, mean, r, mx, cov)
    od, om = O.lvm_kriging((data, mask), g, mean, r, mx, cov)
    assert np.array_equal(out.mask, om)
    if precision == "fp64":
        assert np.array_equal(out.data, od)
    else:
        assert rel_err(out.data, od, 1.0) <= TOL


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_indicator_kriging_five_categories(geo, precision):
    geo.set_precision(precision)
    shape = (40, 40, 12)
    data, mask = scattered(shape, 0.03, 21, law="normal")
    cat, marg = categories(data, mask)
    g = Grid(*shape)
    params = [{"cov_model": Cov(0, (6 + k, 6 + k, 2)), "radiuses": (6 + k, 6 + k, 2), "max_neighbours": 12}
              for k in range(5)]
    pidx, pcnt = _dump_product(5 * g.size, 12)
    out = geo.indicator_kriging((cat, mask, 5), g, [dict(p) for p in params], marg)
    d = O.Dump(5 * g.size, 12)
    try:
        od, om = O.indicator_kriging((cat, mask, 5), g, [dict(p) for p in params], marg)
    finally:
        d.close()
    _same_lists(pidx, pcnt, d.idx, d.cnt)
    assert np.array_equal(out.mask, om)
    flips = int(np.sum(out.data != od))
    if precision == "fp64":
        assert flips == 0, "%d categories differ in FP64 mode" % flips
    else:
        # argmax over probabilities that carry the refinement's ~1e-7
        # difference: only a near-tie can flip (none measured)
        assert flips <= 2, "%d of %d categories differ in mixed mode" % (flips, out.data.size)


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_median_ik(geo, precision):
    geo.set_precision(precision)
    shape = (44, 40, 10)
    data, mask = scattered(shape, 0.03, 31, law="normal")
    ind = (data > 0).astype(np.uint8) * mask
    g = Grid(*shape)
    cov = Cov(0, (9, 9, 3))
    p1 = float(np.mean(ind[mask == 1]))
    out = geo.median_ik((ind, mask, 2), g, (1 - p1, p1), (9, 9, 3), 16, cov)
    od, om = O.median_ik((ind, mask, 2), g, (1 - p1, p1), (9, 9, 3), 16, cov)
    assert np.array_equal(out.mask, om)
    flips = int(np.sum(out.data != od))
    if precision == "fp64":
        assert flips == 0, "%d median IK indicators differ in FP64 mode" % flips
    else:
        # a flip needs a probability within the refinement's ~1e-7 of 0.5
        assert flips <= 2, "%d of %d median IK indicators differ in mixed mode" % (flips, out.data.size)


def test_failure_policies(geo):
    """No neighbours: OK leaves the cell as cloned from the input, SK writes the mean."""
    data = np.full((40, 1, 1), -99, np.float32)
    mask = np.zeros((40, 1, 1), np.uint8)
    data[0], mask[0] = 3.0, 1
    g = Grid(40, 1, 1)
    out = geo.ordinary_kriging((data, mask), g, (2, 1, 1), 4, Cov(0, (2, 1, 1)))
    assert out.mask[39, 0, 0] == 0 and out.data[39, 0, 0] == -99 and out.mask[1, 0, 0] == 1
    out = geo.simple_kriging((data, mask), g, (2, 1, 1), 4, Cov(0, (2, 1, 1)), mean=1.5)
    assert out.mask[39, 0, 0] == 1 and out.data[39, 0, 0] == np.float32(1.5)


def test_problem_7_2_simple_kriging_on_gpu(geo):
    """The reference's golden SK vector through the GPU path (19/20 at print
    precision; node 10 is the std::sort tie witness, equal to the oracle)."""
    here = os.path.dirname(os.path.abspath(__file__))
    d = json.load(open(os.path.join(here, "golden", "problem_7_2.json")))
    data = np.full((20, 1, 1), -99, np.float32)
    mask = np.zeros((20, 1, 1), np.uint8)
    for c, v in zip(d["cells"], d["values"]):
        data[c], mask[c] = v, 1
    out = geo.simple_kriging((data, mask), Grid(20, 1, 1), (5, 1, 1), 2, Cov(1, (5, 1, 1)))
    exp = np.array(d["expected_sk"], np.float32)
    diff = np.abs(out.data.ravel() - exp)
    assert np.all(np.delete(diff, 10) <= 5e-8)
    od, _ = O.simple_kriging((data, mask), Grid(20, 1, 1), (5, 1, 1), 2, Cov(1, (5, 1, 1)))
    assert out.data.ravel()[10] == od.ravel()[10]


def test_simple_kriging_weights_parta(geo):
    here = os.path.dirname(os.path.abspath(__file__))
    d = json.load(open(os.path.join(here, "golden", "gstl_parta.json")))
    pts = np.array(d["points"], float)
    w = geo.simple_kriging_weights([41, 29, 0], pts[:, 0], pts[:, 1], np.zeros(8), ranges=(10, 10, 10), sill=1,
                                   cov_type=geo.covariance.exponential)
    rc, wo, _ = O.simple_kriging_weights([41, 29, 0], pts[:, 0], pts[:, 1], np.zeros(8), Cov(1, (10, 10, 10)))
    assert rc == 0
    assert np.array_equal(w, wo)
    assert np.abs(w - np.array(d["sk_weights"])).max() < 1e-4


def test_mixed_vs_fp64_error_report(geo):
    """Mixed precision stays within 1e-6 relative of the FP64 path on a
    well-conditioned job and reports how many systems went to FP64."""
    shape = (64, 64, 16)
    data, mask = scattered(shape, 0.02, 41, law="normal")
    g = Grid(*shape)
    cov = Cov(1, (12, 12, 4), nugget=0.05)
    geo.set_precision("fp64")
    ref = geo.simple_kriging((data, mask), g, (12, 12, 4), 24, cov, mean=0.0)
    s64 = geo.last_stats()
    geo.set_precision("mixed")
    mix = geo.simple_kriging((data, mask), g, (12, 12, 4), 24, cov, mean=0.0)
    smx = geo.last_stats()
    assert s64["solves_fp32"] == 0 and s64["solves_fp64"] > 0
    assert smx["solves_fp32"] + smx["solves_fp64"] == s64["solves_fp64"]
    assert rel_err(mix.data, ref.data, 1.0) <= 1e-6


@pytest.mark.parametrize("max_n,frac", [(7, 0.01), (19, 0.01), (32, 0.01), (40, 0.04), (64, 0.08)])
def test_uneven_neighbour_counts_buckets(geo, max_n, frac):
    """Neighbour counts across the solve buckets (4..32 on four lanes per
    system, 33..64 on the FP64 kernel with rows in LDS); mixed within 1e-5
    of the oracle, FP64 bit-exact."""
    shape = (36, 30, 10)
    data, mask = scattered(shape, frac, 50 + max_n, law="normal")
    g = Grid(*shape)
    cov = Cov(0, (8, 8, 3), nugget=0.2)
    r = (8, 8, 3)
    od, om = O.simple_kriging((data, mask), g, r, max_n, cov, mean=0.0)
    for precision in ("mixed", "fp64"):
        geo.set_precision(precision)
        pidx, pcnt = _dump_product(g.size, max_n)
        out = geo.simple_kriging((data, mask), g, r, max_n, cov, mean=0.0)
        counts = pcnt[pcnt >= 0]
        assert len(np.unique(counts)) > 3
        if max_n > 32:
            assert np.sum(counts > 32) > 100
        assert np.array_equal(out.mask, om)
        if precision == "fp64":
            assert np.array_equal(out.data, od)
        else:
            assert rel_err(out.data, od, 1.0) <= TOL


def test_ordinary_kriging_wide_systems_fp64_bit_exact(geo):
    """OK with 33..64 neighbours on the analytic model: the covariances come
    from the host-built pair table, so FP64 mode matches the oracle bit for
    bit and mixed mode within 1e-5."""
    shape = (40, 36, 12)
    data, mask = scattered(shape, 0.06, 77)
    g = Grid(*shape)
    cov = Cov(1, (10, 9, 4), angles=(20.0, 0.0, 0.0), sill=1.0, nugget=0.1)
    r, mx = (9, 9, 4), 64
    od, om = O.ordinary_kriging((data, mask), g, r, mx, cov)
    for precision in ("fp64", "mixed"):
        geo.set_precision(precision)
        out = geo.ordinary_kriging((data, mask), g, r, mx, cov)
        assert np.array_equal(out.mask, om)
        if precision == "fp64":
            assert np.array_equal(out.data, od)
        else:
            assert rel_err(out.data[om == 1], od[om == 1], 1.0) <= TOL


@pytest.mark.parametrize("max_n", [24, 64])
def test_ordinary_kriging_device_evaluated_covariance(geo, max_n):
    """OK with radii 42: the pair table over the offsets neighbour pairs can
    take would exceed 2^25 entries, so the FP64 kernels evaluate the model
    on the device (cov_eval) while building each system, for 17..32 and
    33..64 neighbours.  The neighbour lists come from the host-built
    template and rank tables as always, so the masks equal the oracle's;
    the device's exp and square root may differ from the host's in the last
    bit, so values agree within 1e-5 (parity of the device evaluation is
    within tolerance, not bit-exact)."""
    shape = (40, 36, 12)
    data, mask = scattered(shape, 0.05, 91)
    g = Grid(*shape)
    r = (42, 42, 42)
    assert (2 * (4 * r[0] - 2) + 1) ** 3 > 1 << 25
    od, om = O.ordinary_kriging((data, mask), g, r, max_n, Cov(1, (30, 25, 8), angles=(30.0, 0.0, 0.0), nugget=0.1))
    for precision in ("fp64", "mixed"):
        geo.set_precision(precision)
        out = geo.ordinary_kriging((data, mask), g, r, max_n, Cov(1, (30, 25, 8), angles=(30.0, 0.0, 0.0),
                                                                   nugget=0.1))
        st = geo.last_stats()
        assert st["solves_fp64"] > 0 and st["solves_fp32"] == 0
        assert np.array_equal(out.mask, om)
        assert rel_err(out.data[om == 1], od[om == 1], 1.0) <= TOL
    geo.set_precision("mixed")


@pytest.mark.parametrize("case_idx", [0, 1, 2])
def test_cell_binned_ties_and_partial_overflow(geo, case_idx):
    """Lattice data with heavy covariance ties and a few overfull bins, so a
    job mixes both search paths: every neighbour list (order of tied
    candidates included) and every mask equal the oracle's, whose tie order
    is pinned by tests/test_oracle_cluster_ties.py."""
    import test_oracle_cluster_ties as T

    kind, shape, spacing, extra, cov, radii, mx = T.CASES[case_idx]
    g = Grid(*shape)
    data, mask = T.lattice_case(shape, spacing, extra, 7)
    pidx, pcnt = _dump_product(g.size, mx)
    if kind == "sk":
        out = geo.simple_kriging((data, mask), g, radii, mx, cov, mean=0.0)
    else:
        out = geo.ordinary_kriging((data, mask), g, radii, mx, cov)
    st = geo.last_stats()
    d = T._run(kind, g, data, mask, cov, radii, mx)
    assert st["nodes_cluster_path"] > 0 and st["nodes_template_path"] > 0
    _same_lists(pidx, pcnt, d.idx, d.cnt)
    if kind == "sk":
        od, om = O.simple_kriging((data, mask), g, radii, mx, cov, mean=0.0)
    else:
        od, om = O.ordinary_kriging((data, mask), g, radii, mx, cov)
    assert np.array_equal(out.mask, om)
    m = om == 1
    assert rel_err(out.data[m], od[m], 1.0) <= TOL
