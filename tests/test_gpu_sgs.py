"""GPU parity of sequential Gaussian simulation against the CPU oracle.

Per realisation, the neighbour list of every simulated cell (members and
order) must equal the oracle's bit for bit: it depends only on the random
path, the conditioning mask and the template, so any error in the path
LCG, the rank keys or the template scan shows here.  Values within 1e-5
relative (scale floor 1, the normal-score spread) in both precision modes;
in the mixed mode the weights differ from FP64 by the refinement residual
and the difference travels along the path to later cells.
"""
import json
import os

import numpy as np
import pytest

import oracle_binding as O
from cases import Cov, Grid, rel_err, scattered, smooth_field

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
    for r in np.flatnonzero(a_cnt > 0):
        c = a_cnt[r]
        if not np.array_equal(a_idx[r, :c], b_idx[r, :c]):
            raise AssertionError("cell %d: %s vs %s" % (r, a_idx[r, :c], b_idx[r, :c]))


def _normal_scores(shape, frac, seed):
    data, mask = scattered(shape, frac, seed, law="normal")
    return data, mask


def _cdf_of(data, mask):
    v = np.unique(data[mask == 1])
    return v[:-1].astype(np.float32), (np.arange(1, v.size) / float(v.size)).astype(np.float32)


def _run_pair(geo, prop, grid, cdf, radii, max_n, cov, seed, kind="sk", mean=None, mask=None):
    """Product and oracle on the same inputs; returns both results and dumps."""
    n = grid.size
    pidx, pcnt = _dump_product(n, max_n)
    cdf_arg = geo.CdfData(*cdf) if cdf is not None else None
    out = geo.sgs_simulation(prop, grid, cdf_arg, radii, max_n, cov, seed, kriging_type=kind, mean=mean, mask=mask)
    d = O.Dump(n, max_n)
    try:
        od, om = O.sgs_simulation(prop, grid, cdf, radii, max_n, geo.normed_cov_model(cov), seed, kind, mean, mask)
    finally:
        d.close()
    return out, (od, om), (pidx, pcnt), (d.idx, d.cnt)


def test_sgs_problem_7_2_golden_on_gpu(geo):
    """The reference's unconditional SGS vector of problem 7.2 through the
    GPU path (LCG path, mt11213b draws, SK on the box table, Odeh-Evans
    inverse, non-parametric back-transform): all 20 values within 2e-7."""
    here = os.path.dirname(os.path.abspath(__file__))
    d = json.load(open(os.path.join(here, "golden", "problem_7_2.json")))
    data = np.full(20, -99, np.float32)
    mask = np.zeros(20, np.uint8)
    for c, v in zip(d["cells"], d["values"]):
        data[c], mask[c] = v, 1
    cdf = _cdf_of(data, mask)
    s = d["sgs"]
    prop = (np.full((20, 1, 1), -99, np.float32), np.zeros((20, 1, 1), np.uint8))
    out = geo.sgs_simulation(prop, Grid(20, 1, 1), geo.CdfData(*cdf), s["radiuses"], s["max_neighbours"],
                             Cov(s["cov_type"], s["ranges"]), s["seed"])
    assert np.all(out.mask == 1)
    assert np.abs(out.data.ravel() - np.array(d["expected_sgs"], np.float32)).max() <= 2e-7


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_sgs_simple_kriging_conditional(geo, precision):
    """SK SGS with a normal-score transform, 2% conditioning data."""
    geo.set_precision(precision)
    shape = (40, 30, 12)
    g = Grid(*shape)
    data, mask = _normal_scores(shape, 0.02, 2024)
    cov = Cov(0, (12, 9, 4), sill=1.0, nugget=0.0)
    out, (od, om), pd, od_ = _run_pair(geo, (data, mask), g, _cdf_of(data, mask), (12, 9, 4), 16, cov, 777)
    _same_lists(*pd, *od_)
    assert np.array_equal(out.mask, om)
    assert np.all(om == 1)
    assert rel_err(out.data, od, 1.0) <= TOL


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_sgs_wide_neighbourhoods(geo, precision):
    """SK SGS with up to 48 neighbours: systems of 33..48 unknowns take the
    FP64 block-cyclic kernel's SGS epilogue (weights and variance for the
    ordered sweep), and the simulated-neighbour bits reach past bit 31 in
    both template-scan passes.  Lists bit-exact, values within 1e-5."""
    geo.set_precision(precision)
    shape = (30, 26, 10)
    g = Grid(*shape)
    data, mask = _normal_scores(shape, 0.02, 515)
    cov = Cov(1, (10, 9, 5), sill=1.0, nugget=0.0)
    out, (od, om), pd, od_ = _run_pair(geo, (data, mask), g, _cdf_of(data, mask), (10, 9, 5), 48, cov, 4711)
    _same_lists(*pd, *od_)
    assert int(np.max(pd[1])) > 32
    assert np.array_equal(out.mask, om)
    assert rel_err(out.data, od, 1.0) <= TOL
    geo.set_precision("mixed")


def test_sgs_ordinary_kriging_without_transform(geo):
    """OK SGS (kriging kind 0), no cdf, with a nugget (normalised model)."""
    shape = (30, 30, 8)
    g = Grid(*shape)
    data, mask = _normal_scores(shape, 0.01, 99)
    cov = Cov(1, (10, 6, 3), angles=(30.0, 0.0, 0.0), sill=1.2, nugget=0.2)
    out, (od, om), pd, od_ = _run_pair(geo, (data, mask), g, None, (10, 10, 3), 12, cov, 4242, kind="ok")
    _same_lists(*pd, *od_)
    assert np.array_equal(out.mask, om)
    assert rel_err(out.data, od, 1.0) <= TOL


def test_sgs_lvm_and_simulation_mask(geo):
    """LVM SGS (per-cell mean, transformed with the data) on a partial
    simulation mask: cells outside the mask stay uninformed."""
    shape = (32, 24, 10)
    g = Grid(*shape)
    data, mask = _normal_scores(shape, 0.02, 5)
    mean = smooth_field(shape, 8.0, 6) * 0.5
    sim = np.zeros(shape, np.uint8)
    sim[4:28, 2:20, 1:9] = 1
    cov = Cov(2, (9, 9, 3), sill=1.0, nugget=0.0)
    out, (od, om), pd, od_ = _run_pair(geo, (data, mask), g, _cdf_of(data, mask), (9, 9, 3), 10, cov, 31337,
                                       mean=mean, mask=sim)
    _same_lists(*pd, *od_)
    assert np.array_equal(out.mask, om)
    assert not np.all(om == 1)
    assert rel_err(out.data, od, 1.0) <= TOL


@pytest.mark.parametrize("n", [6, 12])
def test_sgs_short_period_path(geo, n):
    """Grids whose path LCG does not have full period (C divides N): the
    reference visits some cells twice and some never; the product follows
    it (the revisits are skipped, unvisited cells stay uninformed)."""
    g = Grid(n, 1, 1)
    data = np.full((n, 1, 1), -99, np.float32)
    mask = np.zeros((n, 1, 1), np.uint8)
    data[1], mask[1] = 0.5, 1
    cov = Cov(1, (4, 1, 1))
    out, (od, om), pd, od_ = _run_pair(geo, (data, mask), g, None, (3, 1, 1), 3, cov, 11)
    _same_lists(*pd, *od_)
    assert np.array_equal(out.mask, om)
    assert rel_err(out.data[om == 1], od[om == 1], 1.0) <= TOL


def test_sgs_batch_equals_single_calls(geo):
    """hpgl_amd_sgs_simulation_batch: realisation r is bit-identical to a
    single call with seeds[r], and its neighbour lists match the oracle's."""
    shape = (24, 20, 10)
    g = Grid(*shape)
    data, mask = _normal_scores(shape, 0.015, 8)
    cdf = _cdf_of(data, mask)
    cov = Cov(0, (8, 8, 4))
    seeds = [3, 1000003, 77, 123456789]
    n, mx = g.size, 12
    pidx, pcnt = _dump_product(len(seeds) * n, mx)
    outs = geo.sgs_simulation_batch((data, mask), g, geo.CdfData(*cdf), (8, 8, 4), mx, cov, seeds)
    for r, s in enumerate(seeds):
        single = geo.sgs_simulation((data, mask), g, geo.CdfData(*cdf), (8, 8, 4), mx, cov, s)
        assert np.array_equal(outs[r].mask, single.mask)
        assert np.array_equal(outs[r].data, single.data), r
        d = O.Dump(n, mx)
        try:
            od, om = O.sgs_simulation((data, mask), g, cdf, (8, 8, 4), mx, cov, s)
        finally:
            d.close()
        _same_lists(pidx[r * n:(r + 1) * n], pcnt[r * n:(r + 1) * n], d.idx, d.cnt)
        assert rel_err(outs[r].data, od, 1.0) <= TOL


def test_sgs_rejects_empty_cdf(geo):
    shape = (8, 8, 4)
    data, mask = _normal_scores(shape, 0.05, 1)
    with pytest.raises(Exception):
        geo.sgs_simulation((data, mask), Grid(*shape), geo.CdfData(np.zeros(0, np.float32), np.zeros(0, np.float32)),
                           (4, 4, 2), 8, Cov(0, (4, 4, 2)), 1)
