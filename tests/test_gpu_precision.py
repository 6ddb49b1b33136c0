"""Precision stress: near-singular and singular systems on the GPU against
the oracle.

The reference decides success from LAPACK's dpotrf/dpotrs info
(my_kriging_weights.h:193-213) and that decision sets the OK mask or the SK
failure value (cont_kriging.h:90-100).  These tests require the product to
take every such decision as the oracle does, in both precision modes:

* the precision-stress workload (informed density 0.1-5%, x-adjacent
  near-duplicate twins, Gaussian covariance of range 25 with nugget 1e-4,
  max_neighbours 64) on a 100x100x50 sub-grid: neighbour lists and masks
  bit-exact, FP64 values bit-exact, mixed values within 1e-5;
* indicator kriging with Gaussian models at 24 neighbours, whose systems
  take every path of the mixed kernel (one refinement step, two, FP64,
  singular): categories equal to the oracle's;
* Gaussian jobs whose FP64 factorisation fails for part of the nodes (SK at
  24 and 32 neighbours, which the mixed FP32 kernels handle first, and OK
  at 64): the singular counts equal the oracle's, mixed mode must send those
  systems to FP64 (solves_fp64 > 0), and every mask and failure value must
  match.
"""
import numpy as np
import pytest

import oracle_binding as O
from cases import Cov, Grid, categories, rel_err, scattered, uneven_twins

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


def _same_lists(p_idx, p_cnt, o_idx, o_cnt):
    """Equal counts, and equal entries in the first count columns of every row."""
    bad = np.flatnonzero(p_cnt != o_cnt)
    assert bad.size == 0, "neighbour counts differ at cells %s" % bad[:10]
    cols = np.arange(p_idx.shape[1])[None, :] < np.maximum(p_cnt, 0)[:, None]
    diff = np.flatnonzero(np.any((p_idx != o_idx) & cols, axis=1))
    assert diff.size == 0, "neighbour lists differ at cells %s" % diff[:10]


@pytest.fixture(scope="module")
def stress_case():
    shape = (100, 100, 50)
    data, mask, twins = uneven_twins(shape, 1003)
    g = Grid(*shape)
    cov = Cov(2, (25, 25, 25), sill=1.0, nugget=1e-4)
    r, mx = (25, 25, 25), 64
    d = O.Dump(g.size, mx)
    try:
        od, om = O.ordinary_kriging((data, mask), g, r, mx, cov)
    finally:
        d.close()
    none, sing = O.last_counts()
    return dict(g=g, data=data, mask=mask, cov=cov, r=r, mx=mx, od=od, om=om, idx=d.idx, cnt=d.cnt, twins=twins,
                none=none, sing=sing)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_precision_stress_subgrid(geo, stress_case, precision):
    c = stress_case
    assert c["twins"] > 0
    geo.set_precision(precision)
    pidx, pcnt = _dump_product(c["g"].size, c["mx"])
    out = geo.ordinary_kriging((c["data"], c["mask"]), c["g"], c["r"], c["mx"], c["cov"])
    st = geo.last_stats()
    _same_lists(pidx, pcnt, c["idx"], c["cnt"])
    assert st["nodes_template_path"] > 0 and st["nodes_cluster_path"] > 0, "both search paths are exercised"
    assert np.array_equal(out.mask, c["om"]), "masks (dpotrf decisions) differ"
    assert st["nodes_singular"] == c["sing"] and st["nodes_no_neighbours"] == c["none"]
    ok = (c["mask"] == 0) & (c["om"] == 1)
    if precision == "fp64":
        diff = np.flatnonzero(out.data[ok] != c["od"][ok])
        assert diff.size == 0, "FP64 values differ from the oracle at %d cells" % diff.size
    else:
        # systems of 33..64 unknowns are factorised in FP64 in both modes
        assert st["solves_fp64"] > 0
        assert rel_err(out.data[ok], c["od"][ok], 1.0) <= TOL


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_indicator_kriging_gaussian_refinement_paths(geo, precision):
    """Indicator kriging with Gaussian models at 24 neighbours: in mixed mode
    the order-24 systems take every path of the mixed kernel (accepted after
    one refinement step by the certificate, after two steps, handed to FP64,
    and singular in FP64), and the categories still equal the oracle's."""
    g = Grid(48, 48, 16)
    data, mask = scattered(g.shape, 0.04, 77, law="normal")
    cat, marg = categories(data, mask, quantiles=(0.3, 0.7), k=3)
    params = [{"cov_model": Cov(2, (10, 10, 4), nugget=0.05), "radiuses": (10, 10, 4), "max_neighbours": 24}
              for _ in range(3)]
    oc, om = O.indicator_kriging((cat, mask, 3), g, [dict(p) for p in params], marg)
    geo.set_precision(precision)
    out = geo.indicator_kriging((cat, mask, 3), g, [dict(p) for p in params], marg)
    st = geo.last_stats()
    assert np.array_equal(out.mask, om)
    flips = int(np.sum(out.data != oc))
    if precision == "fp64":
        assert flips == 0, "%d categories differ in FP64 mode" % flips
    else:
        one_step = st["solves_fp32"] - st["solves_fp32_two_steps"]
        assert one_step > 0 and st["solves_fp32_two_steps"] > 0, "both refinement lengths are exercised"
        assert st["solves_fp64"] > 0 and st["nodes_singular"] > 0, "FP64 hand-over and singular systems occur"
        assert flips <= 2, "%d of %d categories differ in mixed mode" % (flips, out.data.size)


SINGULAR_CASES = [
    # (kind, shape, informed fraction, Gaussian range, radii, max_neighbours)
    ("sk", (40, 40, 20), 0.2, 40, (12, 12, 6), 32),
    ("sk", (60, 50, 20), 0.3, 80, (10, 10, 5), 24),
    ("ok", (40, 40, 20), 0.1, 60, (20, 20, 10), 64),
]


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
@pytest.mark.parametrize("case", SINGULAR_CASES, ids=lambda c: "%s-%d" % (c[0], c[5]))
def test_singular_decisions_equal_oracle(geo, case, precision):
    kind, shape, frac, rng, radii, mx = case
    g = Grid(*shape)
    data, mask = scattered(shape, frac, 11, law="normal")
    cov = Cov(2, (rng, rng, rng), sill=1.0, nugget=1e-4)
    if kind == "sk":
        od, om = O.simple_kriging((data, mask), g, radii, mx, cov, mean=0.0)
    else:
        od, om = O.ordinary_kriging((data, mask), g, radii, mx, cov)
    none, sing = O.last_counts()
    assert sing > 0, "the case must contain systems whose FP64 factorisation fails"
    geo.set_precision(precision)
    if kind == "sk":
        out = geo.simple_kriging((data, mask), g, radii, mx, cov, mean=0.0)
    else:
        out = geo.ordinary_kriging((data, mask), g, radii, mx, cov)
    st = geo.last_stats()
    assert st["nodes_singular"] == sing and st["nodes_no_neighbours"] == none
    assert np.array_equal(out.mask, om)
    unk = mask == 0
    if precision == "fp64":
        assert np.array_equal(out.data[unk], od[unk])
    else:
        assert st["solves_fp64"] > 0, "near-singular systems go to the FP64 factorisation"
        ok = unk & (om == 1)
        assert rel_err(out.data[ok], od[ok], 1.0) <= TOL
