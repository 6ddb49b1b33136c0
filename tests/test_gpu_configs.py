"""GPU parity at the parameters of the measured workloads.

Each test runs one of the benchmark workloads, at its own covariance,
search radii and max_neighbours, against the oracle on the same inputs:

* ordinary kriging on the full 100x100x50 grid (2% informed, lognormal,
  spherical ranges 20/15/5, nugget 0.1, radii 20/20/10, 12 neighbours);
* simple + 5-category indicator kriging on a 128x128x64 sub-grid of the
  headline workload's generator (1% informed, exponential 40/40/8 with
  nugget 0.05 for SK, spherical (30+5k) ranges per IK category, radii =
  ranges, 24 neighbours); IK is compared on every 3rd cell;
* four batched SGS realisations on 100x100x50 (1% normal scores,
  spherical 30/30/6, radii 30/30/6, 32 neighbours, seeds 1000..1003);
* a short Poisson request stream of the serving workload through the
  request loop, every job checked on sampled cells;
* the reference's regression calls (geo_testing/test_scripts/test.py:27-110)
  on its own property files, read through the library's INC reader.

Neighbour lists and masks must be bit-exact; estimates within 1e-5 relative
(scale floor 1) in mixed mode and bit-exact in FP64 mode; categories equal.
In mixed mode the SK systems must all take two refinement steps (continuous
estimates) and the IK systems almost all one (the certificate for
categorical outputs).
"""
import os

import numpy as np
import pytest

import oracle_binding as O
from cases import Cov, Grid, categories, rel_err, scattered

pytestmark = pytest.mark.gpu

TOL = 1e-5
HERE = os.path.dirname(os.path.abspath(__file__))


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


def _same_lists(p_idx, p_cnt, o_idx, o_cnt, rows=None):
    if rows is not None:
        p_idx, p_cnt, o_idx, o_cnt = p_idx[rows], p_cnt[rows], o_idx[rows], o_cnt[rows]
    bad = np.flatnonzero(p_cnt != o_cnt)
    assert bad.size == 0, "neighbour counts differ at rows %s" % bad[:10]
    cols = np.arange(p_idx.shape[1])[None, :] < np.maximum(p_cnt, 0)[:, None]
    diff = np.flatnonzero(np.any((p_idx != o_idx) & cols, axis=1))
    assert diff.size == 0, "neighbour lists differ at rows %s" % diff[:10]


def _values(out, od, om, mask, precision):
    m = (mask == 0) & (om == 1)
    if precision == "fp64":
        bad = int(np.sum(out.data[m] != od[m]))
        assert bad == 0, "%d FP64 values differ" % bad
    else:
        assert rel_err(out.data[m], od[m], 1.0) <= TOL


# ---- ordinary kriging, 100x100x50 ----

@pytest.fixture(scope="module")
def ok_case():
    g = Grid(100, 100, 50)
    data, mask = scattered(g.shape, 0.02, 1001)
    cov, r, mx = Cov(0, (20, 15, 5), sill=1.0, nugget=0.1), (20, 20, 10), 12
    d = O.Dump(g.size, mx)
    try:
        od, om = O.ordinary_kriging((data, mask), g, r, mx, cov)
    finally:
        d.close()
    return g, data, mask, cov, r, mx, od, om, d.idx, d.cnt


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_ordinary_kriging_full_grid(geo, ok_case, precision):
    g, data, mask, cov, r, mx, od, om, oidx, ocnt = ok_case
    geo.set_precision(precision)
    pidx, pcnt = _dump_product(g.size, mx)
    out = geo.ordinary_kriging((data, mask), g, r, mx, cov)
    _same_lists(pidx, pcnt, oidx, ocnt)
    assert np.array_equal(out.mask, om)
    _values(out, od, om, mask, precision)


# ---- simple + indicator kriging, 128x128x64 sub-grid ----

def _sk_ik_case():
    g = Grid(128, 128, 64)
    data, mask = scattered(g.shape, 0.01, 1002, law="normal")
    cat, marg = categories(data, mask)
    sk = (Cov(1, (40, 40, 8), sill=1.0, nugget=0.05), (40, 40, 8), 24)
    ik = [{"cov_model": Cov(0, (30 + 5 * k, 30 + 5 * k, (30 + 5 * k) // 5)),
           "radiuses": (30 + 5 * k, 30 + 5 * k, (30 + 5 * k) // 5), "max_neighbours": 24} for k in range(5)]
    return g, data, mask, cat, marg, sk, ik


@pytest.fixture(scope="module")
def sk_case():
    g, data, mask, cat, marg, sk, ik = _sk_ik_case()
    cov, r, mx = sk
    d = O.Dump(g.size, mx)
    try:
        od, om = O.simple_kriging((data, mask), g, r, mx, cov, mean=0.0)
    finally:
        d.close()
    return g, data, mask, sk, od, om, d.idx, d.cnt


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_simple_kriging_subgrid(geo, sk_case, precision):
    g, data, mask, (cov, r, mx), od, om, oidx, ocnt = sk_case
    geo.set_precision(precision)
    pidx, pcnt = _dump_product(g.size, mx)
    out = geo.simple_kriging((data, mask), g, r, mx, cov, mean=0.0)
    st = geo.last_stats()
    assert st["nodes_cluster_path"] > st["nodes_template_path"]
    if precision == "mixed":
        # continuous estimates: every FP32-accepted system is refined twice
        assert st["solves_fp32"] > 0 and st["solves_fp32_two_steps"] == st["solves_fp32"]
    _same_lists(pidx, pcnt, oidx, ocnt)
    assert np.array_equal(out.mask, om)
    _values(out, od, om, mask, precision)


@pytest.fixture(scope="module")
def ik_case():
    g, data, mask, cat, marg, sk, ik = _sk_ik_case()
    stride = 3
    d = O.Dump(5 * g.size, 24)
    O.set_sample_stride(stride)
    try:
        oc, om = O.indicator_kriging((cat, mask, 5), g, [dict(p) for p in ik], marg)
    finally:
        O.set_sample_stride(1)
        d.close()
    return g, cat, mask, marg, ik, stride, oc, d.idx, d.cnt


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_indicator_kriging_subgrid(geo, ik_case, precision):
    g, cat, mask, marg, ik, stride, oc, oidx, ocnt = ik_case
    geo.set_precision(precision)
    pidx, pcnt = _dump_product(5 * g.size, 24)
    out = geo.indicator_kriging((cat, mask, 5), g, [dict(p) for p in ik], marg)
    st = geo.last_stats()
    if precision == "mixed":
        # categorical outputs: the one-step certificate accepts almost every
        # system after the first refinement step
        assert st["solves_fp32"] > 0 and st["solves_fp32_two_steps"] < st["solves_fp32"] // 100
    samp = np.arange(g.size) % stride == 0
    rows = np.concatenate([np.flatnonzero(samp) + k * g.size for k in range(5)])
    _same_lists(pidx, pcnt, oidx, ocnt, rows)
    got = out.data.ravel(order="F")[samp]
    want = oc.ravel(order="F")[samp]
    flips = int(np.sum(got != want))
    assert np.all(out.mask == 1)
    if precision == "fp64":
        assert flips == 0, "%d categories differ in FP64 mode" % flips
    else:
        assert flips <= 2, "%d of %d categories differ in mixed mode" % (flips, got.size)


# ---- batched SGS, 100x100x50 ----

def test_sgs_four_realisations(geo):
    g = Grid(100, 100, 50)
    data, mask = scattered(g.shape, 0.01, 1008, law="normal")
    cov, r, mx = Cov(0, (30, 30, 6), sill=1.0, nugget=0.0), (30, 30, 6), 32
    seeds = [1000, 1001, 1002, 1003]
    n = g.size
    pidx, pcnt = _dump_product(len(seeds) * n, mx)
    outs = geo.sgs_simulation_batch((data, mask), g, None, r, mx, cov, seeds)
    for k, s in enumerate(seeds):
        d = O.Dump(n, mx)
        try:
            od, om = O.sgs_simulation((data, mask), g, None, r, mx, cov, s)
        finally:
            d.close()
        _same_lists(pidx[k * n:(k + 1) * n], pcnt[k * n:(k + 1) * n], d.idx, d.cnt)
        assert np.array_equal(outs[k].mask, om)
        assert rel_err(outs[k].data, od, 1.0) <= TOL, k


# ---- a short Poisson request stream through the request loop ----

def test_serving_stream_against_oracle():
    import bench
    from hpgl_amd.geo_bsd import _abi as abi
    from hpgl_amd.serving import Server

    jobs = bench.make_serving_trace(n_jobs=8, seed=2024)
    kinds = {"ok": abi.JOB_OK, "sk": abi.JOB_SK, "lvm": abi.JOB_LVM}
    with Server(slots=4) as srv:
        handles = []
        for job in jobs:
            p = bench._serving_params(abi, job)
            if job["kind"] == "ik":
                handles.append(srv.submit_ik(job["shape"], job["cat"], job["mask"], p, 3))
            else:
                handles.append(srv.submit_cont(kinds[job["kind"]], job["shape"], job["data"], job["mask"], p,
                                               mean=job.get("mean")))
        results = [h.wait() for h in handles]
    assert all(r["status"] == 0 for r in results), [r["message"] for r in results]
    for job, h in zip(jobs, handles):
        g = Grid(*job["shape"])
        stride = max(1, g.size // 20000)
        cov = Cov(job["ctype"], job["ranges"], sill=1.0, nugget=0.1)
        O.set_sample_stride(stride)
        try:
            if job["kind"] == "ok":
                od, om = O.ordinary_kriging((job["data"], job["mask"]), g, job["radii"], job["max_n"], cov)
            elif job["kind"] == "sk":
                od, om = O.simple_kriging((job["data"], job["mask"]), g, job["radii"], job["max_n"], cov)
            elif job["kind"] == "lvm":
                od, om = O.lvm_kriging((job["data"], job["mask"]), g, job["mean"], job["radii"], job["max_n"], cov)
            else:
                ikd = [{"cov_model": cov, "radiuses": job["radii"], "max_neighbours": job["max_n"]}
                       for _ in range(3)]
                od, om = O.indicator_kriging((job["cat"], job["mask"], 3), g, ikd, job["marg"])
        finally:
            O.set_sample_stride(1)
        sel = np.arange(g.size) % stride == 0
        got, gm = h.out.ravel(order="F")[sel], h.out_mask.ravel(order="F")[sel]
        want, wm = od.ravel(order="F")[sel], om.ravel(order="F")[sel]
        assert np.array_equal(gm, wm), job["kind"]
        if job["kind"] == "ik":
            assert int(np.sum(got != want)) <= 2
        else:
            ok = (job["mask"].ravel(order="F")[sel] == 0) & (wm == 1)
            assert rel_err(got[ok], want[ok], 1.0) <= TOL, job["kind"]


# ---- the reference's regression calls on its own property files ----

def test_reference_property_files_through_the_reader(geo):
    inc = os.path.join(HERE, "golden", "inc")
    g = Grid(55, 52, 1)
    prop = geo.load_cont_property(os.path.join(inc, "NEW_TEST_PROP.INC"), -99, size=g.shape)
    mean = geo.load_cont_property(os.path.join(inc, "mp_0_8_52_55_1.inc"), -99, size=g.shape).data
    ind = geo.load_ind_property(os.path.join(inc, "NEW_TEST_PROP_01.INC"), -99, [0, 1], size=g.shape)
    data, mask = prop.data, prop.mask
    assert 0 < int(mask.sum()) < g.size
    exp = Cov(1, (10, 10, 10), sill=1.0)
    r = (20, 20, 20)
    out = geo.ordinary_kriging((data, mask), g, r, 12, exp)
    od, om = O.ordinary_kriging((data, mask), g, r, 12, exp)
    assert np.array_equal(out.mask, om)
    _values(out, od, om, mask, "mixed")
    out = geo.simple_kriging((data, mask), g, r, 12, exp, mean=0.487)
    od, om = O.simple_kriging((data, mask), g, r, 12, exp, mean=0.487)
    assert np.array_equal(out.mask, om)
    _values(out, od, om, mask, "mixed")
    out = geo.lvm_kriging((data, mask), g, mean, r, 12, exp)
    od, om = O.lvm_kriging((data, mask), g, mean, r, 12, exp)
    assert np.array_equal(out.mask, om)
    _values(out, od, om, mask, "mixed")
    sph = Cov(0, (10, 10, 10), sill=0.4)
    out = geo.median_ik((ind.data, ind.mask, 2), g, (0.5, 0.5), (10, 10, 10), 12, sph)
    od, om = O.median_ik((ind.data, ind.mask, 2), g, (0.5, 0.5), (10, 10, 10), 12, sph)
    assert np.array_equal(out.mask, om) and int(np.sum(out.data != od)) == 0
