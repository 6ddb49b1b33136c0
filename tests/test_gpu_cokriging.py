"""GPU parity of simple colocated cokriging (Markov models I and II)
against the CPU oracle.

The device solves each (n+1)-system with gauss_solve in the reference's
operation order in FP64 from the same covariance values, so every estimate
must equal the oracle's bit for bit, including the failure value
(no neighbours or a singular system) and systems of 65 unknowns
(max_neighbours 64).
"""
import ctypes as C

import numpy as np
import pytest

import oracle_binding as O
from cases import Cov, Grid

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def geo():
    import hpgl_amd.geo_bsd as g

    return g


def _case(grid, frac, seed, sec_frac=0.7):
    rng = np.random.default_rng(seed)
    n = grid.size
    idx = rng.choice(n, max(1, int(frac * n)), replace=False)
    mask = np.zeros(n, np.uint8)
    mask[idx] = 1
    data = np.full(n, -99.0, np.float32)
    data[idx] = rng.lognormal(0, 0.5, idx.size).astype(np.float32)
    smask = (rng.random(n) < sec_frac).astype(np.uint8)
    sdata = np.where(smask == 1, rng.normal(2.0, 0.4, n), -99.0).astype(np.float32)
    f = lambda a: np.asfortranarray(a.reshape(grid.shape, order="F"))
    return f(data), f(mask), f(sdata), f(smask)


@pytest.mark.parametrize("seed,cov_type,max_n", [(1, 1, 12), (2, 0, 20), (3, 2, 64)])
def test_markI_equals_oracle(geo, seed, cov_type, max_n):
    g = Grid(40, 36, 12)
    data, mask, sdata, smask = _case(g, 0.03 if max_n < 64 else 0.25, seed)
    cm = Cov(cov_type, (10.0, 8.0, 4.0), (30.0, 0.0, 0.0), 1.0, 0.1 if cov_type == 2 else 0.0)
    radii = (8, 8, 4)
    print("cokriging mark1 seed=%d type=%d max_n=%d" % (seed, cov_type, max_n), flush=True)
    gp = geo.simple_cokriging_markI((data, mask), g, (sdata, smask), 1.2, 2.0, 0.16, 0.6, radii, max_n,
                                    geo.CovarianceModel(cm.type, cm.ranges, cm.angles, cm.sill, cm.nugget))
    od, om = O.simple_cokriging_markI((data, mask), g, (sdata, smask), 1.2, 2.0, 0.16, 0.6, radii, max_n, cm)
    assert np.array_equal(gp.mask, om)
    assert np.array_equal(gp.data, od), np.argwhere(gp.data != od)[:10]
    st = geo.last_stats()
    assert st["nodes_kriged"] > 0


def test_markII_equals_oracle(geo):
    g = Grid(36, 30, 10)
    data, mask, sdata, smask = _case(g, 0.04, 7)
    cm1 = Cov(1, (9.0, 9.0, 4.0))
    cm2 = Cov(0, (12.0, 10.0, 5.0), (15.0, 0.0, 0.0), 0.5, 0.05)
    gm = lambda c: geo.CovarianceModel(c.type, c.ranges, c.angles, c.sill, c.nugget)
    print("cokriging mark2", flush=True)
    gp = geo.simple_cokriging_markII(
        g, {"data": geo.ContProperty(data, mask), "mean": 1.1, "cov_model": gm(cm1)},
        {"data": geo.ContProperty(sdata, smask), "mean": 1.9, "cov_model": gm(cm2)}, 0.4, (7, 7, 3), 16)
    od, om = O.simple_cokriging_markII(g, {"data": (data, mask), "mean": 1.1, "cov_model": cm1},
                                       {"data": (sdata, smask), "mean": 1.9, "cov_model": cm2}, 0.4, (7, 7, 3), 16)
    assert np.array_equal(gp.mask, om)
    assert np.array_equal(gp.data, od), np.argwhere(gp.data != od)[:10]


def test_failure_values_equal_oracle(geo):
    g = Grid(20, 16, 6)
    data, mask, sdata, smask = _case(g, 0.01, 9)
    cm = Cov(1, (3.0, 3.0, 3.0))
    gcm = geo.CovarianceModel(1, (3.0, 3.0, 3.0))
    for var2, rho, radii in ((0.0, 0.0, (2, 2, 2)), (0.3, 0.5, (1, 1, 1))):
        gp = geo.simple_cokriging_markI((data, mask), g, (sdata, smask), 0.5, 1.5, var2, rho, radii, 8, gcm)
        od, om = O.simple_cokriging_markI((data, mask), g, (sdata, smask), 0.5, 1.5, var2, rho, radii, 8, cm)
        assert np.array_equal(gp.data, od)
        assert np.array_equal(gp.mask, om)


def test_device_pointer_variant_and_size_check(geo):
    from hip_device import DeviceArray
    from hpgl_amd.geo_bsd import hpgl_wrap as W
    from hpgl_amd.geo_bsd import _abi as abi

    g = Grid(24, 20, 8)
    data, mask, sdata, smask = _case(g, 0.05, 11)
    gcm = geo.CovarianceModel(1, (6.0, 6.0, 3.0))
    host = geo.simple_cokriging_markI((data, mask), g, (sdata, smask), 1.0, 2.0, 0.2, 0.5, (5, 5, 3), 10, gcm)
    f = lambda a: a.ravel(order="F")
    d_in, d_m, d_s, d_sm = DeviceArray(f(data)), DeviceArray(f(mask)), DeviceArray(f(sdata)), DeviceArray(f(smask))
    d_out, d_om = DeviceArray(f(data)), DeviceArray(f(mask))
    p = abi._HPGL_COKRIGING_M1_PARAMS(covariance_type=1, ranges=(C.c_double * 3)(6.0, 6.0, 3.0),
                                      angles=(C.c_double * 3)(0, 0, 0), sill=1.0, nugget=0.0,
                                      radiuses=(C.c_int * 3)(5, 5, 3), max_neighbours=10, primary_mean=1.0,
                                      secondary_mean=2.0, secondary_variance=0.2, correlation_coef=0.5)
    sh = abi._HPGL_SHAPE(data=(C.c_int * 3)(*g.shape), strides=(C.c_int * 3)(1, g.x, g.x * g.y))
    W._hpgl_so.hpgl_amd_simple_cokriging_mark1_dev(d_in.ptr, d_m.ptr, d_s.ptr, d_sm.ptr, C.byref(sh), C.byref(p),
                                                   d_out.ptr, d_om.ptr, None)
    W.check_last_error()
    assert np.array_equal(d_out.numpy(), host.data.ravel(order="F"))
    assert np.array_equal(d_om.numpy(), host.mask.ravel(order="F"))
    small = (sdata[:, :, :4].copy(order="F"), smask[:, :, :4].copy(order="F"))
    # the Python layer checks sizes against the grid before the call
    with pytest.raises(Exception, match="(?i)size of"):
        geo.simple_cokriging_markI((data, mask), g, small, 1.0, 2.0, 0.2, 0.5, (5, 5, 3), 10, gcm)
