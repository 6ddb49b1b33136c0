"""The allocation-free SK weights entry point against the allocating one.

hpgl_amd_simple_kriging_weights_noalloc keeps its system in a static device
array; two calls in a row (the second smaller, so stale scratch from the
first would show) must each give the float weights of
hpgl_simple_kriging_weights bit for bit, oversize inputs must be refused,
and a singular system must report the error and leave the weights alone.
"""
import ctypes as C

import numpy as np
import pytest

from hpgl_amd.geo_bsd import hpgl_wrap
from hpgl_amd.geo_bsd import _abi as abi

lib = hpgl_wrap._hpgl_so


def _cov(type_, ranges, sill=1.0, nugget=0.0, angles=(0.0, 0.0, 0.0)):
    return abi._HPGL_COV_PARAMS(covariance_type=type_, ranges=(C.c_double * 3)(*ranges),
                                angles=(C.c_double * 3)(*angles), sill=sill, nugget=nugget)


def _host_weights(center, x, y, z, cov):
    f = lambda a: np.ascontiguousarray(a, np.float32)
    c, x, y, z = f(center), f(x), f(y), f(z)
    w = np.full(len(x), -7.0, np.float32)
    lib.hpgl_simple_kriging_weights(c.ctypes.data_as(C.POINTER(C.c_float)), x, y, z, len(x), C.byref(cov), w)
    hpgl_wrap.check_last_error()
    return w


def _noalloc_weights(center, x, y, z, cov, fill=-7.0):
    from hip_device import DeviceArray

    f = lambda a: np.ascontiguousarray(a, np.float32)
    dc, dx, dy, dz = (DeviceArray(f(a)) for a in (center, x, y, z))
    dw = DeviceArray(np.full(len(x), fill, np.float32))
    rc = lib.hpgl_amd_simple_kriging_weights_noalloc(dc.ptr, dx.ptr, dy.ptr, dz.ptr, len(x), C.byref(cov), dw.ptr,
                                                     None)
    err = lib.hpgl_amd_get_last_error()
    return rc, err, dw.numpy()


def _points(n, seed, spread=20.0):
    rng = np.random.default_rng(seed)
    pts = rng.uniform(0.0, spread, (n, 3))
    center = rng.uniform(0.0, spread, 3)
    return center, pts[:, 0], pts[:, 1], pts[:, 2]


@pytest.mark.gpu
@pytest.mark.parametrize("cov_args", [(0, (15.0, 10.0, 5.0), 1.0, 0.1), (1, (10.0, 10.0, 10.0), 1.0, 0.0),
                                      (2, (4.0, 3.0, 2.0), 1.0, 0.05, (30.0, 10.0, 5.0))])
def test_noalloc_equals_allocating_entry_point_twice(cov_args):
    cov = _cov(*cov_args)
    for n, seed in ((96, 1), (17, 2)):  # the second call is smaller
        center, x, y, z = _points(n, seed)
        ref = _host_weights(center, x, y, z, cov)
        rc, err, w = _noalloc_weights(center, x, y, z, cov)
        assert rc == 0 and err == 0
        assert np.array_equal(w, ref), "n=%d: max diff %g" % (n, np.max(np.abs(w - ref)))


@pytest.mark.gpu
def test_noalloc_refuses_oversize_and_reports_singular():
    cov = _cov(1, (10.0, 10.0, 10.0))
    center, x, y, z = _points(513, 3)
    rc, err, w = _noalloc_weights(center, x, y, z, cov)
    assert err != 0 and np.all(w == -7.0)
    assert b"at most 512" in lib.hpgl_get_last_exception_message()
    # two coincident points: the system is singular, the weights stay untouched
    x = np.array([1.0, 1.0, 5.0])
    y = np.array([2.0, 2.0, 3.0])
    z = np.zeros(3)
    rc, err, w = _noalloc_weights([0.0, 0.0, 0.0], x, y, z, cov)
    assert err != 0 and np.all(w == -7.0)
    # and a regular call afterwards still works
    center, x, y, z = _points(40, 4)
    ref = _host_weights(center, x, y, z, cov)
    rc, err, w = _noalloc_weights(center, x, y, z, cov)
    assert err == 0 and np.array_equal(w, ref)
