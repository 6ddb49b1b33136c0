"""ctypes binding of the CPU oracle (oracle/_build/liboracle.so) for the tests.

Test infrastructure only: the oracle is the checker.  The functions take
and return the same Python objects as hpgl_amd.geo_bsd (ContProperty-like
(data, mask) pairs in Fortran order, CovarianceModel-like objects), so a
parity test reads ``oracle.ordinary_kriging(...)`` beside
``geo_bsd.ordinary_kriging(...)``.
"""
import ctypes as C
import importlib.util
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_SO = os.path.join(ORACLE_DIR, "_build", "liboracle.so")


def _load_abi():
    # The struct mirrors only; importing the package would load the GPU library.
    spec = importlib.util.spec_from_file_location("_hpgl_abi", os.path.join(ROOT, "hpgl_amd", "geo_bsd", "_abi.py"))
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    return m


abi = _load_abi()


def _load_cvar_abi():
    spec = importlib.util.spec_from_file_location("_cvar_abi", os.path.join(ROOT, "hpgl_amd", "geo_bsd",
                                                                             "_cvar_abi.py"))
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    return m


cvabi = _load_cvar_abi()


def _load():
    # ORACLE_SO may name another build of the same source (the CPU baseline
    # loads the -march=native one)
    path = os.environ.get("ORACLE_SO", ORACLE_SO)
    src = os.path.join(ORACLE_DIR, "hpgl_oracle.cpp")
    if path == ORACLE_SO and (not os.path.exists(ORACLE_SO) or os.path.getmtime(ORACLE_SO) < os.path.getmtime(src)):
        subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)
    return C.CDLL(path)


lib = _load()
_P = C.POINTER
lib.oracle_last_error.restype = C.c_char_p
lib.oracle_ordinary_kriging.argtypes = [_P(abi._HPGL_CONT_MASKED_ARRAY), _P(abi._HPGL_OK_PARAMS),
                                        _P(abi._HPGL_CONT_MASKED_ARRAY)]
lib.oracle_simple_kriging.argtypes = [C.c_void_p, C.c_void_p, _P(abi._HPGL_SHAPE), _P(abi._HPGL_SK_PARAMS),
                                      C.c_void_p, C.c_void_p, _P(abi._HPGL_SHAPE)]
lib.oracle_lvm_kriging.argtypes = [C.c_void_p, C.c_void_p, _P(abi._HPGL_SHAPE), C.c_void_p, _P(abi._HPGL_SHAPE),
                                   _P(abi._HPGL_OK_PARAMS), C.c_void_p, C.c_void_p, _P(abi._HPGL_SHAPE)]
lib.oracle_indicator_kriging.argtypes = [_P(abi._HPGL_IND_MASKED_ARRAY), _P(abi._HPGL_IND_MASKED_ARRAY),
                                         _P(abi._HPGL_IK_PARAMS), C.c_int]
lib.oracle_median_ik.argtypes = [_P(abi._HPGL_IND_MASKED_ARRAY), _P(abi._HPGL_MEDIAN_IK_PARAMS),
                                 _P(abi._HPGL_IND_MASKED_ARRAY)]
lib.oracle_last_counts.argtypes = [C.c_void_p]
lib.oracle_sgs_simulation_batch.argtypes = [C.c_void_p, C.c_void_p, _P(abi._HPGL_SHAPE), _P(abi._HPGL_SGS_PARAMS),
                                            C.c_void_p, C.c_int]
lib.oracle_sgs_simulation.argtypes = [_P(abi._HPGL_CONT_MASKED_ARRAY), _P(abi._HPGL_SGS_PARAMS),
                                      _P(abi.hpgl_non_parametric_cdf_t), _P(C.c_double), _P(abi._HPGL_UBYTE_ARRAY)]
lib.oracle_sgs_lvm_simulation.argtypes = [_P(abi._HPGL_CONT_MASKED_ARRAY), _P(abi._HPGL_SGS_PARAMS),
                                          _P(abi.hpgl_non_parametric_cdf_t), _P(abi._HPGL_FLOAT_ARRAY),
                                          _P(abi._HPGL_UBYTE_ARRAY)]
lib.oracle_sis_simulation.argtypes = [_P(abi._HPGL_IND_MASKED_ARRAY), _P(abi._HPGL_IK_PARAMS), C.c_int, C.c_int,
                                      _P(abi._HPGL_UBYTE_ARRAY)]
lib.oracle_sis_simulation_lvm.argtypes = [_P(abi._HPGL_IND_MASKED_ARRAY), _P(abi._HPGL_IK_PARAMS),
                                          _P(abi._HPGL_FLOAT_ARRAY), C.c_int, C.c_int, _P(abi._HPGL_UBYTE_ARRAY),
                                          C.c_int]
lib.oracle_simple_cokriging_mark1.argtypes = [_P(abi._HPGL_CONT_MASKED_ARRAY), _P(abi._HPGL_CONT_MASKED_ARRAY),
                                              _P(abi._HPGL_COKRIGING_M1_PARAMS), _P(abi._HPGL_CONT_MASKED_ARRAY)]
lib.oracle_simple_cokriging_mark2.argtypes = [_P(abi._HPGL_CONT_MASKED_ARRAY), _P(abi._HPGL_CONT_MASKED_ARRAY),
                                              _P(abi._HPGL_COKRIGING_M2_PARAMS), _P(abi._HPGL_CONT_MASKED_ARRAY)]
lib.oracle_gauss_solve.restype = C.c_int
lib.oracle_gauss_solve.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
lib.oracle_calc_variograms.argtypes = [_P(cvabi.variogram_search_template_t), _P(cvabi.hard_data_t), C.c_void_p,
                                       C.c_int]
lib.oracle_calc_variograms_from_point_set.argtypes = [_P(cvabi.variogram_search_template_t),
                                                      _P(cvabi.cont_point_set_t), C.c_void_p, C.c_int]
lib.oracle_fill_ellipsoid_directions.argtypes = [_P(cvabi.ellipsoid_t), C.c_double, C.c_double, C.c_double]
lib.oracle_stack_layers.argtypes = [_P(cvabi.float_data_t), _P(C.c_int), C.c_int, C.c_int, C.c_float, C.c_int,
                                    _P(cvabi.float_data_t)]
lib.oracle_simple_kriging_weights.restype = C.c_int
lib.oracle_simple_kriging_weights.argtypes = [C.c_void_p] * 4 + [C.c_int, _P(abi._HPGL_COV_PARAMS), C.c_void_p,
                                                                  _P(C.c_double)]
lib.oracle_ok_weights_int.restype = C.c_int
lib.oracle_ok_weights_int.argtypes = [C.c_void_p, C.c_void_p, C.c_int, _P(abi._HPGL_COV_PARAMS), C.c_void_p,
                                      _P(C.c_double)]
lib.oracle_cholesky_solve.restype = C.c_int
lib.oracle_cholesky_solve.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
lib.oracle_covariance.restype = C.c_double
lib.oracle_covariance.argtypes = [_P(abi._HPGL_COV_PARAMS), C.c_double, C.c_double, C.c_double]
lib.oracle_template.restype = C.c_int
lib.oracle_template.argtypes = [_P(abi._HPGL_COV_PARAMS), C.c_void_p, C.c_int, C.c_void_p, C.c_int]
lib.oracle_sgs_path.argtypes = [C.c_long, C.c_long, C.c_void_p, C.c_long]
lib.oracle_mt11213b.argtypes = [C.c_uint, C.c_void_p, C.c_long]
lib.oracle_gauss_inverse.restype = C.c_double
lib.oracle_gauss_inverse.argtypes = [C.c_double]
lib.oracle_gauss_prob.restype = C.c_double
lib.oracle_gauss_prob.argtypes = [C.c_double]
lib.oracle_set_neighbour_dump.argtypes = [C.c_void_p, C.c_void_p, C.c_longlong, C.c_int]
lib.oracle_set_threads.argtypes = [C.c_int]
lib.oracle_set_progress_tick.argtypes = [C.c_int]
lib.oracle_set_sample_stride.argtypes = [C.c_int]


class OracleError(RuntimeError):
    pass


def _check():
    msg = lib.oracle_last_error()
    if msg:
        raise OracleError(msg.decode())


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


def _shape(grid):
    return abi._HPGL_SHAPE(data=(C.c_int * 3)(grid.x, grid.y, grid.z),
                           strides=(C.c_int * 3)(1, grid.x, grid.x * grid.y))


def _cov_kwargs(cm):
    return dict(covariance_type=cm.type, ranges=(C.c_double * 3)(*cm.ranges), angles=(C.c_double * 3)(*cm.angles),
                sill=cm.sill, nugget=cm.nugget)


def cov_params(cm):
    return abi._HPGL_COV_PARAMS(**_cov_kwargs(cm))


def _cont(data, mask):
    return (np.require(data, np.float32, "F").copy(order="F"), np.require(mask, np.uint8, "F").copy(order="F"))


def _cma(data, mask, grid):
    return abi._HPGL_CONT_MASKED_ARRAY(data=data.ctypes.data_as(_P(C.c_float)), mask=mask.ctypes.data_as(_P(C.c_ubyte)),
                                       shape=_shape(grid))


def _ima(data, mask, k, grid):
    return abi._HPGL_IND_MASKED_ARRAY(data=data.ctypes.data_as(_P(C.c_ubyte)), mask=mask.ctypes.data_as(_P(C.c_ubyte)),
                                      shape=_shape(grid), indicator_count=k)


class Dump:
    """Collects per-cell neighbour lists from the next oracle call."""

    def __init__(self, n_cells, max_n):
        self.idx = np.full((n_cells, max_n), -1, dtype=np.int32)
        self.cnt = np.full(n_cells, -1, dtype=np.int32)
        lib.oracle_set_neighbour_dump(_ptr(self.idx), _ptr(self.cnt), n_cells, max_n)

    def close(self):
        lib.oracle_set_neighbour_dump(None, None, 0, 0)

    def lists(self):
        return [tuple(self.idx[i, :c]) if c >= 0 else None for i, c in enumerate(self.cnt)]


def ordinary_kriging(prop, grid, radiuses, max_neighbours, cov_model):
    data, mask = prop
    ind, inm = _cont(data, mask)
    od, om = ind.copy(order="F"), inm.copy(order="F")
    p = abi._HPGL_OK_PARAMS(radiuses=(C.c_int * 3)(*radiuses), max_neighbours=max_neighbours, **_cov_kwargs(cov_model))
    lib.oracle_ordinary_kriging(C.byref(_cma(ind, inm, grid)), C.byref(p), C.byref(_cma(od, om, grid)))
    _check()
    return od, om


def simple_kriging(prop, grid, radiuses, max_neighbours, cov_model, mean=None):
    data, mask = prop
    ind, inm = _cont(data, mask)
    od, om = ind.copy(order="F"), inm.copy(order="F")
    p = abi._HPGL_SK_PARAMS(radiuses=(C.c_int * 3)(*radiuses), max_neighbours=max_neighbours,
                            automatic_mean=(mean is None), mean=(mean if mean is not None else 0),
                            **_cov_kwargs(cov_model))
    sh = _shape(grid)
    lib.oracle_simple_kriging(_ptr(ind), _ptr(inm), C.byref(sh), C.byref(p), _ptr(od), _ptr(om), C.byref(sh))
    _check()
    return od, om


def lvm_kriging(prop, grid, mean_data, radiuses, max_neighbours, cov_model):
    data, mask = prop
    ind, inm = _cont(data, mask)
    od, om = ind.copy(order="F"), inm.copy(order="F")
    md = np.require(mean_data, np.float32, "F")
    p = abi._HPGL_OK_PARAMS(radiuses=(C.c_int * 3)(*radiuses), max_neighbours=max_neighbours, **_cov_kwargs(cov_model))
    sh = _shape(grid)
    lib.oracle_lvm_kriging(_ptr(ind), _ptr(inm), C.byref(sh), _ptr(md), C.byref(sh), C.byref(p), _ptr(od), _ptr(om),
                           C.byref(sh))
    _check()
    return od, om


def indicator_kriging(prop, grid, data, marginal_probs):
    """K >= 3 categories (the K == 2 case goes to median_ik, as in geo.py:493)."""
    d, m, k = prop
    ind = np.require(d, np.uint8, "F").copy(order="F")
    inm = np.require(m, np.uint8, "F").copy(order="F")
    od, om = ind.copy(order="F"), inm.copy(order="F")
    ps = (abi._HPGL_IK_PARAMS * len(data))()
    for i, dd in enumerate(data):
        ps[i] = abi._HPGL_IK_PARAMS(radiuses=(C.c_int * 3)(*dd["radiuses"]), max_neighbours=dd["max_neighbours"],
                                    marginal_prob=marginal_probs[i], **_cov_kwargs(dd["cov_model"]))
    lib.oracle_indicator_kriging(C.byref(_ima(ind, inm, k, grid)), C.byref(_ima(od, om, k, grid)), ps, len(data))
    _check()
    return od, om


def median_ik(prop, grid, marginal_probs, radiuses, max_neighbours, cov_model):
    d, m, k = prop
    ind = np.require(d, np.uint8, "F").copy(order="F")
    inm = np.require(m, np.uint8, "F").copy(order="F")
    od, om = ind.copy(order="F"), inm.copy(order="F")
    p = abi._HPGL_MEDIAN_IK_PARAMS(radiuses=(C.c_int * 3)(*radiuses), max_neighbours=max_neighbours,
                                   marginal_probs=(C.c_double * 2)(*marginal_probs), **_cov_kwargs(cov_model))
    lib.oracle_median_ik(C.byref(_ima(ind, inm, k, grid)), C.byref(p), C.byref(_ima(od, om, k, grid)))
    _check()
    return od, om


def sgs_simulation(prop, grid, cdf, radiuses, max_neighbours, cov_model, seed, kriging_type="sk", mean=None,
                   mask=None):
    """Runs the C-level SGS on an already-normalised covariance model.

    cdf is None or (values, probs) float32 arrays; mean is None (stationary)
    or an array (locally varying mean)."""
    data, m = prop
    od, om = _cont(data, m)
    p = abi._HPGL_SGS_PARAMS(radiuses=(C.c_int * 3)(*radiuses), max_neighbours=max_neighbours,
                             kriging_kind={"sk": 1, "ok": 0}[kriging_type], seed=seed, min_neighbours=0,
                             **_cov_kwargs(cov_model))
    keep = []
    c = None
    if cdf is not None:
        v = np.ascontiguousarray(cdf[0], np.float32)
        pr = np.ascontiguousarray(cdf[1], np.float32)
        keep += [v, pr]
        c = C.byref(abi.hpgl_non_parametric_cdf_t(values=v.ctypes.data_as(_P(C.c_float)),
                                                  probs=pr.ctypes.data_as(_P(C.c_float)), size=v.size))
    sm = None
    if mask is not None:
        mk = np.require(mask, np.uint8, "F")
        keep.append(mk)
        sm = C.byref(abi._HPGL_UBYTE_ARRAY(data=mk.ctypes.data_as(_P(C.c_ubyte)), shape=_shape(grid)))
    arr = _cma(od, om, grid)
    if mean is None or np.isscalar(mean):
        lib.oracle_sgs_simulation(C.byref(arr), C.byref(p), c, None, sm)
    else:
        md = np.require(mean, np.float32, "F")
        keep.append(md)
        fa = abi._HPGL_FLOAT_ARRAY(data=md.ctypes.data_as(_P(C.c_float)), shape=_shape(grid))
        lib.oracle_sgs_lvm_simulation(C.byref(arr), C.byref(p), c, C.byref(fa), sm)
    _check()
    return od, om


def sgs_simulation_batch(prop, grid, radiuses, max_neighbours, cov_model, seeds, kriging_type="sk"):
    """len(seeds) realisations of the C-level SGS (no cdf), one per OpenMP
    thread at a time; returns (values, masks) of shape (R, N), each row a
    Fortran-order property."""
    data, m = prop
    d, mk = _cont(data, m)
    R = len(seeds)
    od = np.tile(d.ravel(order="F"), (R, 1))
    om = np.tile(mk.ravel(order="F"), (R, 1))
    p = abi._HPGL_SGS_PARAMS(radiuses=(C.c_int * 3)(*radiuses), max_neighbours=max_neighbours,
                             kriging_kind={"sk": 1, "ok": 0}[kriging_type], seed=int(seeds[0]), min_neighbours=0,
                             **_cov_kwargs(cov_model))
    sd = np.ascontiguousarray(seeds, np.int64)
    lib.oracle_sgs_simulation_batch(_ptr(od), _ptr(om), C.byref(_shape(grid)), C.byref(p), _ptr(sd), R)
    _check()
    return od, om


def sis_simulation(prop, grid, data, seed, marginal_probs, use_correlogram=True, mask=None):
    """C-level SIS on (data, mask, indicator_count); marginal_probs are K
    scalars or K per-cell arrays (LVM).  Returns the simulated (data, mask)."""
    d, m, k = prop
    od = np.require(d, np.uint8, "F").copy(order="F")
    om = np.require(m, np.uint8, "F").copy(order="F")
    K = len(data)
    lvm = not np.isscalar(marginal_probs[0])
    ps = (abi._HPGL_IK_PARAMS * K)()
    for i, dd in enumerate(data):
        ps[i] = abi._HPGL_IK_PARAMS(radiuses=(C.c_int * 3)(*dd["radiuses"]), max_neighbours=dd["max_neighbours"],
                                    marginal_prob=0 if lvm else marginal_probs[i], **_cov_kwargs(dd["cov_model"]))
    keep = []
    sm = None
    if mask is not None:
        mk = np.require(mask, np.uint8, "F")
        keep.append(mk)
        sm = C.byref(abi._HPGL_UBYTE_ARRAY(data=mk.ctypes.data_as(_P(C.c_ubyte)), shape=_shape(grid)))
    arr = _ima(od, om, k, grid)
    if not lvm:
        lib.oracle_sis_simulation(C.byref(arr), ps, K, seed, sm)
    else:
        means = (abi._HPGL_FLOAT_ARRAY * K)()
        for i, mp in enumerate(marginal_probs):
            md = np.require(mp, np.float32, "F")
            keep.append(md)
            means[i] = abi._HPGL_FLOAT_ARRAY(data=md.ctypes.data_as(_P(C.c_float)), shape=_shape(grid))
        lib.oracle_sis_simulation_lvm(C.byref(arr), ps, means, K, seed, sm, 1 if use_correlogram else 0)
    _check()
    return od, om


def simple_cokriging_markI(prop, grid, secondary, primary_mean, secondary_mean, secondary_variance,
                           correlation_coef, radiuses, max_neighbours, cov_model):
    ind, inm = _cont(*prop)
    sd, sm = _cont(*secondary)
    od, om = ind.copy(order="F"), inm.copy(order="F")
    p = abi._HPGL_COKRIGING_M1_PARAMS(radiuses=(C.c_int * 3)(*radiuses), max_neighbours=max_neighbours,
                                      primary_mean=primary_mean, secondary_mean=secondary_mean,
                                      secondary_variance=secondary_variance, correlation_coef=correlation_coef,
                                      **_cov_kwargs(cov_model))
    lib.oracle_simple_cokriging_mark1(C.byref(_cma(ind, inm, grid)), C.byref(_cma(sd, sm, grid)), C.byref(p),
                                      C.byref(_cma(od, om, grid)))
    _check()
    return od, om


def simple_cokriging_markII(grid, primary, secondary, correlation_coef, radiuses, max_neighbours):
    """primary/secondary: dicts with "data" ((values, mask)), "mean", "cov_model"."""
    ind, inm = _cont(*primary["data"])
    sd, sm = _cont(*secondary["data"])
    od, om = ind.copy(order="F"), inm.copy(order="F")
    p = abi._HPGL_COKRIGING_M2_PARAMS(primary_cov_params=cov_params(primary["cov_model"]),
                                      secondary_cov_params=cov_params(secondary["cov_model"]),
                                      radiuses=(C.c_int * 3)(*radiuses), max_neighbours=max_neighbours,
                                      primary_mean=primary["mean"], secondary_mean=secondary["mean"],
                                      correlation_coef=correlation_coef)
    lib.oracle_simple_cokriging_mark2(C.byref(_cma(ind, inm, grid)), C.byref(_cma(sd, sm, grid)), C.byref(p),
                                      C.byref(_cma(od, om, grid)))
    _check()
    return od, om


def gauss_solve(a, b):
    a = np.ascontiguousarray(a, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    x = np.zeros_like(b)
    ok = lib.oracle_gauss_solve(len(b), _ptr(a), _ptr(b), _ptr(x))
    return bool(ok), x


def ellipsoid(R1, R2, R3, azimuth, dip, rotation):
    e = cvabi.ellipsoid_t(R1=R1, R2=R2, R3=R3)
    lib.oracle_fill_ellipsoid_directions(C.byref(e), azimuth, dip, rotation)
    return e


def variogram_template(lag_width, lag_separation, tol_distance, num_lags, first_lag_distance, ell):
    return cvabi.variogram_search_template_t(lag_width=lag_width, lag_separation=lag_separation,
                                             tol_distance=tol_distance, num_lags=num_lags,
                                             first_lag_distance=first_lag_distance, ellipsoid=ell)


def calc_variograms(templ, values, mask):
    """Gridded semivariogram of (values, mask) arrays (any strides)."""
    v = np.asarray(values, np.float32)
    m = np.asarray(mask, np.uint8)
    st = lambda a: (C.c_int * 3)(*[s // a.itemsize for s in a.strides])
    hd = cvabi.hard_data_t(data=v.ctypes.data_as(_P(C.c_float)), mask=m.ctypes.data_as(_P(C.c_ubyte)),
                           data_shape=(C.c_int * 3)(*v.shape), data_strides=st(v),
                           mask_shape=(C.c_int * 3)(*m.shape), mask_strides=st(m))
    out = np.zeros(templ.num_lags, np.float32)
    lib.oracle_calc_variograms(C.byref(templ), C.byref(hd), _ptr(out), out.size)
    _check()
    return out


def calc_variograms_from_point_set(templ, xs, ys, zs, values):
    cols = [np.ascontiguousarray(a, np.float32) for a in (xs, ys, zs, values)]
    ps = cvabi.cont_point_set_t(xs=cols[0].ctypes.data_as(_P(C.c_float)), ys=cols[1].ctypes.data_as(_P(C.c_float)),
                                zs=cols[2].ctypes.data_as(_P(C.c_float)),
                                values=cols[3].ctypes.data_as(_P(C.c_float)), values_count=cols[3].size)
    out = np.zeros(templ.num_lags, np.float32)
    lib.oracle_calc_variograms_from_point_set(C.byref(templ), C.byref(ps), _ptr(out), out.size)
    _check()
    return out


def stack_layers(layers, markers, nz, scalez, blank_value, result):
    fd = lambda a: cvabi.float_data_t(data=a.ctypes.data_as(_P(C.c_float)), data_shape=(C.c_int * 3)(*a.shape),
                                      data_strides=(C.c_int * 3)(*[s // a.itemsize for s in a.strides]))
    ls = (cvabi.float_data_t * len(layers))(*[fd(a) for a in layers])
    lib.oracle_stack_layers(ls, (C.c_int * len(markers))(*markers), len(layers), nz, scalez, blank_value,
                            C.byref(fd(result)))


def simple_kriging_weights(center, xs, ys, zs, cm):
    c = np.asarray(center, np.float32)
    x, y, z = (np.ascontiguousarray(a, np.float32) for a in (xs, ys, zs))
    w = np.zeros(len(x), np.float32)
    var = C.c_double()
    rc = lib.oracle_simple_kriging_weights(_ptr(c), _ptr(x), _ptr(y), _ptr(z), len(x), C.byref(cov_params(cm)), _ptr(w),
                                           C.byref(var))
    return rc, w, var.value


def ok_weights_int(center, coords, cm):
    c = np.asarray(center, np.int32)
    xyz = np.ascontiguousarray(coords, np.int32)
    w = np.zeros(len(xyz), np.float64)
    var = C.c_double()
    ok = lib.oracle_ok_weights_int(_ptr(c), _ptr(xyz), len(xyz), C.byref(cov_params(cm)), _ptr(w), C.byref(var))
    return bool(ok), w, var.value


def cholesky_solve(a, b):
    a = np.asfortranarray(a, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    x = np.zeros_like(b)
    f = np.zeros_like(a, order="F")
    info = lib.oracle_cholesky_solve(a.shape[0], _ptr(a), _ptr(b), _ptr(x), _ptr(f))
    return info, x, f


def covariance(cm, dx, dy, dz):
    return lib.oracle_covariance(C.byref(cov_params(cm)), dx, dy, dz)


def template(cm, radiuses, table):
    r = np.asarray(radiuses, np.int32)
    n = lib.oracle_template(C.byref(cov_params(cm)), _ptr(r), int(table), None, 0)
    out = np.zeros((max(n, 1), 3), np.int32)
    lib.oracle_template(C.byref(cov_params(cm)), _ptr(r), int(table), _ptr(out), n)
    return out[:n]


def sgs_path(size, seed, count):
    out = np.zeros(count, np.int32)
    lib.oracle_sgs_path(size, seed, _ptr(out), count)
    return out


def mt11213b(seed, count):
    out = np.zeros(count, np.uint32)
    lib.oracle_mt11213b(seed, _ptr(out), count)
    return out


def gauss_inverse(p):
    return lib.oracle_gauss_inverse(p)


def gauss_prob(x):
    return lib.oracle_gauss_prob(x)


def last_counts():
    """(nodes without neighbours, singular systems) of the last OK/SK/LVM call."""
    out = np.zeros(2, np.int64)
    lib.oracle_last_counts(_ptr(out))
    return int(out[0]), int(out[1])


def set_threads(n):
    lib.oracle_set_threads(n)


def set_sample_stride(stride):
    """Estimate only cells with cell % stride == 0 (CPU-baseline sampling)."""
    lib.oracle_set_sample_stride(int(stride))


def set_progress_tick(on):
    lib.oracle_set_progress_tick(1 if on else 0)
