"""Kriging API, a Python 3 drop-in for the reference's ``geo_bsd/geo.py``.

Same names, arguments, argument meaning and return values as the
reference (geo.py:85-635); the work runs in ``hpgl_amd/hpgl.so`` on the GPU.
Python 2 constructs of the reference are restated for Python 3 without
changing what the functions do:

* ``__get_strides`` used ``/`` (geo.py:21-24) and yields integers here;
* ``accepts_tuple`` used ``dict.has_key`` (geo.py:266);
* ``set_progress_handler(None)`` cleared the *output* handler
  (geo.py:628-632); here it clears the progress handler.

* file names and property names are passed to the library as bytes.

Property files (``load_cont_property``, ``write_property`` ...) are read and
written by the library's host code (INC and GSLIB formats, geo.py:168-362).
"""
import ctypes as C
import os

import numpy

from .hpgl_wrap import (
    _HPGL_SHAPE,
    _HPGL_CONT_MASKED_ARRAY,
    _HPGL_IND_MASKED_ARRAY,
    _HPGL_UBYTE_ARRAY,
    _HPGL_FLOAT_ARRAY,
    _HPGL_OK_PARAMS,
    _HPGL_SK_PARAMS,
    _HPGL_IK_PARAMS,
    _HPGL_MEDIAN_IK_PARAMS,
    _HPGL_COV_PARAMS,
    _HPGL_COKRIGING_M1_PARAMS,
    _HPGL_COKRIGING_M2_PARAMS,
    _hpgl_so,
    check_last_error,
    hpgl_output_handler,
    hpgl_progress_handler,
)


def _c_array(ar_type, size, values):
    if len(values) != size:
        raise RuntimeError("%s values specified for array of %s elements" % (len(values), size))
    return (ar_type * size)(*values)


def _create_hpgl_shape(shape, strides=None):
    if strides is None:
        strides = (1, shape[0], shape[0] * shape[1])
    return _HPGL_SHAPE(data=_c_array(C.c_int, 3, shape), strides=_c_array(C.c_int, 3, strides))


def _get_strides(a):
    return (a.strides[0] // a.itemsize, a.strides[1] // a.itemsize, a.strides[2] // a.itemsize)


def _checked_create(T, **kargs):
    fields = [f for f, _ in T._fields_]
    for k in kargs:
        if k in fields:
            fields.remove(k)
    assert len(fields) == 0, "No values for parameters: %s" % fields
    return T(**kargs)


def checkFWA(a):
    """Checks for fortran-order, writable and aligned flags."""
    assert a.flags["F"] and a.flags["W"] and a.flags["A"]


def _grid_shape(prop_data, grid):
    if grid is None:
        return _create_hpgl_shape(prop_data.shape, _get_strides(prop_data))
    if grid.x * grid.y * grid.z != prop_data.size:
        raise RuntimeError(
            "Invalid data size. Size of data = %s. Size of grid = %s" % (prop_data.size, grid.x * grid.y * grid.z))
    return _create_hpgl_shape((grid.x, grid.y, grid.z))


def _create_hpgl_cont_masked_array(prop, grid):
    return _HPGL_CONT_MASKED_ARRAY(
        data=prop.data.ctypes.data_as(C.POINTER(C.c_float)),
        mask=prop.mask.ctypes.data_as(C.POINTER(C.c_ubyte)),
        shape=_grid_shape(prop.data, grid))


def _create_hpgl_ind_masked_array(prop, grid):
    if grid is None:
        assert prop.data.strides == prop.mask.strides
    return _HPGL_IND_MASKED_ARRAY(
        data=prop.data.ctypes.data_as(C.POINTER(C.c_ubyte)),
        mask=prop.mask.ctypes.data_as(C.POINTER(C.c_ubyte)),
        shape=_grid_shape(prop.data, grid),
        indicator_count=prop.indicator_count)


def _create_hpgl_ubyte_array(array, grid):
    checkFWA(array)
    return _HPGL_UBYTE_ARRAY(data=array.ctypes.data_as(C.POINTER(C.c_ubyte)), shape=_grid_shape(array, grid))


def _create_hpgl_float_array(array, grid):
    checkFWA(array)
    return _HPGL_FLOAT_ARRAY(data=array.ctypes.data_as(C.POINTER(C.c_float)), shape=_grid_shape(array, grid))


class ContProperty:
    def __init__(self, data, mask):
        self.data = numpy.require(data, "float32", "F")
        self.mask = numpy.require(mask, "uint8", "F")

    def validate(self):
        checkFWA(self.data)
        checkFWA(self.mask)
        assert self.data.shape == self.mask.shape

    def fix_shape(self, grid):
        n = grid.x * grid.y * grid.z
        if self.data.ndim != 3 and self.data.size == n:
            self.data = self.data.reshape((grid.x, grid.y, grid.z), order="F")
        if self.mask.ndim != 3 and self.mask.size == n:
            self.mask = self.mask.reshape((grid.x, grid.y, grid.z), order="F")

    def __getitem__(self, idx):
        if idx == 0:
            return self.data
        elif idx == 1:
            return self.mask
        raise RuntimeError("Index out of range.")


class IndProperty:
    def __init__(self, data, mask, indicator_count):
        self.data = numpy.require(data, "uint8", "F")
        self.mask = numpy.require(mask, "uint8", "F")
        self.indicator_count = indicator_count
        # the reference's range check (geo.py:114); on uint8 inputs the
        # converted Fortran-order arrays hold the same values, and checking
        # them avoids strided passes over differently ordered arrays
        d0, m0 = numpy.asarray(data), numpy.asarray(mask)
        d, m = (self.data, self.mask) if (d0.dtype == numpy.uint8 and m0.dtype == numpy.uint8) else (d0, m0)
        if numpy.any(numpy.logical_and(m > 0, d >= indicator_count)):
            raise RuntimeError("Property contains some indicators outside of [0..%s] range." % (indicator_count - 1))
        assert numpy.shape(data) == numpy.shape(mask)

    def validate(self):
        checkFWA(self.data)
        checkFWA(self.mask)
        assert self.data.shape == self.mask.shape

    def __getitem__(self, idx):
        if idx == 0:
            return self.data
        elif idx == 1:
            return self.mask
        elif idx == 2:
            return self.indicator_count
        raise RuntimeError("Index out of range.")


def _prop_to_tuple_(prop):
    if isinstance(prop, ContProperty):
        return (prop.data, prop.mask)
    elif isinstance(prop, IndProperty):
        return (prop.data, prop.mask, prop.indicator_count)
    assert False
    return prop


def append_mask(prop, mask):
    infs = prop[1]
    infs &= mask
    infs.choose(-99, prop[0], out=prop[0])


class covariance:
    spherical = 0
    exponential = 1
    gaussian = 2


class SugarboxGrid:
    def __init__(self, x, y, z):
        self.x = x
        self.y = y
        self.z = z


class CovarianceModel:
    def __init__(self, type=0, ranges=(0, 0, 0), angles=(0.0, 0.0, 0.0), sill=1.0, nugget=0.0):
        self.type = type
        self.ranges = ranges
        self.angles = angles
        self.sill = sill
        self.nugget = nugget
        if nugget > sill:
            raise Exception("Nugget should be less or equal to Sill")


def _create_cont_prop(size):
    return ContProperty(numpy.zeros(size, dtype="float32"), numpy.zeros(size, dtype="uint8"))


def _create_ind_prop(size, indicator_count):
    return IndProperty(numpy.zeros(size, dtype="uint8"), numpy.zeros(size, dtype="uint8"), indicator_count)


def _empty_clone(prop):
    data2 = prop.data.copy("F")
    data2.fill(0)
    mask2 = prop.mask.copy("F")
    mask2.fill(0)
    if isinstance(prop, IndProperty):
        return IndProperty(data2, mask2, prop.indicator_count)
    elif isinstance(prop, ContProperty):
        return ContProperty(data2, mask2)
    assert False


def _clone_prop(prop):
    data2 = prop.data.copy("F")
    mask2 = prop.mask.copy("F")
    if isinstance(prop, IndProperty):
        # a copy of a checked property: the range check need not run again
        out = IndProperty.__new__(IndProperty)
        out.data, out.mask, out.indicator_count = data2, mask2, prop.indicator_count
        return out
    elif isinstance(prop, ContProperty):
        return ContProperty(data2, mask2)
    assert False, "Unknown prop type"


def _require_cont_data(data):
    if data is None:
        return None
    return numpy.require(data, dtype="float32", requirements="F")


def _requite_ind_data(data):
    if data is None:
        return None
    return numpy.require(data, dtype="uint8", requirements="F")


def accepts_tuple(arg_name, arg_pos):
    def tuple_to_prop(t):
        if isinstance(t, tuple):
            if len(t) == 3:
                return IndProperty(*t)
            elif len(t) == 2:
                return ContProperty(*t)
            assert False
        assert isinstance(t, (ContProperty, IndProperty))
        return t

    def decorator(f):
        def new_f(*args, **kargs):
            if arg_name in kargs:
                kargs[arg_name] = tuple_to_prop(kargs[arg_name])
            elif len(args) > arg_pos:
                args = args[:arg_pos] + (tuple_to_prop(args[arg_pos]),) + args[arg_pos + 1:]
            else:
                assert False
            return f(*args, **kargs)

        new_f.__name__ = f.__name__
        new_f.__doc__ = f.__doc__
        return new_f

    return decorator


def _bytes(s):
    return s if isinstance(s, bytes) else os.fsencode(s)


def _load_prop_cont_slow(filename, undefined_value):
    """geo.py:168-186: every token that parses as a float, skipping lines
    that start with "--"."""
    values = []
    mask = []
    with open(filename) as f:
        for line in f:
            if line.strip().startswith("--"):
                continue
            for part in line.split():
                try:
                    val = float(part.strip())
                except ValueError:
                    continue
                values.append(val)
                mask.append(0 if val == undefined_value else 1)
    return ContProperty(numpy.array(values, dtype="float32"), numpy.array(mask, dtype="uint8"))


def _load_prop_ind_slow(filename, undefined_value, ind_values):
    """geo.py:188-212: integer tokens mapped to their index in ind_values;
    the undefined value gives 255 with mask 0."""
    dict_map = {v: i for i, v in enumerate(ind_values)}
    values = []
    mask = []
    with open(filename) as f:
        for line in f:
            if line.strip().startswith("--"):
                continue
            for part in line.split():
                try:
                    val = int(part.strip())
                    if val == undefined_value:
                        values.append(255)
                        mask.append(0)
                    else:
                        values.append(dict_map[val])
                        mask.append(1)
                except (ValueError, KeyError):
                    pass
    return IndProperty(numpy.array(values, dtype="uint8", order="F"), numpy.array(mask, dtype="uint8", order="F"),
                       len(ind_values))


@accepts_tuple("prop", 0)
def write_property(prop, filename, prop_name, undefined_value, indicator_values=[]):
    """geo.py:277-305: INC file, one value per line, "/" at the end."""
    if prop.data.ndim == 3:
        sh = _create_hpgl_shape(prop.data.shape)
    else:
        sh = _create_hpgl_shape((prop.data.size, 1, 1))
    if isinstance(prop, ContProperty):
        marr = _HPGL_CONT_MASKED_ARRAY(data=prop.data.ctypes.data_as(C.POINTER(C.c_float)),
                                       mask=prop.mask.ctypes.data_as(C.POINTER(C.c_ubyte)), shape=sh)
        _hpgl_so.hpgl_write_inc_file_float(_bytes(filename), C.byref(marr), undefined_value, _bytes(prop_name))
    else:
        marr = _HPGL_IND_MASKED_ARRAY(data=prop.data.ctypes.data_as(C.POINTER(C.c_ubyte)),
                                      mask=prop.mask.ctypes.data_as(C.POINTER(C.c_ubyte)), shape=sh,
                                      indicator_count=prop.indicator_count)
        vals = numpy.array(indicator_values, dtype=numpy.uint8)
        _hpgl_so.hpgl_write_inc_file_byte(_bytes(filename), C.byref(marr), undefined_value, _bytes(prop_name),
                                          vals.ctypes.data_as(C.POINTER(C.c_ubyte)), len(indicator_values))
    check_last_error()


@accepts_tuple("prop", 0)
def write_gslib_property(prop, filename, prop_name, undefined_value, indicator_values=[]):
    """geo.py:307-322: GSLIB file with one variable."""
    if isinstance(prop, ContProperty):
        _hpgl_so.hpgl_write_gslib_cont_property(C.byref(_create_hpgl_cont_masked_array(prop, None)), _bytes(filename),
                                                _bytes(prop_name), undefined_value)
    else:
        _hpgl_so.hpgl_write_gslib_byte_property(
            C.byref(_create_hpgl_ind_masked_array(prop, None)), _bytes(filename), _bytes(prop_name), undefined_value,
            _c_array(C.c_ubyte, len(indicator_values), indicator_values), len(indicator_values))
    check_last_error()


def load_cont_property(filename, undefined_value, size=None):
    """geo.py:324-329: without a size the pure-Python reader is used."""
    if size is None:
        return _load_prop_cont_slow(filename, undefined_value)
    return read_inc_file_float(filename, undefined_value, size)


def read_inc_file_float(filename, undefined_value, size):
    """geo.py:331-342."""
    data = numpy.zeros(size, dtype="float32", order="F")
    mask = numpy.zeros(size, dtype="uint8", order="F")
    _hpgl_so.hpgl_read_inc_file_float(_bytes(filename), undefined_value, size[0] * size[1] * size[2], data, mask)
    check_last_error()
    return ContProperty(data, mask)


def read_inc_file_byte(filename, undefined_value, size, indicator_values):
    """geo.py:344-355: values are replaced by their index in
    indicator_values."""
    data = numpy.zeros(size, dtype="uint8", order="F")
    mask = numpy.zeros(size, dtype="uint8", order="F")
    _hpgl_so.hpgl_read_inc_file_byte(_bytes(filename), undefined_value, size[0] * size[1] * size[2], data, mask,
                                     numpy.array(indicator_values, dtype="uint8"), len(indicator_values))
    check_last_error()
    return IndProperty(data, mask, len(indicator_values))


def load_ind_property(filename, undefined_value, indicator_values, size=None):
    """geo.py:357-362."""
    if size is None:
        return _load_prop_ind_slow(filename, undefined_value, indicator_values)
    return read_inc_file_byte(filename, undefined_value, size, indicator_values)


def get_gslib_property(prop_dict, prop_name, undefined_value):
    """geo.py:608-616: (values, informed) from a dict of arrays."""
    prop = prop_dict[prop_name]
    informed_array = (prop != undefined_value).astype(numpy.uint8)
    return (prop_dict[prop_name], informed_array)


def set_thread_num(num):
    _hpgl_so.hpgl_set_thread_num(num)


def get_thread_num():
    return _hpgl_so.hpgl_get_thread_num()


@accepts_tuple("prop", 0)
def calc_mean(prop):
    m = prop.mask.ravel(order="F") == 1
    return float(numpy.sum(prop.data.ravel(order="F")[m].astype(numpy.float64)) / numpy.count_nonzero(m))


def _cov_kwargs(cov_model):
    return dict(
        covariance_type=cov_model.type,
        ranges=_c_array(C.c_double, 3, cov_model.ranges),
        angles=_c_array(C.c_double, 3, cov_model.angles),
        sill=cov_model.sill,
        nugget=cov_model.nugget)


@accepts_tuple("prop", 0)
def ordinary_kriging(prop, grid, radiuses, max_neighbours, cov_model):
    """Ordinary kriging of every uninformed cell (geo.py:381-399)."""
    out_prop = _clone_prop(prop)
    okp = _HPGL_OK_PARAMS(
        radiuses=_c_array(C.c_int, 3, radiuses), max_neighbours=max_neighbours, **_cov_kwargs(cov_model))
    _hpgl_so.hpgl_ordinary_kriging(
        C.byref(_create_hpgl_cont_masked_array(prop, grid)), C.byref(okp),
        C.byref(_create_hpgl_cont_masked_array(out_prop, grid)))
    check_last_error()
    return out_prop


@accepts_tuple("prop", 0)
def simple_kriging(prop, grid, radiuses, max_neighbours, cov_model, mean=None):
    """Simple kriging; mean=None computes the mean of the informed cells (geo.py:401-425)."""
    out_prop = _clone_prop(prop)
    skp = _HPGL_SK_PARAMS(
        radiuses=_c_array(C.c_int, 3, radiuses), max_neighbours=max_neighbours,
        automatic_mean=(mean is None), mean=(mean if mean is not None else 0), **_cov_kwargs(cov_model))
    sh = _HPGL_SHAPE(data=(C.c_int * 3)(grid.x, grid.y, grid.z))
    _hpgl_so.hpgl_simple_kriging(
        prop.data, prop.mask, C.byref(sh), C.byref(skp), out_prop.data, out_prop.mask, C.byref(sh))
    check_last_error()
    return out_prop


@accepts_tuple("prop", 0)
def lvm_kriging(prop, grid, mean_data, radiuses, max_neighbours, cov_model):
    """Simple kriging with a locally varying mean (geo.py:427-450)."""
    out_prop = _clone_prop(prop)
    okp = _HPGL_OK_PARAMS(
        radiuses=_c_array(C.c_int, 3, radiuses), max_neighbours=max_neighbours, **_cov_kwargs(cov_model))
    sh = _HPGL_SHAPE(data=(C.c_int * 3)(grid.x, grid.y, grid.z))
    mean_data = numpy.require(mean_data, "float32", "F")
    _hpgl_so.hpgl_lvm_kriging(
        prop.data, prop.mask, C.byref(sh), mean_data, C.byref(sh), C.byref(okp),
        out_prop.data, out_prop.mask, C.byref(sh))
    check_last_error()
    return out_prop


@accepts_tuple("prop", 0)
def median_ik(prop, grid, marginal_probs, radiuses, max_neighbours, cov_model):
    """Median indicator kriging for two indicators (geo.py:452-469)."""
    out_prop = _clone_prop(prop)
    miksp = _HPGL_MEDIAN_IK_PARAMS(
        radiuses=_c_array(C.c_int, 3, radiuses), max_neighbours=max_neighbours,
        marginal_probs=_c_array(C.c_double, 2, marginal_probs), **_cov_kwargs(cov_model))
    inp = _create_hpgl_ind_masked_array(prop, grid)
    outp = _create_hpgl_ind_masked_array(out_prop, grid)
    _hpgl_so.hpgl_median_ik(C.byref(inp), C.byref(miksp), C.byref(outp))
    check_last_error()
    return out_prop


def _create_hpgl_ik_params(data, indicator_count, is_lvm, marginal_probs):
    assert len(data) == indicator_count
    ikps = []
    for i in range(indicator_count):
        ikd = data[i]
        ikps.append(_checked_create(
            _HPGL_IK_PARAMS,
            radiuses=_c_array(C.c_int, 3, ikd["radiuses"]),
            max_neighbours=ikd["max_neighbours"],
            marginal_prob=0 if is_lvm else marginal_probs[i],
            **_cov_kwargs(ikd["cov_model"])))
    return _c_array(_HPGL_IK_PARAMS, indicator_count, ikps)


@accepts_tuple("prop", 0)
def indicator_kriging(prop, grid, data, marginal_probs):
    """Indicator kriging; two indicators go to median IK (geo.py:489-509).

    ``data`` is a list of dicts with keys cov_model, radiuses,
    max_neighbours; like the reference, 'marginal_prob' is written into them.
    """
    for i in range(len(data)):
        data[i]["marginal_prob"] = marginal_probs[i]
    if len(data) == 2:
        return median_ik(
            prop, grid, (data[0]["marginal_prob"], 1 - data[0]["marginal_prob"]),
            data[0]["radiuses"], data[0]["max_neighbours"], data[0]["cov_model"])
    out_prop = _clone_prop(prop)
    _hpgl_so.hpgl_indicator_kriging(
        C.byref(_create_hpgl_ind_masked_array(prop, grid)),
        C.byref(_create_hpgl_ind_masked_array(out_prop, grid)),
        _create_hpgl_ik_params(data, len(data), False, marginal_probs),
        len(data))
    check_last_error()
    return out_prop


@accepts_tuple("prop", 0)
def simple_cokriging_markI(prop, grid, secondary_data, primary_mean, secondary_mean, secondary_variance,
                           correlation_coef, radiuses, max_neighbours, cov_model):
    """Simple colocated cokriging, Markov model I (geo.py:511-534).

    secondary_data: ContProperty (or (data, mask)) on the same grid; cells
    where it is not informed use secondary_mean.  Every cell of the result
    is informed.
    """
    if isinstance(secondary_data, tuple):
        secondary_data = ContProperty(*secondary_data)
    out_prop = _clone_prop(prop)
    p = _checked_create(
        _HPGL_COKRIGING_M1_PARAMS, radiuses=_c_array(C.c_int, 3, radiuses), max_neighbours=max_neighbours,
        primary_mean=primary_mean, secondary_mean=secondary_mean, secondary_variance=secondary_variance,
        correlation_coef=correlation_coef, **_cov_kwargs(cov_model))
    _hpgl_so.hpgl_simple_cokriging_mark1(
        C.byref(_create_hpgl_cont_masked_array(prop, grid)),
        C.byref(_create_hpgl_cont_masked_array(secondary_data, grid)), C.byref(p),
        C.byref(_create_hpgl_cont_masked_array(out_prop, grid)))
    check_last_error()
    return out_prop


def simple_cokriging_markII(grid, primary_data, secondary_data, correlation_coef, radiuses, max_neighbours):
    """Simple colocated cokriging, Markov model II (geo.py:536-572).

    primary_data, secondary_data: dicts with keys "data" (ContProperty),
    "mean" and "cov_model".
    """
    out_prop = _clone_prop(primary_data["data"])
    pcp = primary_data["cov_model"]
    scp = secondary_data["cov_model"]
    p = _checked_create(
        _HPGL_COKRIGING_M2_PARAMS,
        primary_cov_params=_checked_create(_HPGL_COV_PARAMS, **_cov_kwargs(pcp)),
        secondary_cov_params=_checked_create(_HPGL_COV_PARAMS, **_cov_kwargs(scp)),
        radiuses=_c_array(C.c_int, 3, radiuses), max_neighbours=max_neighbours,
        primary_mean=primary_data["mean"], secondary_mean=secondary_data["mean"],
        correlation_coef=correlation_coef)
    _hpgl_so.hpgl_simple_cokriging_mark2(
        C.byref(_create_hpgl_cont_masked_array(primary_data["data"], grid)),
        C.byref(_create_hpgl_cont_masked_array(secondary_data["data"], grid)), C.byref(p),
        C.byref(_create_hpgl_cont_masked_array(out_prop, grid)))
    check_last_error()
    return out_prop


def simple_kriging_weights(center_point, n_x, n_y, n_z, ranges=(100000, 100000, 100000), sill=1,
                           cov_type=covariance.exponential, nugget=None, angles=None):
    """SK weights at arbitrary real coordinates (geo.py:574-606)."""
    if angles is None:
        angles = (0, 0, 0)
    if nugget is None:
        nugget = 0
    covp = C.byref(_checked_create(
        _HPGL_COV_PARAMS, covariance_type=cov_type, ranges=_c_array(C.c_double, 3, ranges),
        angles=_c_array(C.c_double, 3, angles), sill=sill, nugget=nugget))
    if len(n_x) != len(n_y) or len(n_x) != len(n_z):
        raise RuntimeError("Invalid pointset. %s,%s,%s." % (len(n_x), len(n_y), len(n_z)))
    if nugget > sill:
        raise RuntimeError("Nugget value must be less or equal to Sill value.")
    weights = numpy.zeros(len(n_x), dtype="float32")
    _hpgl_so.hpgl_simple_kriging_weights(
        _c_array(C.c_float, 3, center_point),
        numpy.ascontiguousarray(n_x, dtype="float32"),
        numpy.ascontiguousarray(n_y, dtype="float32"),
        numpy.ascontiguousarray(n_z, dtype="float32"),
        len(n_x), covp, weights)
    check_last_error()
    return weights


_handlers = {}


def set_output_handler(handler, param):
    if handler is None:
        _hpgl_so.hpgl_set_output_handler(hpgl_output_handler(), None)
        _handlers.pop("output", None)
    else:
        h = hpgl_output_handler(handler)
        _handlers["output"] = h
        _hpgl_so.hpgl_set_output_handler(h, param)


def set_progress_handler(handler, param):
    if handler is None:
        _hpgl_so.hpgl_set_progress_handler(hpgl_progress_handler(), None)
        _handlers.pop("progress", None)
    else:
        h = hpgl_progress_handler(handler)
        _handlers["progress"] = h
        _hpgl_so.hpgl_set_progress_handler(h, param)
