"""Experimental variograms, a Python 3 drop-in for the reference's
``geo_bsd/cvariogram.py``, backed by hpgl_amd/_cvariogram.so.

Restated for Python 3 without changing behaviour, except where the
reference could not run at all:

* strides are integer element counts (``//``; the reference's ``/`` gives
  floats under Python 3);
* ``CalcVariogramsFromPointSet`` reads the values from
  ``point_set["Property"]`` and counts them (the reference calls
  ``point_set("Property")`` and reads an undefined ``values``);
* ``_create_float_data`` uses the module's own ``checked_create``;
* a failed call raises ``RuntimeError`` with the library's message.
"""
import ctypes as C
import os

import numpy

from ._cvar_abi import cont_point_set_t, ellipsoid_t, float_data_t, hard_data_t, variogram_search_template_t, vector_t

_LIB_PATH = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "_cvariogram.so")
if not os.path.exists(_LIB_PATH):
    raise ImportError("%s is missing: build it with python hpgl_amd/build.py" % _LIB_PATH)
cvar = C.CDLL(_LIB_PATH)

_f32 = numpy.ctypeslib.ndpointer(dtype=numpy.float32)

cvar.fill_ellipsoid_directions.restype = None
cvar.fill_ellipsoid_directions.argtypes = [C.POINTER(ellipsoid_t), C.c_double, C.c_double, C.c_double]
cvar.calc_variograms.restype = None
cvar.calc_variograms.argtypes = [C.POINTER(variogram_search_template_t), C.POINTER(hard_data_t), _f32, C.c_int,
                                 C.c_int]
cvar.calc_variograms_from_point_set.restype = None
cvar.calc_variograms_from_point_set.argtypes = [C.POINTER(variogram_search_template_t),
                                                C.POINTER(cont_point_set_t), _f32, C.c_int]
cvar.cvar_stack_layers.restype = None
cvar.cvar_stack_layers.argtypes = [C.POINTER(float_data_t), C.POINTER(C.c_int), C.c_int, C.c_int, C.c_float,
                                   C.c_int, C.POINTER(float_data_t)]
cvar.cvar_amd_last_error.restype = C.c_char_p
cvar.cvar_amd_last_error.argtypes = []
cvar.cvar_amd_last_stats.restype = None
cvar.cvar_amd_last_stats.argtypes = [C.POINTER(C.c_longlong), C.POINTER(C.c_longlong), C.POINTER(C.c_double)]


def _check():
    msg = cvar.cvar_amd_last_error()
    if msg:
        raise RuntimeError(msg.decode())


def checked_create(T, **kargs):
    fields = [f for f, _ in T._fields_]
    for k in kargs:
        if k in fields:
            fields.remove(k)
    assert len(fields) == 0, "No values for parameters: %s" % fields
    return T(**kargs)


def _strides(array):
    return tuple(s // array.itemsize for s in array.strides)


def _c_array(t, size, values):
    return (t * size)(*values)


class Ellipsoid:
    def __init__(self, R1, R2, R3, azimuth, dip, rotation):
        vec = checked_create(vector_t, data=_c_array(C.c_double, 3, (0, 0, 0)))
        self.ell = checked_create(ellipsoid_t, direction1=vec, direction2=vec, direction3=vec, R1=R1, R2=R2, R3=R3)
        cvar.fill_ellipsoid_directions(C.byref(self.ell), azimuth, dip, rotation)
        _check()


class VariogramSearchTemplate:
    def __init__(self, lag_width, lag_separation, tol_distance, num_lags, first_lag_distance, ellipsoid):
        self.templ = checked_create(
            variogram_search_template_t, lag_width=lag_width, lag_separation=lag_separation,
            tol_distance=tol_distance, num_lags=num_lags, first_lag_distance=first_lag_distance,
            ellipsoid=ellipsoid.ell)
        self.num_lags = num_lags
        self.lag_separation = lag_separation


def CalcVariograms(templ, hard_data, percent=100):
    """Semivariogram of gridded data (values, mask) (cvariogram.py:129-153);
    returns (lag borders, float32 variogram)."""
    variogram = numpy.zeros(templ.num_lags, dtype="float32")
    values, mask = hard_data[0], hard_data[1]
    hd = checked_create(
        hard_data_t,
        data=values.ctypes.data_as(C.POINTER(C.c_float)),
        mask=mask.ctypes.data_as(C.POINTER(C.c_ubyte)),
        data_shape=_c_array(C.c_int, 3, values.shape),
        mask_shape=_c_array(C.c_int, 3, mask.shape),
        data_strides=_c_array(C.c_int, 3, _strides(values)),
        mask_strides=_c_array(C.c_int, 3, _strides(mask)))
    cvar.calc_variograms(C.byref(templ.templ), C.byref(hd), variogram, variogram.size, percent)
    _check()
    lags_borders = numpy.array([k * templ.lag_separation for k in range(templ.num_lags)], dtype=float)
    return (lags_borders, variogram)


def CalcVariogramsFromPointSet(templ, point_set, variogram=None):
    """Semivariogram of a point set {"X", "Y", "Z", "Property"}
    (cvariogram.py:155-177); returns (lag borders, float32 variogram)."""
    variogram = numpy.zeros(templ.num_lags, dtype="float32")
    cols = [numpy.ascontiguousarray(point_set[k], dtype="float32") for k in ("X", "Y", "Z", "Property")]
    ps = checked_create(
        cont_point_set_t,
        xs=cols[0].ctypes.data_as(C.POINTER(C.c_float)),
        ys=cols[1].ctypes.data_as(C.POINTER(C.c_float)),
        zs=cols[2].ctypes.data_as(C.POINTER(C.c_float)),
        values=cols[3].ctypes.data_as(C.POINTER(C.c_float)),
        values_count=cols[3].size)
    cvar.calc_variograms_from_point_set(C.byref(templ.templ), C.byref(ps), variogram, variogram.size)
    _check()
    lags_borders = numpy.array([k * templ.lag_separation for k in range(templ.num_lags)], dtype=float)
    return (lags_borders, variogram)


def last_stats():
    """Offsets in the tunnel, candidate pairs tested and device ms of the
    last CalcVariograms call in this thread (no reference counterpart)."""
    o, p, ms = C.c_longlong(), C.c_longlong(), C.c_double()
    cvar.cvar_amd_last_stats(C.byref(o), C.byref(p), C.byref(ms))
    return {"offsets": o.value, "pairs_tested": p.value, "device_ms": ms.value}


def _create_float_data(array):
    return checked_create(
        float_data_t,
        data=array.ctypes.data_as(C.POINTER(C.c_float)),
        data_shape=_c_array(C.c_int, 3, array.shape),
        data_strides=_c_array(C.c_int, 3, _strides(array)))


def CStackLayers(layers, markers, nz, scalez, blank_value, result):
    """Fills the float32 cube ``result`` from layer thickness maps
    (cvariogram.py:186-198)."""
    layers2 = [_create_float_data(layer) for layer in layers]
    result2 = _create_float_data(result)
    cvar.cvar_stack_layers(_c_array(float_data_t, len(layers2), layers2), _c_array(C.c_int, len(markers), markers),
                           len(layers2), nz, scalez, blank_value, C.byref(result2))
    _check()
