"""ctypes binding of the MI355X kriging library (``hpgl_amd/hpgl.so``).

Mirrors the reference's ``geo_bsd/hpgl_wrap.py``: the same ``Structure``
mirrors of the C structs (hpgl_wrap.py:18-139 there, ``include/hpgl.h``
here) and the same argtypes (hpgl_wrap.py:148-311).  The library is loaded
at import, as in the reference (hpgl_wrap.py:143-146); a missing library is
an ImportError naming the build step, never a silent fallback.
"""
import ctypes as C
import os

import numpy

from ._abi import (  # noqa: F401
    hpgl_output_handler, hpgl_progress_handler, _HPGL_MEAN_KIND, _HPGL_KRIGING_KIND, _HPGL_SHAPE,
    _HPGL_CONT_MASKED_ARRAY, _HPGL_IND_MASKED_ARRAY, _HPGL_UBYTE_ARRAY, _HPGL_FLOAT_ARRAY, _HPGL_OK_PARAMS,
    _HPGL_SK_PARAMS, _HPGL_SGS_PARAMS, _HPGL_MEDIAN_IK_PARAMS, _HPGL_IK_PARAMS, _HPGL_COV_PARAMS,
    __hpgl_cov_params_t, hpgl_non_parametric_cdf_t, _HPGL_AMD_STATS, _HPGL_AMD_JOB, _HPGL_AMD_JOB_RESULT,
    _HPGL_COKRIGING_M1_PARAMS, _HPGL_COKRIGING_M2_PARAMS)

_PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBRARY_PATH = os.path.join(_PKG_DIR, "hpgl.so")


def _load():
    if not os.path.exists(LIBRARY_PATH):
        raise ImportError(
            "hpgl_amd: native library %s is missing; build it with "
            "`python -c 'import __graft_entry__; __graft_entry__.build()'` "
            "(or `python hpgl_amd/build.py`)" % LIBRARY_PATH
        )
    return C.CDLL(LIBRARY_PATH)


_hpgl_so = _load()

_f32 = numpy.ctypeslib.ndpointer(dtype=numpy.float32, flags="C_CONTIGUOUS")
_f32F = numpy.ctypeslib.ndpointer(dtype=numpy.float32, flags="F_CONTIGUOUS")
_u8F = numpy.ctypeslib.ndpointer(dtype=numpy.uint8, flags="F_CONTIGUOUS")
_P = C.POINTER

_hpgl_so.hpgl_get_last_exception_message.restype = C.c_char_p
_hpgl_so.hpgl_get_last_exception_message.argtypes = []
_hpgl_so.hpgl_set_output_handler.restype = None
_hpgl_so.hpgl_set_output_handler.argtypes = [hpgl_output_handler, C.py_object]
_hpgl_so.hpgl_set_progress_handler.restype = None
_hpgl_so.hpgl_set_progress_handler.argtypes = [hpgl_progress_handler, C.py_object]
_hpgl_so.hpgl_set_thread_num.restype = None
_hpgl_so.hpgl_set_thread_num.argtypes = [C.c_int]
_hpgl_so.hpgl_get_thread_num.restype = C.c_int
_hpgl_so.hpgl_get_thread_num.argtypes = []

# property files (hpgl_wrap.py:228-277)
_u8W = numpy.ctypeslib.ndpointer(dtype=numpy.uint8, flags=["F", "W", "A"])
_f32W = numpy.ctypeslib.ndpointer(dtype=numpy.float32, flags=["F", "W", "A"])
_hpgl_so.hpgl_read_inc_file_float.restype = None
_hpgl_so.hpgl_read_inc_file_float.argtypes = [C.c_char_p, C.c_float, C.c_int, _f32W, _u8W]
_hpgl_so.hpgl_read_inc_file_byte.restype = None
_hpgl_so.hpgl_read_inc_file_byte.argtypes = [C.c_char_p, C.c_int, C.c_int, _u8W, _u8W, _u8W, C.c_int]
_hpgl_so.hpgl_write_inc_file_float.restype = None
_hpgl_so.hpgl_write_inc_file_float.argtypes = [C.c_char_p, _P(_HPGL_CONT_MASKED_ARRAY), C.c_float, C.c_char_p]
_hpgl_so.hpgl_write_inc_file_byte.restype = None
_hpgl_so.hpgl_write_inc_file_byte.argtypes = [
    C.c_char_p, _P(_HPGL_IND_MASKED_ARRAY), C.c_int, C.c_char_p, _P(C.c_ubyte), C.c_int]
_hpgl_so.hpgl_write_gslib_cont_property.restype = None
_hpgl_so.hpgl_write_gslib_cont_property.argtypes = [_P(_HPGL_CONT_MASKED_ARRAY), C.c_char_p, C.c_char_p, C.c_double]
_hpgl_so.hpgl_write_gslib_byte_property.restype = None
_hpgl_so.hpgl_write_gslib_byte_property.argtypes = [
    _P(_HPGL_IND_MASKED_ARRAY), C.c_char_p, C.c_char_p, C.c_double, _P(C.c_ubyte), C.c_int]

_hpgl_so.hpgl_ordinary_kriging.restype = None
_hpgl_so.hpgl_ordinary_kriging.argtypes = [_P(_HPGL_CONT_MASKED_ARRAY), _P(_HPGL_OK_PARAMS), _P(_HPGL_CONT_MASKED_ARRAY)]

_hpgl_so.hpgl_simple_kriging.restype = None
_hpgl_so.hpgl_simple_kriging.argtypes = [
    _f32F, _u8F, _P(_HPGL_SHAPE), _P(_HPGL_SK_PARAMS), _f32F, _u8F, _P(_HPGL_SHAPE)]

_hpgl_so.hpgl_lvm_kriging.restype = None
_hpgl_so.hpgl_lvm_kriging.argtypes = [
    _f32F, _u8F, _P(_HPGL_SHAPE), _f32F, _P(_HPGL_SHAPE), _P(_HPGL_OK_PARAMS), _f32F, _u8F, _P(_HPGL_SHAPE)]

_hpgl_so.hpgl_simple_kriging_weights.restype = C.c_int
_hpgl_so.hpgl_simple_kriging_weights.argtypes = [
    _P(C.c_float), _f32, _f32, _f32, C.c_int, _P(__hpgl_cov_params_t), _f32]

_hpgl_so.hpgl_indicator_kriging.restype = None
_hpgl_so.hpgl_indicator_kriging.argtypes = [
    _P(_HPGL_IND_MASKED_ARRAY), _P(_HPGL_IND_MASKED_ARRAY), _P(_HPGL_IK_PARAMS), C.c_int]

_hpgl_so.hpgl_median_ik.restype = None
_hpgl_so.hpgl_median_ik.argtypes = [_P(_HPGL_IND_MASKED_ARRAY), _P(_HPGL_MEDIAN_IK_PARAMS), _P(_HPGL_IND_MASKED_ARRAY)]

_hpgl_so.hpgl_sgs_simulation.restype = None
_hpgl_so.hpgl_sgs_simulation.argtypes = [
    _P(_HPGL_CONT_MASKED_ARRAY), _P(_HPGL_SGS_PARAMS), _P(hpgl_non_parametric_cdf_t), _P(C.c_double),
    _P(_HPGL_UBYTE_ARRAY)]

_hpgl_so.hpgl_sgs_lvm_simulation.restype = None
_hpgl_so.hpgl_sgs_lvm_simulation.argtypes = [
    _P(_HPGL_CONT_MASKED_ARRAY), _P(_HPGL_SGS_PARAMS), _P(hpgl_non_parametric_cdf_t), _P(_HPGL_FLOAT_ARRAY),
    _P(_HPGL_UBYTE_ARRAY)]

_hpgl_so.hpgl_amd_sgs_simulation_batch.restype = None
_hpgl_so.hpgl_amd_sgs_simulation_batch.argtypes = [
    _P(_HPGL_CONT_MASKED_ARRAY), _P(_HPGL_SGS_PARAMS), _P(hpgl_non_parametric_cdf_t), _P(_HPGL_FLOAT_ARRAY),
    _P(_HPGL_UBYTE_ARRAY), _P(C.c_long), C.c_int, C.c_void_p, C.c_void_p]
_hpgl_so.hpgl_amd_sgs_simulation_batch_dev.restype = None
_hpgl_so.hpgl_amd_sgs_simulation_batch_dev.argtypes = [
    C.c_void_p, C.c_void_p, _P(_HPGL_SHAPE), _P(_HPGL_SGS_PARAMS), _P(hpgl_non_parametric_cdf_t), C.c_void_p,
    C.c_void_p, _P(C.c_long), C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]

_hpgl_so.hpgl_sis_simulation.restype = None
_hpgl_so.hpgl_sis_simulation.argtypes = [
    _P(_HPGL_IND_MASKED_ARRAY), _P(_HPGL_IK_PARAMS), C.c_int, C.c_int, _P(_HPGL_UBYTE_ARRAY)]
_hpgl_so.hpgl_sis_simulation_lvm.restype = None
_hpgl_so.hpgl_sis_simulation_lvm.argtypes = [
    _P(_HPGL_IND_MASKED_ARRAY), _P(_HPGL_IK_PARAMS), _P(_HPGL_FLOAT_ARRAY), C.c_int, C.c_int,
    _P(_HPGL_UBYTE_ARRAY), C.c_int]
_hpgl_so.hpgl_amd_sis_simulation_batch.restype = None
_hpgl_so.hpgl_amd_sis_simulation_batch.argtypes = [
    _P(_HPGL_IND_MASKED_ARRAY), _P(_HPGL_IK_PARAMS), C.c_int, _P(_HPGL_FLOAT_ARRAY), C.c_int,
    _P(_HPGL_UBYTE_ARRAY), _P(C.c_int), C.c_int, C.c_void_p, C.c_void_p]

_hpgl_so.hpgl_amd_sis_simulation_batch_dev.restype = None
_hpgl_so.hpgl_amd_sis_simulation_batch_dev.argtypes = [
    C.c_void_p, C.c_void_p, _P(_HPGL_SHAPE), _P(_HPGL_IK_PARAMS), C.c_int, C.c_void_p, C.c_int, C.c_void_p,
    _P(C.c_int), C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]

_hpgl_so.hpgl_simple_cokriging_mark1.restype = None
_hpgl_so.hpgl_simple_cokriging_mark1.argtypes = [
    _P(_HPGL_CONT_MASKED_ARRAY), _P(_HPGL_CONT_MASKED_ARRAY), _P(_HPGL_COKRIGING_M1_PARAMS),
    _P(_HPGL_CONT_MASKED_ARRAY)]
_hpgl_so.hpgl_simple_cokriging_mark2.restype = None
_hpgl_so.hpgl_simple_cokriging_mark2.argtypes = [
    _P(_HPGL_CONT_MASKED_ARRAY), _P(_HPGL_CONT_MASKED_ARRAY), _P(_HPGL_COKRIGING_M2_PARAMS),
    _P(_HPGL_CONT_MASKED_ARRAY)]
_hpgl_so.hpgl_amd_simple_cokriging_mark1_dev.restype = None
_hpgl_so.hpgl_amd_simple_cokriging_mark1_dev.argtypes = [
    C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, _P(_HPGL_SHAPE), _P(_HPGL_COKRIGING_M1_PARAMS), C.c_void_p,
    C.c_void_p, C.c_void_p]
_hpgl_so.hpgl_amd_simple_cokriging_mark2_dev.restype = None
_hpgl_so.hpgl_amd_simple_cokriging_mark2_dev.argtypes = [
    C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, _P(_HPGL_SHAPE), _P(_HPGL_COKRIGING_M2_PARAMS), C.c_void_p,
    C.c_void_p, C.c_void_p]

_hpgl_so.hpgl_amd_server_create.restype = C.c_void_p
_hpgl_so.hpgl_amd_server_create.argtypes = [C.c_int, C.c_longlong]
_hpgl_so.hpgl_amd_server_submit.restype = C.c_longlong
_hpgl_so.hpgl_amd_server_submit.argtypes = [C.c_void_p, _P(_HPGL_AMD_JOB)]
_hpgl_so.hpgl_amd_server_wait.restype = C.c_int
_hpgl_so.hpgl_amd_server_wait.argtypes = [C.c_void_p, C.c_longlong, _P(_HPGL_AMD_JOB_RESULT)]
_hpgl_so.hpgl_amd_server_destroy.restype = None
_hpgl_so.hpgl_amd_server_destroy.argtypes = [C.c_void_p]

_hpgl_so.hpgl_amd_set_precision.restype = None
_hpgl_so.hpgl_amd_set_precision.argtypes = [C.c_int]
_hpgl_so.hpgl_amd_get_precision.restype = C.c_int
_hpgl_so.hpgl_amd_get_precision.argtypes = []
_hpgl_so.hpgl_amd_get_last_error.restype = C.c_int
_hpgl_so.hpgl_amd_get_last_error.argtypes = []
_hpgl_so.hpgl_amd_get_last_stats.restype = None
_hpgl_so.hpgl_amd_get_last_stats.argtypes = [_P(_HPGL_AMD_STATS)]
_hpgl_so.hpgl_amd_simple_kriging_weights_noalloc.restype = C.c_int
_hpgl_so.hpgl_amd_simple_kriging_weights_noalloc.argtypes = [C.c_void_p] * 4 + [C.c_int, _P(_HPGL_COV_PARAMS),
                                                                               C.c_void_p, C.c_void_p]
_hpgl_so.hpgl_amd_pool_selftest.restype = C.c_int
_hpgl_so.hpgl_amd_pool_selftest.argtypes = []
_hpgl_so.hpgl_amd_set_neighbour_dump.restype = None
_hpgl_so.hpgl_amd_set_neighbour_dump.argtypes = [C.c_void_p, C.c_void_p, C.c_longlong, C.c_int]


class HpglError(RuntimeError):
    """A call into hpgl.so failed; the message is the library's."""


def check_last_error():
    """Raise HpglError if the last library call on this thread failed."""
    if _hpgl_so.hpgl_amd_get_last_error():
        msg = _hpgl_so.hpgl_get_last_exception_message()
        raise HpglError(msg.decode("utf-8", "replace") if msg else "hpgl error")


def last_stats():
    """Statistics of the last kriging/simulation call, as a dict."""
    s = _HPGL_AMD_STATS()
    _hpgl_so.hpgl_amd_get_last_stats(C.byref(s))
    return {name: getattr(s, name) for name, _ in _HPGL_AMD_STATS._fields_}


def set_precision(mode):
    """'mixed' (FP32 + FP64 refinement, default) or 'fp64'."""
    _hpgl_so.hpgl_amd_set_precision({"mixed": 0, "fp64": 1}[mode])


def get_precision():
    return {0: "mixed", 1: "fp64"}[_hpgl_so.hpgl_amd_get_precision()]
