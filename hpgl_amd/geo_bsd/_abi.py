"""ctypes mirrors of the C structs declared in include/hpgl.h.

Same field names and order as the reference's geo_bsd/hpgl_wrap.py:6-139,
kept in a module of their own so that code which only needs the layouts
(the oracle binding of the tests) does not load the GPU library.
"""
import ctypes as C

hpgl_output_handler = C.CFUNCTYPE(C.c_int, C.c_char_p, C.py_object)
hpgl_progress_handler = C.CFUNCTYPE(C.c_int, C.c_char_p, C.c_int, C.py_object)


class _HPGL_MEAN_KIND:
    stationary_auto = 0
    stationary = 1
    varying = 2


class _HPGL_KRIGING_KIND:
    ordinary = 0
    simple = 1


class _HPGL_SHAPE(C.Structure):
    _fields_ = [("data", C.c_int * 3), ("strides", C.c_int * 3)]


class _HPGL_CONT_MASKED_ARRAY(C.Structure):
    _fields_ = [("data", C.POINTER(C.c_float)), ("mask", C.POINTER(C.c_ubyte)), ("shape", _HPGL_SHAPE)]


class _HPGL_IND_MASKED_ARRAY(C.Structure):
    _fields_ = [
        ("data", C.POINTER(C.c_ubyte)),
        ("mask", C.POINTER(C.c_ubyte)),
        ("shape", _HPGL_SHAPE),
        ("indicator_count", C.c_int),
    ]


class _HPGL_UBYTE_ARRAY(C.Structure):
    _fields_ = [("data", C.POINTER(C.c_ubyte)), ("shape", _HPGL_SHAPE)]


class _HPGL_FLOAT_ARRAY(C.Structure):
    _fields_ = [("data", C.POINTER(C.c_float)), ("shape", _HPGL_SHAPE)]


_COV_FIELDS = [
    ("covariance_type", C.c_int),
    ("ranges", C.c_double * 3),
    ("angles", C.c_double * 3),
    ("sill", C.c_double),
    ("nugget", C.c_double),
]
_NB_FIELDS = [("radiuses", C.c_int * 3), ("max_neighbours", C.c_int)]


class _HPGL_OK_PARAMS(C.Structure):
    _fields_ = _COV_FIELDS + _NB_FIELDS


class _HPGL_SK_PARAMS(C.Structure):
    _fields_ = _COV_FIELDS + _NB_FIELDS + [("automatic_mean", C.c_ubyte), ("mean", C.c_double)]


class _HPGL_SGS_PARAMS(C.Structure):
    _fields_ = _COV_FIELDS + _NB_FIELDS + [
        ("kriging_kind", C.c_int),
        ("seed", C.c_long),
        ("min_neighbours", C.c_int),
    ]


class _HPGL_MEDIAN_IK_PARAMS(C.Structure):
    _fields_ = _COV_FIELDS + _NB_FIELDS + [("marginal_probs", C.c_double * 2)]


class _HPGL_IK_PARAMS(C.Structure):
    _fields_ = _COV_FIELDS + _NB_FIELDS + [("marginal_prob", C.c_double)]


class __hpgl_cov_params_t(C.Structure):
    _fields_ = list(_COV_FIELDS)


class __hpgl_cockriging_m1_params_t(C.Structure):
    _fields_ = _COV_FIELDS + _NB_FIELDS + [
        ("primary_mean", C.c_double),
        ("secondary_mean", C.c_double),
        ("secondary_variance", C.c_double),
        ("correlation_coef", C.c_double),
    ]


class __hpgl_cockriging_m2_params_t(C.Structure):
    _fields_ = [
        ("primary_cov_params", globals()["__hpgl_cov_params_t"]),
        ("secondary_cov_params", globals()["__hpgl_cov_params_t"]),
    ] + _NB_FIELDS + [
        ("primary_mean", C.c_double),
        ("secondary_mean", C.c_double),
        ("correlation_coef", C.c_double),
    ]


class hpgl_non_parametric_cdf_t(C.Structure):
    _fields_ = [("values", C.POINTER(C.c_float)), ("probs", C.POINTER(C.c_float)), ("size", C.c_longlong)]


class _HPGL_AMD_STATS(C.Structure):
    _fields_ = [
        ("nodes_target", C.c_longlong),
        ("nodes_kriged", C.c_longlong),
        ("nodes_no_neighbours", C.c_longlong),
        ("nodes_singular", C.c_longlong),
        ("solves_fp32", C.c_longlong),
        ("solves_fp64", C.c_longlong),
        ("nodes_template_path", C.c_longlong),
        ("nodes_cluster_path", C.c_longlong),
        ("ms_setup_host", C.c_double),
        ("ms_search", C.c_double),
        ("ms_solve", C.c_double),
        ("ms_device_total", C.c_double),
        ("ms_wall", C.c_double),
        ("ms_k_template", C.c_double),
        ("ms_k_cluster", C.c_double),
        ("ms_k_solve_mixed", C.c_double),
        ("ms_k_solve_fp64", C.c_double),
        ("n_k_template", C.c_longlong),
        ("n_k_cluster", C.c_longlong),
        ("n_k_solve_mixed", C.c_longlong),
        ("n_k_solve_fp64", C.c_longlong),
        ("cluster_candidates", C.c_longlong),
        ("neighbours_total", C.c_longlong),
        ("cluster_lists_spilled", C.c_longlong),
        ("solves_fp32_two_steps", C.c_longlong),
    ]


class _HPGL_AMD_JOB(C.Structure):
    """struct hpgl_amd_job_t (include/hpgl.h): one request of the request loop."""
    _fields_ = [
        ("kind", C.c_int),
        ("shape", _HPGL_SHAPE),
        ("cont_data", C.POINTER(C.c_float)),
        ("cont_mask", C.POINTER(C.c_ubyte)),
        ("ind_data", C.POINTER(C.c_ubyte)),
        ("ind_mask", C.POINTER(C.c_ubyte)),
        ("indicator_count", C.c_int),
        ("mean", C.POINTER(C.c_float)),
        ("params", C.c_void_p),
        ("cdf", C.POINTER(hpgl_non_parametric_cdf_t)),
        ("sim_mask", C.POINTER(C.c_ubyte)),
        ("out_cont", C.POINTER(C.c_float)),
        ("out_cont_mask", C.POINTER(C.c_ubyte)),
        ("out_ind", C.POINTER(C.c_ubyte)),
        ("out_ind_mask", C.POINTER(C.c_ubyte)),
    ]


class _HPGL_AMD_JOB_RESULT(C.Structure):
    """struct hpgl_amd_job_result_t (include/hpgl.h)."""
    _fields_ = [
        ("status", C.c_int),
        ("queue_ms", C.c_double),
        ("latency_ms", C.c_double),
        ("run_ms", C.c_double),
        ("nodes", C.c_longlong),
        ("hbm_bytes", C.c_longlong),
        ("start_ms", C.c_double),
        ("end_ms", C.c_double),
        ("message", C.c_char * 256),
    ]


JOB_OK, JOB_SK, JOB_LVM, JOB_IK, JOB_MEDIAN_IK, JOB_SGS = range(6)


# Module-level alias without name mangling (the reference relies on the
# double-underscore name only inside this module).
_HPGL_COV_PARAMS = __hpgl_cov_params_t
_HPGL_COKRIGING_M1_PARAMS = __hpgl_cockriging_m1_params_t
_HPGL_COKRIGING_M2_PARAMS = __hpgl_cockriging_m2_params_t

