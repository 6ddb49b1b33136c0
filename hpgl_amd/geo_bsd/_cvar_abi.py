"""ctypes mirrors of the structs of include/cvariogram.h (the reference's
geo_bsd/cvariogram.py:9-52), in a module of their own so that code which
only needs the layouts does not load the GPU library."""
import ctypes as C


class vector_t(C.Structure):
    _fields_ = [("data", C.c_double * 3)]


class ellipsoid_t(C.Structure):
    _fields_ = [
        ("direction1", vector_t),
        ("direction2", vector_t),
        ("direction3", vector_t),
        ("R1", C.c_double),
        ("R2", C.c_double),
        ("R3", C.c_double)]


class variogram_search_template_t(C.Structure):
    _fields_ = [
        ("lag_width", C.c_double),
        ("lag_separation", C.c_double),
        ("tol_distance", C.c_double),
        ("num_lags", C.c_int),
        ("first_lag_distance", C.c_double),
        ("ellipsoid", ellipsoid_t)]


class hard_data_t(C.Structure):
    _fields_ = [
        ("data", C.POINTER(C.c_float)),
        ("mask", C.POINTER(C.c_ubyte)),
        ("data_shape", (C.c_int * 3)),
        ("data_strides", (C.c_int * 3)),
        ("mask_shape", (C.c_int * 3)),
        ("mask_strides", (C.c_int * 3))]


class cont_point_set_t(C.Structure):
    _fields_ = [
        ("xs", C.POINTER(C.c_float)),
        ("ys", C.POINTER(C.c_float)),
        ("zs", C.POINTER(C.c_float)),
        ("values", C.POINTER(C.c_float)),
        ("values_count", C.c_int)]


class float_data_t(C.Structure):
    _fields_ = [
        ("data", C.POINTER(C.c_float)),
        ("data_shape", (C.c_int * 3)),
        ("data_strides", (C.c_int * 3))]
