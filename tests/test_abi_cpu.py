"""The C ABI library loads, exports every symbol include/hpgl.h declares,
and its struct layouts match the ctypes mirrors (no GPU calls)."""
import ctypes as C
import os
import re

import pytest

import oracle_binding as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_functions():
    src = open(os.path.join(ROOT, "include", "hpgl.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(hpgl_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    from hpgl_amd.geo_bsd import hpgl_wrap

    lib = hpgl_wrap._hpgl_so
    names = _declared_functions()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_variogram_library_exports_every_declared_symbol():
    src = open(os.path.join(ROOT, "include", "cvariogram.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    names = sorted(set(re.findall(r"\b([a-z_]+)\s*\(", src)) - {"if", "for"})
    lib = C.CDLL(os.path.join(ROOT, "hpgl_amd", "_cvariogram.so"))
    assert {"calc_variograms", "calc_variograms_from_point_set", "fill_ellipsoid_directions",
            "cvar_stack_layers"} <= set(names)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_variogram_struct_layouts():
    """_cvariogram/api.h:23-90 as the C compiler lays them out from include/cvariogram.h."""
    import subprocess
    import tempfile

    v = O.cvabi
    prog = r"""
#include <stdio.h>
#include <stddef.h>
#include "cvariogram.h"
int main(void) {
    printf("%zu %zu %zu %zu %zu %zu\n", sizeof(struct ellipsoid_t), sizeof(struct variogram_search_template_t),
           offsetof(struct variogram_search_template_t, m_ellipsoid), sizeof(struct cont_point_set_t),
           sizeof(struct hard_data_t), sizeof(struct float_data_t));
    return 0;
}
"""
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "l.c")
        open(c, "w").write(prog)
        exe = os.path.join(d, "l")
        subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), c, "-o", exe], check=True)
        got = [int(x) for x in subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split()]
    want = [C.sizeof(v.ellipsoid_t), C.sizeof(v.variogram_search_template_t),
            v.variogram_search_template_t.ellipsoid.offset, C.sizeof(v.cont_point_set_t), C.sizeof(v.hard_data_t),
            C.sizeof(v.float_data_t)]
    assert got == want == [96, 136, 40, 40, 64, 32]


def test_struct_sizes_match_reference_layouts():
    # LP64 layouts of geo_bsd/hpgl/api.h measured for the reference
    a = O.abi
    sizes = {"_HPGL_SHAPE": 24, "_HPGL_CONT_MASKED_ARRAY": 40, "_HPGL_IND_MASKED_ARRAY": 48, "_HPGL_FLOAT_ARRAY": 32,
             "_HPGL_UBYTE_ARRAY": 32, "_HPGL_OK_PARAMS": 88, "_HPGL_SK_PARAMS": 104, "_HPGL_COV_PARAMS": 72,
             "_HPGL_SGS_PARAMS": 112, "_HPGL_MEDIAN_IK_PARAMS": 104, "_HPGL_IK_PARAMS": 96,
             "hpgl_non_parametric_cdf_t": 24, "_HPGL_COKRIGING_M1_PARAMS": 120, "_HPGL_COKRIGING_M2_PARAMS": 184}
    for name, size in sizes.items():
        assert C.sizeof(getattr(a, name)) == size, name
    assert a._HPGL_SK_PARAMS.automatic_mean.offset == 88 and a._HPGL_SK_PARAMS.mean.offset == 96
    assert a._HPGL_SGS_PARAMS.kriging_kind.offset == 88 and a._HPGL_SGS_PARAMS.seed.offset == 96
    assert a._HPGL_OK_PARAMS.radiuses.offset == 72 and a._HPGL_OK_PARAMS.max_neighbours.offset == 84


def test_extension_struct_layouts_match_the_header():
    """hpgl_amd_job_t / hpgl_amd_job_result_t / hpgl_amd_stats_t as the C
    compiler lays them out from include/hpgl.h."""
    import subprocess
    import tempfile

    from hpgl_amd.geo_bsd import _abi as a

    prog = r"""
#include <stddef.h>
#include <stdio.h>
#include "hpgl.h"
int main(void) {
    printf("%zu %zu %zu %zu %zu %zu\n", sizeof(struct hpgl_amd_job_t), offsetof(struct hpgl_amd_job_t, params),
           offsetof(struct hpgl_amd_job_t, out_ind_mask), sizeof(struct hpgl_amd_job_result_t),
           offsetof(struct hpgl_amd_job_result_t, message), sizeof(struct hpgl_amd_stats_t));
    return 0;
}
"""
    with tempfile.TemporaryDirectory() as d:
        src, exe = os.path.join(d, "l.c"), os.path.join(d, "l")
        open(src, "w").write(prog)
        subprocess.check_call(["gcc", "-I", os.path.join(ROOT, "include"), src, "-o", exe])
        got = [int(v) for v in subprocess.check_output([exe]).split()]
    want = [C.sizeof(a._HPGL_AMD_JOB), a._HPGL_AMD_JOB.params.offset, a._HPGL_AMD_JOB.out_ind_mask.offset,
            C.sizeof(a._HPGL_AMD_JOB_RESULT), a._HPGL_AMD_JOB_RESULT.message.offset, C.sizeof(a._HPGL_AMD_STATS)]
    assert got == want


def test_errors_are_reported_not_thrown():
    """A failing call returns and sets the last error instead of throwing
    through the C ABI (the reference throws; api.cpp:54-67)."""
    from hpgl_amd.geo_bsd import hpgl_wrap

    lib = C.CDLL(hpgl_wrap.LIBRARY_PATH)  # own handle: argtypes of the shared one stay as declared
    lib.hpgl_read_inc_file_float.argtypes = [C.c_char_p, C.c_float, C.c_int, C.c_void_p, C.c_void_p]
    lib.hpgl_amd_get_last_error.restype = C.c_int
    lib.hpgl_get_last_exception_message.restype = C.c_char_p
    lib.hpgl_read_inc_file_float(b"no-such-dir/x.inc", -99.0, 0, None, None)
    assert lib.hpgl_amd_get_last_error() == 1
    assert b"Error opening file" in lib.hpgl_get_last_exception_message()


def test_python_api_surface():
    """The geo_bsd names a reference user calls on the kriging path exist."""
    import hpgl_amd.geo_bsd as g

    for name in ["ordinary_kriging", "simple_kriging", "lvm_kriging", "indicator_kriging", "median_ik",
                 "sgs_simulation", "simple_kriging_weights", "CovarianceModel", "SugarboxGrid", "ContProperty",
                 "IndProperty", "CdfData", "calc_cdf", "covariance", "set_thread_num", "get_thread_num",
                 "set_output_handler", "set_progress_handler", "normed_cov_model", "calc_mean"]:
        assert hasattr(g, name), name
    cm = g.CovarianceModel(type=g.covariance.exponential, ranges=(5, 1, 1), sill=2.0, nugget=0.5)
    n = g.normed_cov_model(cm)
    assert abs(n.sill - 2.0 / 1.5) < 1e-15 and abs(n.nugget - 0.5 / 1.5) < 1e-15
    try:
        g.CovarianceModel(sill=1.0, nugget=2.0)
        assert False
    except Exception as e:
        assert "Nugget" in str(e)


def test_indicator_property_range_check():
    """IndProperty raises exactly when the reference's expression
    sum((mask > 0) & (data >= indicator_count)) > 0 holds (geo.py:114),
    whatever the inputs' dtype and memory order; the clone of a checked
    property keeps its arrays and count."""
    import numpy as np

    from hpgl_amd.geo_bsd import geo

    rng = np.random.default_rng(3)
    shape = (7, 5, 4)
    for order in ("C", "F"):
        for ddt, mdt in (("uint8", "uint8"), ("int32", "uint8"), ("uint8", "float32"), ("int64", "int16")):
            data = np.asarray(rng.integers(0, 4, shape), dtype=ddt, order=order)
            mask = np.asarray(rng.integers(0, 2, shape), dtype=mdt, order=order)
            p = geo.IndProperty(data, mask, 4)
            assert p.data.flags["F_CONTIGUOUS"] and p.mask.flags["F_CONTIGUOUS"]
            c = geo._clone_prop(p)
            assert c.indicator_count == 4 and np.array_equal(c.data, p.data) and np.array_equal(c.mask, p.mask)
            assert c.data is not p.data
            bad = data.copy(order=order)
            i = np.flatnonzero(np.asarray(mask).ravel() > 0)[0]
            bad[np.unravel_index(i, shape)] = 9
            want = np.sum(np.bitwise_and(np.asarray(mask) > 0, np.asarray(bad) >= 4)) > 0
            assert want
            with pytest.raises(RuntimeError):
                geo.IndProperty(bad, mask, 4)
            # an out-of-range value where the mask is 0 is accepted
            ok = data.copy(order=order)
            j = np.flatnonzero(np.asarray(mask).ravel() == 0)[0]
            ok[np.unravel_index(j, shape)] = 9
            geo.IndProperty(ok, mask, 4)
