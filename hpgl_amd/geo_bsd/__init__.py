"""geo_bsd-compatible kriging API backed by the MI355X library hpgl_amd/hpgl.so.

``from hpgl_amd.geo_bsd import *`` provides what the reference's
``from geo_bsd import *`` provides on the kriging path (geo_bsd/__init__.py).
"""
from .geo import *  # noqa: F401,F403
from .geo import (  # noqa: F401
    _clone_prop, _empty_clone, _create_hpgl_cont_masked_array, _create_hpgl_ind_masked_array,
    _create_hpgl_float_array, _create_hpgl_ubyte_array, _create_hpgl_shape, _c_array)
from .sgs import sgs_simulation, sgs_simulation_batch, normed_cov_model  # noqa: F401
from .sis import sis_simulation, sis_simulation_batch  # noqa: F401
from .cdf import CdfData, calc_cdf  # noqa: F401
from .hpgl_wrap import last_stats, set_precision, get_precision, HpglError  # noqa: F401
