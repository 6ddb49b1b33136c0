"""Sequential Gaussian simulation, a Python 3 drop-in for the reference's ``geo_bsd/sgs.py``.

Restated for Python 3 without changing behaviour, except where the
reference could not run at all:

* ``cdf_data=None`` skips the normal-score transform, as the docstring and
  the C ABI say, instead of tripping the unconditional CdfData assert
  (sgs.py:100);
* ``mask != None`` on arrays (sgs.py:98) is ``mask is not None``;
* ``cdf_data`` given as data (a ``(values, mask)`` tuple or ContProperty)
  is turned into a CdfData with ``calc_cdf``, as the docstring describes
  ("data to use for calculating it").
"""
import ctypes as C

import numpy

from .cdf import CdfData, calc_cdf
from .geo import (
    CovarianceModel,
    ContProperty,
    accepts_tuple,
    _checked_create,
    _clone_prop,
    _create_hpgl_cont_masked_array,
    _create_hpgl_float_array,
    _create_hpgl_ubyte_array,
    _empty_clone,
    _require_cont_data,
    _requite_ind_data,
    _c_array,
)
from .hpgl_wrap import _HPGL_KRIGING_KIND, _HPGL_SGS_PARAMS, _hpgl_so, check_last_error, hpgl_non_parametric_cdf_t


def __prepare_sgs(prop, mean=None, use_harddata=True, mask=None):
    out_prop = _clone_prop(prop) if use_harddata else _empty_clone(prop)
    if mean is not None and not numpy.isscalar(mean):
        mean = _require_cont_data(mean)
    if mask is not None:
        mask = _requite_ind_data(mask)
    return out_prop, mean, mask


def _create_hpgl_nonparam_cdf(cdf_data):
    assert isinstance(cdf_data, CdfData)
    return _checked_create(
        hpgl_non_parametric_cdf_t,
        values=cdf_data.values.ctypes.data_as(C.POINTER(C.c_float)),
        probs=cdf_data.probs.ctypes.data_as(C.POINTER(C.c_float)),
        size=cdf_data.values.size)


def normed_cov_model(cov_model):
    """Scale sill and nugget by 1/(sill - nugget) (sgs.py:34-41)."""
    coef = cov_model.sill - cov_model.nugget
    return CovarianceModel(
        cov_model.type, cov_model.ranges, cov_model.angles, cov_model.sill / coef, cov_model.nugget / coef)


def _as_cdf(cdf_data):
    if cdf_data is None or isinstance(cdf_data, CdfData):
        return cdf_data
    if isinstance(cdf_data, tuple):
        cdf_data = ContProperty(*cdf_data)
    return calc_cdf(cdf_data)


def _sgs_params(cov_model, radiuses, max_neighbours, kriging_type, seed, min_neighbours):
    return _HPGL_SGS_PARAMS(
        covariance_type=cov_model.type,
        ranges=_c_array(C.c_double, 3, cov_model.ranges),
        angles=_c_array(C.c_double, 3, cov_model.angles),
        sill=cov_model.sill,
        nugget=cov_model.nugget,
        radiuses=_c_array(C.c_int, 3, radiuses),
        max_neighbours=max_neighbours,
        kriging_kind={"sk": _HPGL_KRIGING_KIND.simple, "ok": _HPGL_KRIGING_KIND.ordinary}[kriging_type],
        seed=seed,
        min_neighbours=min_neighbours)


@accepts_tuple("prop", 0)
def sgs_simulation(prop, grid, cdf_data, radiuses, max_neighbours, cov_model, seed, kriging_type="sk", mean=None,
                   use_harddata=True, use_regions=False, region_size=None, mask=None, force_single_thread=False,
                   force_parallel=False, min_neighbours=0, **params):
    """Performs Sequential Gaussian Simulation (sgs.py:43-120).

    cdf_data: None, CdfData, or data to compute one from.
    radiuses: search radii by X, Y, Z.  max_neighbours: neighbour limit.
    cov_model: CovarianceModel (normalised to unit sill - nugget).
    seed: random seed.  kriging_type: "sk" (mean 0) or "ok".
    mean: None, scalar (stationary) or array (locally varying mean).
    use_harddata: keep the informed cells of prop as conditioning data.
    mask: uint8 array, 1 where cells are simulated; None simulates all.
    """
    prop.fix_shape(grid)
    cov_model = normed_cov_model(cov_model)
    out_prop, mean, mask = __prepare_sgs(prop=prop, mean=mean, use_harddata=use_harddata, mask=mask)
    sgsp = _sgs_params(cov_model, radiuses, max_neighbours, kriging_type, seed, min_neighbours)
    cdf = _as_cdf(cdf_data)
    hpgl_cdf = C.byref(_create_hpgl_nonparam_cdf(cdf)) if cdf is not None else None
    hpgl_mask = C.byref(_create_hpgl_ubyte_array(mask, grid)) if mask is not None else None
    if mean is None or numpy.isscalar(mean):
        _hpgl_so.hpgl_sgs_simulation(
            C.byref(_create_hpgl_cont_masked_array(out_prop, grid)), C.byref(sgsp), hpgl_cdf,
            C.byref(C.c_double(mean)) if mean is not None else None, hpgl_mask)
    else:
        _hpgl_so.hpgl_sgs_lvm_simulation(
            C.byref(_create_hpgl_cont_masked_array(out_prop, grid)), C.byref(sgsp), hpgl_cdf,
            C.byref(_create_hpgl_float_array(mean, grid)), hpgl_mask)
    check_last_error()
    return out_prop


@accepts_tuple("prop", 0)
def sgs_simulation_batch(prop, grid, cdf_data, radiuses, max_neighbours, cov_model, seeds, kriging_type="sk",
                         mean=None, use_harddata=True, mask=None, min_neighbours=0, **params):
    """Several realisations of ``sgs_simulation`` in one call, one per seed.

    Takes the arguments of ``sgs_simulation`` with a sequence of seeds in
    place of one; realisation r equals ``sgs_simulation(..., seed=seeds[r])``.
    Returns a list of ContProperty, one per seed.  (No counterpart in the
    reference, which runs one realisation per call.)
    """
    prop.fix_shape(grid)
    cov_model = normed_cov_model(cov_model)
    base, mean, mask = __prepare_sgs(prop=prop, mean=mean, use_harddata=use_harddata, mask=mask)
    seeds = [int(s) for s in seeds]
    if not seeds:
        raise ValueError("at least one seed is needed")
    sgsp = _sgs_params(cov_model, radiuses, max_neighbours, kriging_type, seeds[0], min_neighbours)
    cdf = _as_cdf(cdf_data)
    hpgl_cdf = C.byref(_create_hpgl_nonparam_cdf(cdf)) if cdf is not None else None
    hpgl_mask = C.byref(_create_hpgl_ubyte_array(mask, grid)) if mask is not None else None
    lvm = None
    if mean is not None and not numpy.isscalar(mean):
        lvm = C.byref(_create_hpgl_float_array(mean, grid))
    R = len(seeds)
    shape = base.data.shape
    out = numpy.empty((R,) + tuple(shape[::-1]), dtype="float32")
    out_mask = numpy.empty((R,) + tuple(shape[::-1]), dtype="uint8")
    c_seeds = (C.c_long * R)(*seeds)
    _hpgl_so.hpgl_amd_sgs_simulation_batch(
        C.byref(_create_hpgl_cont_masked_array(base, grid)), C.byref(sgsp), hpgl_cdf, lvm, hpgl_mask, c_seeds, R,
        out.ctypes.data, out_mask.ctypes.data)
    check_last_error()
    # realisation r holds its N cells in Fortran order
    return [ContProperty(out[r].transpose(), out_mask[r].transpose()) for r in range(R)]
