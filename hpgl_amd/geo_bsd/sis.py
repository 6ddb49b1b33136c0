"""Sequential indicator simulation, a Python 3 drop-in for the reference's ``geo_bsd/sis.py``.

Restated for Python 3 without changing behaviour, except that
``mask != None`` on arrays (sis.py:62, 73) is ``mask is not None``.
"""
import ctypes as C

import numpy

from .geo import (
    IndProperty,
    accepts_tuple,
    _c_array,
    _clone_prop,
    _create_hpgl_float_array,
    _create_hpgl_ik_params,
    _create_hpgl_ind_masked_array,
    _create_hpgl_ubyte_array,
    _empty_clone,
    _require_cont_data,
    _requite_ind_data,
)
from .hpgl_wrap import _HPGL_FLOAT_ARRAY, _hpgl_so, check_last_error


def __prepare_sis(prop, data, marginal_probs, mask, use_harddata):
    """sis.py:11-28: LVM when the marginal probabilities are arrays; like
    the reference, 'marginal_prob' is written into the dicts of ``data``."""
    is_lvm = not numpy.isscalar(marginal_probs[0])
    out_prop = _clone_prop(prop) if use_harddata else _empty_clone(prop)
    if is_lvm:
        marginal_probs = [_require_cont_data(m) for m in marginal_probs]
    for i in range(len(data)):
        data[i]["marginal_prob"] = 0 if is_lvm else marginal_probs[i]
    if mask is not None:
        mask = _requite_ind_data(mask)
    return out_prop, is_lvm, marginal_probs, mask


@accepts_tuple("prop", 0)
def sis_simulation(prop, grid, data, seed, marginal_probs, use_correlogram=True, mask=None,
                   force_single_thread=False, force_parallel=False, use_harddata=True, use_regions=False,
                   region_size=(), min_neighbours=0):
    """Sequential indicator simulation (sis.py:49-80).

    data: one dict per category with keys cov_model, radiuses,
    max_neighbours.  marginal_probs: one probability per category, or one
    per-cell array per category (locally varying mean; use_correlogram
    then selects the correlogram weights).  Two categories run median SIS.
    mask: uint8 array, 1 where cells are simulated; None simulates all.
    use_harddata: keep the informed cells of prop as conditioning data.
    """
    out_prop, is_lvm, marginal_probs, mask = __prepare_sis(prop, data, marginal_probs, mask, use_harddata)
    prop_2 = _create_hpgl_ind_masked_array(out_prop, grid)
    ikps = _create_hpgl_ik_params(data, len(data), is_lvm, marginal_probs)
    hpgl_mask = C.byref(_create_hpgl_ubyte_array(mask, grid)) if mask is not None else None
    if not is_lvm:
        _hpgl_so.hpgl_sis_simulation(C.byref(prop_2), ikps, len(data), seed, hpgl_mask)
    else:
        means = [_create_hpgl_float_array(m, grid) for m in marginal_probs]
        _hpgl_so.hpgl_sis_simulation_lvm(
            C.byref(prop_2), ikps, _c_array(_HPGL_FLOAT_ARRAY, len(data), means), len(data), seed, hpgl_mask,
            1 if use_correlogram else 0)
    check_last_error()
    return out_prop


@accepts_tuple("prop", 0)
def sis_simulation_batch(prop, grid, data, seeds, marginal_probs, use_correlogram=True, mask=None,
                         use_harddata=True):
    """Several realisations of ``sis_simulation`` in one call, one per seed.

    Realisation r equals ``sis_simulation(..., seed=seeds[r])``; returns a
    list of IndProperty.  (No counterpart in the reference, which runs one
    realisation per call.)
    """
    base, is_lvm, marginal_probs, mask = __prepare_sis(prop, data, marginal_probs, mask, use_harddata)
    seeds = [int(s) for s in seeds]
    if not seeds:
        raise ValueError("at least one seed is needed")
    ikps = _create_hpgl_ik_params(data, len(data), is_lvm, marginal_probs)
    hpgl_mask = C.byref(_create_hpgl_ubyte_array(mask, grid)) if mask is not None else None
    means = None
    if is_lvm:
        means = _c_array(_HPGL_FLOAT_ARRAY, len(data), [_create_hpgl_float_array(m, grid) for m in marginal_probs])
    R = len(seeds)
    shape = base.data.shape
    out = numpy.empty((R,) + tuple(shape[::-1]), dtype="uint8")
    out_mask = numpy.empty((R,) + tuple(shape[::-1]), dtype="uint8")
    c_seeds = (C.c_int * R)(*seeds)
    _hpgl_so.hpgl_amd_sis_simulation_batch(
        C.byref(_create_hpgl_ind_masked_array(base, grid)), ikps, len(data), means, 1 if use_correlogram else 0,
        hpgl_mask, c_seeds, R, out.ctypes.data, out_mask.ctypes.data)
    check_last_error()
    # realisation r holds its N cells in Fortran order
    return [IndProperty(out[r].transpose(), out_mask[r].transpose(), base.indicator_count) for r in range(R)]
