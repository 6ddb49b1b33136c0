"""Python front end of the library's request loop (``hpgl_amd_server_*``).

A ``Server`` owns worker slots, each with its own HIP stream; jobs wait in
arrival order and start as soon as a slot is free and their estimated HBM
footprint fits the unused budget (continuous batching), so jobs submitted
while others run overlap on the device.  Each job's result is the one the
corresponding ``geo_bsd`` call returns alone.

Jobs take the same arguments as the ``geo_bsd`` calls, already in the C
ABI's terms (parameter structs from ``hpgl_amd.geo_bsd._abi``); the arrays
of a job are kept alive by its ``Job`` object until ``wait`` returns.
"""
import ctypes as C

import numpy as np

from .geo_bsd import _abi as abi
from .geo_bsd.hpgl_wrap import HpglError, _hpgl_so

_F = C.POINTER(C.c_float)
_B = C.POINTER(C.c_ubyte)


def _fptr(a):
    return a.ctypes.data_as(_F) if a is not None else None


def _bptr(a):
    return a.ctypes.data_as(_B) if a is not None else None


def _shape(shape):
    x, y, z = (int(v) for v in shape)
    return abi._HPGL_SHAPE(data=(C.c_int * 3)(x, y, z), strides=(C.c_int * 3)(1, x, x * y))


class Job:
    """A submitted request; ``wait()`` returns its result dict and leaves the
    outputs in ``out`` / ``out_mask`` (or the SGS property in place)."""

    def __init__(self, server, jid, keep, out, out_mask):
        self._server, self.id, self._keep = server, jid, keep
        self.out, self.out_mask = out, out_mask
        self.result = None

    def wait(self):
        if self.result is None:
            r = abi._HPGL_AMD_JOB_RESULT()
            if _hpgl_so.hpgl_amd_server_wait(self._server._h, self.id, C.byref(r)) != 0:
                raise HpglError("unknown job %d" % self.id)
            self.result = {"status": r.status, "queue_ms": r.queue_ms, "latency_ms": r.latency_ms,
                           "run_ms": r.run_ms, "nodes": r.nodes, "hbm_bytes": r.hbm_bytes,
                           "start_ms": r.start_ms, "end_ms": r.end_ms,
                           "message": r.message.decode(errors="replace")}
            self._keep = None
        return self.result


class Server:
    def __init__(self, slots=4, hbm_budget_bytes=0):
        self._h = _hpgl_so.hpgl_amd_server_create(int(slots), int(hbm_budget_bytes))
        if not self._h:
            raise HpglError(_hpgl_so.hpgl_get_last_exception_message())

    def close(self):
        if self._h:
            _hpgl_so.hpgl_amd_server_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _submit(self, job, keep, out, out_mask):
        jid = _hpgl_so.hpgl_amd_server_submit(self._h, C.byref(job))
        if jid < 0:
            raise HpglError(_hpgl_so.hpgl_get_last_exception_message())
        return Job(self, jid, keep + [job], out, out_mask)

    def submit_cont(self, kind, shape, data, mask, params, mean=None):
        """OK (params: _HPGL_OK_PARAMS), SK (_HPGL_SK_PARAMS) or LVM
        (_HPGL_OK_PARAMS + per-cell mean); data float32 and mask uint8 in
        Fortran order.  The output is prefilled with the input, as the
        reference's front end does."""
        data = np.require(data, np.float32, "F")
        mask = np.require(mask, np.uint8, "F")
        out, out_mask = data.copy(order="F"), mask.copy(order="F")
        mean = np.require(mean, np.float32, "F") if mean is not None else None
        j = abi._HPGL_AMD_JOB(kind=kind, shape=_shape(shape), cont_data=_fptr(data), cont_mask=_bptr(mask),
                              mean=_fptr(mean), params=C.cast(C.pointer(params), C.c_void_p),
                              out_cont=_fptr(out), out_cont_mask=_bptr(out_mask))
        return self._submit(j, [data, mask, mean, params], out, out_mask)

    def submit_ik(self, shape, data, mask, params, count):
        """Indicator kriging (params: array of `count` _HPGL_IK_PARAMS) or,
        with kind median, a single _HPGL_MEDIAN_IK_PARAMS and count 2."""
        data = np.require(data, np.uint8, "F")
        mask = np.require(mask, np.uint8, "F")
        out, out_mask = data.copy(order="F"), mask.copy(order="F")
        median = isinstance(params, abi._HPGL_MEDIAN_IK_PARAMS)
        pp = C.cast(C.pointer(params) if median else params, C.c_void_p)
        j = abi._HPGL_AMD_JOB(kind=abi.JOB_MEDIAN_IK if median else abi.JOB_IK, shape=_shape(shape),
                              ind_data=_bptr(data), ind_mask=_bptr(mask), indicator_count=int(count), params=pp,
                              out_ind=_bptr(out), out_ind_mask=_bptr(out_mask))
        return self._submit(j, [data, mask, params], out, out_mask)

    def submit_sgs(self, shape, data, mask, params, cdf=None, mean=None, sim_mask=None):
        """SGS in place on copies of data / mask (returned as out / out_mask);
        cdf is None or a (values, probs) pair of float32 arrays."""
        out = np.require(data, np.float32, "F").copy(order="F")
        out_mask = np.require(mask, np.uint8, "F").copy(order="F")
        keep = [params]
        c = None
        if cdf is not None:
            v = np.ascontiguousarray(cdf[0], np.float32)
            p = np.ascontiguousarray(cdf[1], np.float32)
            c = abi.hpgl_non_parametric_cdf_t(values=v.ctypes.data_as(_F), probs=p.ctypes.data_as(_F), size=v.size)
            keep += [v, p, c]
        mean = np.require(mean, np.float32, "F") if mean is not None else None
        sim_mask = np.require(sim_mask, np.uint8, "F") if sim_mask is not None else None
        keep += [mean, sim_mask]
        j = abi._HPGL_AMD_JOB(kind=abi.JOB_SGS, shape=_shape(shape), cont_data=_fptr(out), cont_mask=_bptr(out_mask),
                              mean=_fptr(mean), params=C.cast(C.pointer(params), C.c_void_p),
                              cdf=C.pointer(c) if c is not None else None, sim_mask=_bptr(sim_mask))
        return self._submit(j, keep, out, out_mask)
