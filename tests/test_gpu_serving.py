"""The request loop: every job's result in a mix of overlapping jobs equals
its result alone.

Alone = the job submitted to a one-slot server and waited for before the
next is submitted (nothing else on the device).  Mix = all jobs submitted
at once to a four-slot server.  Tolerance: bit-identical (each job runs the
same kernels on the same inputs; only the interleaving differs), which is
inside the stated 1e-5 relative bound for estimates.
"""
import ctypes as C

import numpy as np
import pytest

from cases import categories, scattered, smooth_field

pytestmark = pytest.mark.gpu


def _cov_kw(type_, ranges, sill=1.0, nugget=0.0, angles=(0.0, 0.0, 0.0)):
    return dict(covariance_type=type_, ranges=(C.c_double * 3)(*ranges), angles=(C.c_double * 3)(*angles),
                sill=sill, nugget=nugget)


def _jobs():
    from hpgl_amd.geo_bsd import _abi as abi

    out = []
    for i, shape in enumerate([(40, 36, 12), (52, 30, 10)]):
        data, mask = scattered(shape, 0.02 + 0.01 * i, 300 + i, law="normal")
        r = (8 + 2 * i, 8, 3)
        okp = abi._HPGL_OK_PARAMS(radiuses=(C.c_int * 3)(*r), max_neighbours=12 + 4 * i,
                                  **_cov_kw(i % 3, (10, 8, 3), nugget=0.1))
        out.append(("ok", shape, (abi.JOB_OK, data, mask, okp)))
        skp = abi._HPGL_SK_PARAMS(radiuses=(C.c_int * 3)(*r), max_neighbours=16, automatic_mean=1, mean=0.0,
                                  **_cov_kw(1, (12, 10, 4), nugget=0.05))
        out.append(("sk", shape, (abi.JOB_SK, data, mask, skp)))
        mean = smooth_field(shape, 6.0, 400 + i) * 0.3
        out.append(("lvm", shape, (abi.JOB_LVM, data, mask, okp, mean)))
        cat, marg = categories(data, mask, quantiles=(0.33, 0.66), k=3)
        ikp = (abi._HPGL_IK_PARAMS * 3)()
        for k in range(3):
            ikp[k] = abi._HPGL_IK_PARAMS(radiuses=(C.c_int * 3)(*r), max_neighbours=12, marginal_prob=marg[k],
                                         **_cov_kw(0, (10 + 2 * k, 10 + 2 * k, 3)))
        out.append(("ik", shape, (cat, mask, ikp, 3)))
        cat2 = (data > np.median(data[mask == 1])).astype(np.uint8) * mask
        mp = abi._HPGL_MEDIAN_IK_PARAMS(radiuses=(C.c_int * 3)(*r), max_neighbours=12,
                                        marginal_probs=(C.c_double * 2)(0.5, 0.5), **_cov_kw(0, (10, 10, 3)))
        out.append(("median", shape, (cat2, mask, mp, 2)))
        sgp = abi._HPGL_SGS_PARAMS(radiuses=(C.c_int * 3)(*r), max_neighbours=12, kriging_kind=1, seed=77 + i,
                                   min_neighbours=0, **_cov_kw(0, (10, 10, 3)))
        v = np.unique(data[mask == 1])
        cdf = (v[:-1].astype(np.float32), (np.arange(1, v.size) / float(v.size)).astype(np.float32))
        out.append(("sgs", shape, (data, mask, sgp, cdf)))
    return out


def _submit(server, kind, shape, a):
    if kind in ("ok", "sk"):
        return server.submit_cont(a[0], shape, a[1], a[2], a[3])
    if kind == "lvm":
        return server.submit_cont(a[0], shape, a[1], a[2], a[3], mean=a[4])
    if kind in ("ik", "median"):
        return server.submit_ik(shape, a[0], a[1], a[2], a[3])
    return server.submit_sgs(shape, a[0], a[1], a[2], cdf=a[3])


def test_results_alone_equal_results_in_the_mix():
    from hpgl_amd.serving import Server

    jobs = _jobs()
    alone = []
    with Server(slots=1) as s:
        for kind, shape, a in jobs:
            j = _submit(s, kind, shape, a)
            r = j.wait()
            assert r["status"] == 0, (kind, r["message"])
            alone.append((j.out.copy(), j.out_mask.copy()))
    with Server(slots=4) as s:
        handles = [_submit(s, kind, shape, a) for kind, shape, a in jobs]
        results = [h.wait() for h in handles]
    for (kind, _, _), h, r, (o, om) in zip(jobs, handles, results, alone):
        assert r["status"] == 0, (kind, r["message"])
        assert np.array_equal(h.out_mask, om), kind
        assert np.array_equal(h.out, o), kind
    # jobs ran at the same time: sorted by start, some job starts before the
    # previous one has ended
    spans = sorted((r["start_ms"], r["end_ms"]) for r in results)
    assert any(b0 < a1 for (_, a1), (b0, _) in zip(spans, spans[1:]))


def test_admission_by_footprint_serialises_jobs_over_budget():
    """With a budget smaller than any job, jobs run one at a time (each
    admitted only when nothing else runs) and still complete correctly."""
    from hpgl_amd.serving import Server

    jobs = _jobs()[:3]
    with Server(slots=4, hbm_budget_bytes=1) as s:
        handles = [_submit(s, kind, shape, a) for kind, shape, a in jobs]
        results = [h.wait() for h in handles]
    assert all(r["status"] == 0 for r in results)
    spans = sorted((r["start_ms"], r["end_ms"]) for r in results)
    for (a0, a1), (b0, b1) in zip(spans, spans[1:]):
        assert b0 >= a1  # no two jobs ran at the same time
