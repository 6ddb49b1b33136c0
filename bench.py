"""Benchmark of the MI355X kriging path: kriged nodes/s on one GPU.

Workload (default): simple + indicator kriging on a 256x256x128 grid
(8,388,608 cells), 1% of cells informed (uniform, seed 1002), values
standard normal.
  * SK: mean 0, exponential covariance, ranges 40/40/8, sill 1, nugget 0.05,
    search radii 40/40/8 (radii are not given by the workload; chosen equal
    to the ranges), max_neighbours 24.
  * IK: 5 categories from thresholds at the data quantiles 0.1/0.3/0.5/0.7/0.9
    (category = min(searchsorted(t, v), 4)), marginal probabilities = the
    category frequencies; category k has a spherical covariance with ranges
    (30+5k, 30+5k, (30+5k)/5), sill 1, radii = ranges, max_neighbours 24.
One step = one SK job + one IK job on data already resident in HBM, through
the device-pointer C ABI (the output arrays are re-initialised from the
input inside the step, as the Python API clones the input).  Kriged nodes
per step = uninformed cells x (1 + 5).

Prints one JSON line.  --gpus N > 1 exits with a message: this build
measures one GPU.
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3
FP64_PEAK_TFLOPS = 78.6


def make_cfg2(seed=1002):
    from cases import Cov, Grid, categories, scattered

    g = Grid(256, 256, 128)
    data, mask = scattered(g.shape, 0.01, seed, law="normal")
    cat, marg = categories(data, mask)
    sk = dict(cov=Cov(1, (40, 40, 8), sill=1.0, nugget=0.05), radii=(40, 40, 8), max_n=24, mean=0.0)
    ik = [dict(cov=Cov(0, (30 + 5 * k, 30 + 5 * k, (30 + 5 * k) // 5)),
               radii=(30 + 5 * k, 30 + 5 * k, (30 + 5 * k) // 5), max_n=24, marginal=marg[k]) for k in range(5)]
    return g, data, mask, cat, marg, sk, ik


def _params(abi, sk, ik):
    def cov_kw(c):
        return dict(covariance_type=c.type, ranges=(C.c_double * 3)(*c.ranges), angles=(C.c_double * 3)(*c.angles),
                    sill=c.sill, nugget=c.nugget)

    skp = abi._HPGL_SK_PARAMS(radiuses=(C.c_int * 3)(*sk["radii"]), max_neighbours=sk["max_n"], automatic_mean=0,
                              mean=sk["mean"], **cov_kw(sk["cov"]))
    ikp = (abi._HPGL_IK_PARAMS * len(ik))()
    for i, p in enumerate(ik):
        ikp[i] = abi._HPGL_IK_PARAMS(radiuses=(C.c_int * 3)(*p["radii"]), max_neighbours=p["max_n"],
                                     marginal_prob=p["marginal"], **cov_kw(p["cov"]))
    return skp, ikp


def cpu_baseline(g, data, mask, cat, marg, sk, ik, budget_s=15.0, threads=16):
    """The CPU oracle (a C++/OpenMP restatement of the reference's CPU path)
    on a bounded sample of the same workload: every s-th cell of the grid is
    estimated, s chosen from a short pilot so the sample takes ~budget_s."""
    import oracle_binding as O

    O.set_threads(threads)
    n_unk = int(np.sum(mask == 0))
    ikd = [{"cov_model": p["cov"], "radiuses": p["radii"], "max_neighbours": p["max_n"]} for p in ik]

    def run(stride):
        O.set_sample_stride(stride)
        t0 = time.perf_counter()
        O.simple_kriging((data, mask), g, sk["radii"], sk["max_n"], sk["cov"], mean=sk["mean"])
        t1 = time.perf_counter()
        O.indicator_kriging((cat, mask, 5), g, ikd, marg)
        t2 = time.perf_counter()
        O.set_sample_stride(1)
        return t1 - t0, t2 - t1

    pilot = 4096
    a, b = run(pilot)
    per_sample = (a + b) / (g.size / pilot)
    stride = max(1, int(g.size * per_sample / budget_s))
    a, b = run(stride)
    sampled_unk = n_unk / stride  # uniform sampling of cells
    nodes = sampled_unk * (1 + 5)
    return {"value": nodes / (a + b), "unit": "kriged nodes/s", "cores": threads, "kind": "port",
            "sample": "every %d-th cell of the same SK+IK workload (%.0f kriged nodes, SK %.1f s + IK %.1f s)"
                      % (stride, nodes, a, b)}


def make_sis_workload(seed=1006):
    """The reference's published SIS benchmark shape (doc/hpgl.sourceforge.net/index.html:103-128):
    166x141x225 grid (5,266,350 cells), 20% informed, two indicators (median
    SIS), exponential covariance ranges 10/10/10, search radii 20/20/20, at
    most 12 neighbours.  Categories: 1 above the 0.8 quantile of a standard
    normal sample, else 0 (marginals 0.8/0.2 as in the reference's SIS test
    scripts)."""
    from cases import Cov, Grid, categories, scattered

    g = Grid(166, 141, 225)
    data, mask = scattered(g.shape, 0.20, seed, law="normal")
    cat, marg = categories(data, mask, quantiles=(0.8,), k=2)
    params = [dict(cov=Cov(1, (10, 10, 10)), radii=(20, 20, 20), max_n=12, marginal=marg[k]) for k in range(2)]
    return g, cat, mask, marg, params


# 4,213,080 simulated cells in 1 min 30 s on one Core Duo 1.8 GHz core
# (doc/hpgl.sourceforge.net/index.html:128)
SIS_PUBLISHED_CELLS_PER_S = 0.8 * 5266350 / 90.0


def sis_bench(torch, lib, abi, hpgl_wrap, steps=2, warmup=1, cpu=True, seed=3241347):
    """Sequential indicator simulation of one realisation on the published
    SIS workload, inputs resident in HBM; simulated cells/s.  The CPU leg
    runs the oracle on the same workload and seed (one thread) and checks
    the GPU realisation against it."""
    g, cat, mask, marg, params = make_sis_workload()
    _, ikp = _params(abi, dict(cov=params[0]["cov"], radii=(1, 1, 1), max_n=1, mean=0.0), params)
    dev = torch.device("cuda:0")
    d_cat = torch.from_numpy(cat.ravel(order="F").copy()).to(dev)
    d_mask = torch.from_numpy(mask.ravel(order="F").copy()).to(dev)
    d_out = torch.empty_like(d_cat)
    d_out_mask = torch.empty_like(d_mask)
    shape = abi._HPGL_SHAPE(data=(C.c_int * 3)(*g.shape), strides=(C.c_int * 3)(1, g.x, g.x * g.y))
    stream = torch.cuda.current_stream(dev)
    sp = C.c_void_p(stream.cuda_stream)
    seeds = (C.c_int * 1)(seed)
    n_sim = int(np.sum(mask == 0))

    def run():
        lib.hpgl_amd_sis_simulation_batch_dev(d_cat.data_ptr(), d_mask.data_ptr(), C.byref(shape), ikp, 2, None, 0,
                                              None, seeds, 1, d_out.data_ptr(), d_out_mask.data_ptr(), sp)
        hpgl_wrap.check_last_error()
        return hpgl_wrap.last_stats()

    for _ in range(warmup):
        run()
    torch.cuda.synchronize()
    acc = {}
    t0 = time.perf_counter()
    for _ in range(steps):
        st = run()
        for k, v in st.items():
            acc[k] = acc.get(k, 0) + v
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) / steps
    res = {"workload": "median SIS, 166x141x225 grid, 20% informed, exponential ranges 10/10/10, radii 20/20/20, "
                       "max_neighbours 12, one realisation",
           "simulated_cells": n_sim, "value": n_sim / wall, "unit": "simulated cells/s", "ms_per_realisation": wall * 1e3,
           "ms_template_scan": acc.get("ms_k_template", 0.0) / steps, "ms_solve": acc.get("ms_solve", 0.0) / steps,
           "ms_device_total": acc.get("ms_device_total", 0.0) / steps,
           "published_reference": {"value": SIS_PUBLISHED_CELLS_PER_S, "unit": "simulated cells/s",
                                   "hardware": "1 core, Intel Core Duo 1.8 GHz",
                                   "source": "doc/hpgl.sourceforge.net/index.html:128"}}
    res["vs_published"] = res["value"] / SIS_PUBLISHED_CELLS_PER_S
    if cpu:
        import oracle_binding as O

        gpu_cat = d_out.cpu().numpy().reshape(g.shape, order="F")
        hpgl_wrap.set_precision("fp64")
        run()
        torch.cuda.synchronize()
        hpgl_wrap.set_precision("mixed")
        gpu64 = d_out.cpu().numpy().reshape(g.shape, order="F")
        data = [{"cov_model": p["cov"], "radiuses": p["radii"], "max_neighbours": p["max_n"]} for p in params]
        O.set_threads(1)
        t0 = time.perf_counter()
        od, om = O.sis_simulation((cat, mask, 2), g, data, seed, marg)
        dt = time.perf_counter() - t0
        res["cpu_baseline"] = {"value": n_sim / dt, "unit": "simulated cells/s", "cores": 1, "kind": "port",
                               "sample": "the whole workload (one realisation, same seed)"}
        res["fp64_equals_oracle"] = bool(np.array_equal(gpu64, od))
        res["mixed_agreement_vs_oracle"] = float(np.mean(gpu_cat == od))
    return res


def make_cokriging_workload(seed=1007):
    """Colocated cokriging (Markov model I) on the grid of the reference's
    published benchmarks (166x141x225, 20% informed, exponential ranges
    10/10/10, radii 20/20/20, max_neighbours 12); the secondary variable is
    a smooth random field informed everywhere; primary mean 0, secondary
    mean 0, secondary variance 0.07 and correlation 0.25 as in the
    reference's regression script (geo_testing/test_compare.py:165-166)."""
    from cases import Cov, Grid, scattered, smooth_field

    g = Grid(166, 141, 225)
    data, mask = scattered(g.shape, 0.20, seed, law="normal")
    sec = smooth_field(g.shape, 10.0, seed + 1)
    smask = np.ones(g.shape, np.uint8, order="F")
    params = dict(cov=Cov(1, (10, 10, 10)), radii=(20, 20, 20), max_n=12, primary_mean=0.0, secondary_mean=0.0,
                  secondary_variance=0.07, correlation=0.25)
    return g, data, mask, sec, smask, params


def cokriging_bench(torch, lib, abi, hpgl_wrap, steps=2, warmup=1, cpu=True):
    """Estimated cells/s of simple colocated cokriging, inputs resident in
    HBM.  The CPU leg runs the oracle on every 64th cell (16 threads) and the
    GPU result is compared with it on those cells."""
    g, data, mask, sec, smask, prm = make_cokriging_workload()
    dev = torch.device("cuda:0")
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a.ravel(order="F"))).to(dev)
    d_in, d_m, d_s, d_sm = t(data), t(mask), t(sec), t(smask)
    d_out, d_om = torch.empty_like(d_in), torch.empty_like(d_m)
    c = prm["cov"]
    p = abi._HPGL_COKRIGING_M1_PARAMS(covariance_type=c.type, ranges=(C.c_double * 3)(*c.ranges),
                                      angles=(C.c_double * 3)(*c.angles), sill=c.sill, nugget=c.nugget,
                                      radiuses=(C.c_int * 3)(*prm["radii"]), max_neighbours=prm["max_n"],
                                      primary_mean=prm["primary_mean"], secondary_mean=prm["secondary_mean"],
                                      secondary_variance=prm["secondary_variance"],
                                      correlation_coef=prm["correlation"])
    shape = abi._HPGL_SHAPE(data=(C.c_int * 3)(*g.shape), strides=(C.c_int * 3)(1, g.x, g.x * g.y))
    stream = torch.cuda.current_stream(dev)
    sp = C.c_void_p(stream.cuda_stream)
    n_est = int(np.sum(mask == 0))

    def run():
        d_out.copy_(d_in)
        d_om.copy_(d_m)
        lib.hpgl_amd_simple_cokriging_mark1_dev(d_in.data_ptr(), d_m.data_ptr(), d_s.data_ptr(), d_sm.data_ptr(),
                                                C.byref(shape), C.byref(p), d_out.data_ptr(), d_om.data_ptr(), sp)
        hpgl_wrap.check_last_error()
        return hpgl_wrap.last_stats()

    for _ in range(warmup):
        run()
    torch.cuda.synchronize()
    acc = {}
    t0 = time.perf_counter()
    for _ in range(steps):
        st = run()
        for k, v in st.items():
            acc[k] = acc.get(k, 0) + v
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) / steps
    res = {"workload": "simple colocated cokriging (Markov model I), 166x141x225 grid, 20% informed, exponential "
                       "ranges 10/10/10, radii 20/20/20, max_neighbours 12, secondary informed everywhere",
           "estimated_cells": n_est, "value": n_est / wall, "unit": "estimated cells/s", "ms_per_job": wall * 1e3,
           "ms_template_scan": acc.get("ms_k_template", 0.0) / steps,
           "ms_solve": acc.get("ms_k_solve_fp64", 0.0) / steps,
           "singular": int(acc.get("nodes_singular", 0) / steps), "dtype": "f64"}
    if cpu:
        import oracle_binding as O

        stride = 64
        out = d_out.cpu().numpy().reshape(g.shape, order="F")
        O.set_threads(16)
        O.set_sample_stride(stride)
        t0 = time.perf_counter()
        od, om = O.simple_cokriging_markI((data, mask), g, (sec, smask), prm["primary_mean"], prm["secondary_mean"],
                                          prm["secondary_variance"], prm["correlation"], prm["radii"], prm["max_n"],
                                          c)
        dt = time.perf_counter() - t0
        O.set_sample_stride(1)
        cells = np.arange(g.size) % stride == 0
        flat_m = mask.ravel(order="F") == 0
        sampled = int(np.sum(cells & flat_m))
        res["cpu_baseline"] = {"value": sampled / dt, "unit": "estimated cells/s", "cores": 16, "kind": "port",
                               "sample": "every 64th cell of the same job (%d estimated cells)" % sampled}
        sel = cells
        res["equals_oracle_on_sample"] = bool(np.array_equal(out.ravel(order="F")[sel], od.ravel(order="F")[sel]))
    return res


def variogram_bench(cpu=True):
    """Experimental semivariogram of gridded data on the 256x256x128, 1%-informed
    grid: a horizontal tunnel (azimuth 0, dip 0, half-bandwidths 10 and 2
    cells, tolerance 2) of 20 lags of 4 cells.  Throughput = candidate pairs
    (informed cell x tunnel offset) per second of the whole call (host arrays
    in, PCIe copies included).  The CPU leg runs the oracle on the first 8 z
    layers of the same data."""
    from hpgl_amd.geo_bsd import cvariogram as cv

    g, data, mask, _, _, _, _ = make_cfg2()
    args = (4.0, 4.0, 2.0, 20, 0.0)
    ell = (1, 10.0, 2.0, 0, 0, 0)
    t = cv.VariogramSearchTemplate(*args, cv.Ellipsoid(*ell))
    cv.CalcVariograms(t, (data, mask))  # warm-up
    t0 = time.perf_counter()
    _, vario = cv.CalcVariograms(t, (data, mask))
    wall = time.perf_counter() - t0
    st = cv.last_stats()
    res = {"workload": "gridded semivariogram, 256x256x128 grid, 1% informed, horizontal tunnel (half-bandwidths "
                       "10/2 cells, tolerance 2), 20 lags of 4 cells",
           "offsets": st["offsets"], "candidate_pairs": st["pairs_tested"],
           "value": st["pairs_tested"] / wall, "unit": "candidate pairs/s", "ms_per_call": wall * 1e3,
           "device_ms": st["device_ms"], "variogram": [round(float(v), 6) for v in vario]}
    if cpu:
        import oracle_binding as O

        sub = (np.asfortranarray(data[:, :, :8]), np.asfortranarray(mask[:, :, :8]))
        ot = O.variogram_template(*args[:4], args[4], O.ellipsoid(*ell))
        t0 = time.perf_counter()
        want = O.calc_variograms(ot, *sub)
        dt = time.perf_counter() - t0
        pairs = int(np.sum(sub[1] != 0)) * st["offsets"]
        _, got = cv.CalcVariograms(t, sub)
        fin = np.isfinite(want)
        res["cpu_baseline"] = {"value": pairs / dt, "unit": "candidate pairs/s", "cores": 1, "kind": "port",
                               "sample": "the first 8 z layers (256x256x8) of the same data"}
        res["max_rel_diff_vs_oracle_on_sample"] = float(
            np.max(np.abs(got[fin] - want[fin]) / np.maximum(np.abs(want[fin]), 1e-6), initial=0.0))
    return res


def sgs_bench(torch, lib, abi, hpgl_wrap, R=64, cpu=True):
    """Sequential Gaussian simulation, R realisations batched in one call
    (200x200x100 grid, 1% conditioning data as normal scores, spherical
    ranges 30/30/6, sill 1, nugget 0, radii 30/30/6, max_neighbours 32, SK,
    no transform, seeds 1000..1000+R-1), inputs resident in HBM; simulated
    cells/s over all realisations.  The CPU leg runs the oracle on one
    realisation of a 100x100x50 grid with the same density and model (the
    reference's simulation is serial)."""
    from cases import Cov, Grid, scattered

    g = Grid(200, 200, 100)
    data, mask = scattered(g.shape, 0.01, 1008, law="normal")
    cov = Cov(0, (30, 30, 6), sill=1.0, nugget=0.0)
    p = abi._HPGL_SGS_PARAMS(covariance_type=cov.type, ranges=(C.c_double * 3)(*cov.ranges),
                             angles=(C.c_double * 3)(*cov.angles), sill=cov.sill, nugget=cov.nugget,
                             radiuses=(C.c_int * 3)(30, 30, 6), max_neighbours=32, kriging_kind=1, seed=1000,
                             min_neighbours=0)
    dev = torch.device("cuda:0")
    d_in = torch.from_numpy(data.ravel(order="F").copy()).to(dev)
    d_m = torch.from_numpy(mask.ravel(order="F").copy()).to(dev)
    d_out = torch.empty(R * g.size, dtype=torch.float32, device=dev)
    d_om = torch.empty(R * g.size, dtype=torch.uint8, device=dev)
    shape = abi._HPGL_SHAPE(data=(C.c_int * 3)(*g.shape), strides=(C.c_int * 3)(1, g.x, g.x * g.y))
    seeds = (C.c_long * R)(*[1000 + r for r in range(R)])
    stream = torch.cuda.current_stream(dev)
    sp = C.c_void_p(stream.cuda_stream)

    def run(n):
        lib.hpgl_amd_sgs_simulation_batch_dev(d_in.data_ptr(), d_m.data_ptr(), C.byref(shape), C.byref(p), None,
                                              None, None, seeds, n, d_out.data_ptr(), d_om.data_ptr(), sp)
        hpgl_wrap.check_last_error()
        return hpgl_wrap.last_stats()

    run(1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    st = run(R)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    n_sim = int(np.sum(mask == 0)) * R
    res = {"workload": "SGS, %d realisations in one call, 200x200x100 grid, 1%% conditioning (normal scores), "
                       "spherical ranges 30/30/6, radii 30/30/6, max_neighbours 32, SK" % R,
           "simulated_cells": n_sim, "value": n_sim / wall, "unit": "simulated cells/s",
           "ms_per_call": wall * 1e3, "ms_per_realisation": wall * 1e3 / R,
           "ms_template_scan": st["ms_k_template"], "ms_solve": st["ms_solve"],
           "solves": {"fp32_accepted": int(st["solves_fp32"]), "fp64": int(st["solves_fp64"])}}
    if cpu:
        import oracle_binding as O

        gs = Grid(100, 100, 50)
        sd, sm = scattered(gs.shape, 0.01, 1009, law="normal")
        O.set_threads(1)
        t0 = time.perf_counter()
        O.sgs_simulation((sd, sm), gs, None, (30, 30, 6), 32, cov, 1000, kriging_type="sk")
        dt = time.perf_counter() - t0
        res["cpu_baseline"] = {"value": int(np.sum(sm == 0)) / dt, "unit": "simulated cells/s", "cores": 1,
                               "kind": "port", "sample": "one realisation of a 100x100x50 grid, same density "
                                                         "and model (%.1f s)" % dt}
    return res


def make_serving_trace(n_jobs=60, rate=20.0, seed=1005):
    """Poisson request stream of mixed OK / SK / IK / LVM jobs: grid axes
    uniform in [32, 256] x [32, 256] x [32, 128], informed fraction
    U(0.5%, 3%), covariance type U{0, 1, 2}, horizontal ranges U(5, 50) and
    vertical ranges a quarter of that, radii = ranges, max_neighbours
    U{8..32}; IK with 3 categories (data terciles); the LVM mean is a smooth
    Gaussian field of correlation length 50 / 3 cells (range 50)."""
    from cases import categories, scattered, smooth_field

    rng = np.random.default_rng(seed)
    t = 0.0
    jobs = []
    for i in range(n_jobs):
        t += rng.exponential(1.0 / rate)
        kind = ["ok", "sk", "ik", "lvm"][int(rng.integers(4))]
        shape = (int(rng.integers(32, 257)), int(rng.integers(32, 257)), int(rng.integers(32, 129)))
        frac = float(rng.uniform(0.005, 0.03))
        ctype = int(rng.integers(3))
        h = float(rng.uniform(5, 50))
        ranges = (h, h, max(h / 4, 1.0))
        radii = tuple(max(1, int(round(v))) for v in ranges)
        max_n = int(rng.integers(8, 33))
        data, mask = scattered(shape, frac, int(rng.integers(1 << 30)), law="normal")
        job = dict(t=t, kind=kind, shape=shape, ctype=ctype, ranges=ranges, radii=radii, max_n=max_n,
                   data=data, mask=mask)
        if kind == "ik":
            job["cat"], job["marg"] = categories(data, mask, quantiles=(1 / 3, 2 / 3), k=3)
        if kind == "lvm":
            job["mean"] = smooth_field(shape, 50.0 / 3.0, int(rng.integers(1 << 30)))
        jobs.append(job)
    return jobs


def _serving_params(abi, job):
    def cov_kw():
        return dict(covariance_type=job["ctype"], ranges=(C.c_double * 3)(*job["ranges"]),
                    angles=(C.c_double * 3)(0.0, 0.0, 0.0), sill=1.0, nugget=0.1)

    r = (C.c_int * 3)(*job["radii"])
    if job["kind"] in ("ok", "lvm"):
        return abi._HPGL_OK_PARAMS(radiuses=r, max_neighbours=job["max_n"], **cov_kw())
    if job["kind"] == "sk":
        return abi._HPGL_SK_PARAMS(radiuses=r, max_neighbours=job["max_n"], automatic_mean=1, mean=0.0, **cov_kw())
    ikp = (abi._HPGL_IK_PARAMS * 3)()
    for k in range(3):
        ikp[k] = abi._HPGL_IK_PARAMS(radiuses=(C.c_int * 3)(*job["radii"]), max_neighbours=job["max_n"],
                                     marginal_prob=job["marg"][k], **cov_kw())
    return ikp


def property_io_bench(g, data, mask):
    """Writes the SK result (every cell) as an INC file and reads it back
    through the library's host code; values/s and MB/s of each."""
    import tempfile

    shape = data.shape
    d = np.asfortranarray(data)
    m = np.asfortranarray(mask)
    with tempfile.TemporaryDirectory() as tmp:
        path = os.path.join(tmp, "sk.inc")
        t0 = time.perf_counter()
        g.write_property((d, m), path, "SK_ESTIMATE", -99)
        tw = time.perf_counter() - t0
        size = os.path.getsize(path)
        t0 = time.perf_counter()
        back = g.load_cont_property(path, -99, size=shape)
        tr = time.perf_counter() - t0
        same_mask = bool(np.array_equal(back.mask, m))
    n = d.size
    return {"format": "INC, one %E value per line", "values": int(n), "file_mb": round(size / 1e6, 1),
            "write_values_per_s": round(n / tw), "write_mb_per_s": round(size / tw / 1e6, 1),
            "read_values_per_s": round(n / tr), "read_mb_per_s": round(size / tr / 1e6, 1),
            "round_trip_mask_equal": same_mask, "threads": 1}


def serving_bench(slots=4):
    """Replays the request stream against the library's request loop:
    jobs are submitted at their arrival times and overlap on the GPU."""
    from hpgl_amd.geo_bsd import _abi as abi
    from hpgl_amd.serving import Server

    jobs = make_serving_trace()
    kinds = {"ok": abi.JOB_OK, "sk": abi.JOB_SK, "lvm": abi.JOB_LVM}
    with Server(slots=slots) as srv:
        # warm the resident covariance assets and kernels with one small job
        w = make_serving_trace(n_jobs=1, seed=7)[0]
        srv.submit_cont(abi.JOB_SK, w["shape"], w["data"], w["mask"],
                        _serving_params(abi, dict(w, kind="sk"))).wait()
        handles = []
        t0 = time.perf_counter()
        for job in jobs:
            delay = job["t"] - (time.perf_counter() - t0)
            if delay > 0:
                time.sleep(delay)
            p = _serving_params(abi, job)
            if job["kind"] == "ik":
                handles.append(srv.submit_ik(job["shape"], job["cat"], job["mask"], p, 3))
            else:
                handles.append(srv.submit_cont(kinds[job["kind"]], job["shape"], job["data"], job["mask"], p,
                                               mean=job.get("mean")))
        results = [h.wait() for h in handles]
        makespan = time.perf_counter() - t0
    failed = [r["message"] for r in results if r["status"] != 0]
    if failed:
        raise RuntimeError("serving job failed: %s" % failed[0])
    lat = np.array([r["latency_ms"] for r in results])
    nodes = sum(r["nodes"] for r in results)
    return jobs, {"jobs": len(jobs), "arrival_rate_per_s": 20.0, "slots": slots,
                  "p50_latency_ms": round(float(np.percentile(lat, 50)), 2),
                  "p99_latency_ms": round(float(np.percentile(lat, 99)), 2),
                  "mean_queue_ms": round(float(np.mean([r["queue_ms"] for r in results])), 2),
                  "kriged_nodes": int(nodes), "nodes_per_s": nodes / makespan, "makespan_s": round(makespan, 3),
                  "trace": "Poisson 20 jobs/s, OK/SK/IK/LVM, grids 32..256 x 32..256 x 32..128, seed 1005"}


def serving_cpu_baseline(jobs, threads=16, nodes_per_sample=20000):
    """The same trace processed first-in first-out by the CPU oracle: each
    job's service time is measured on a sample of its cells (every s-th
    cell, about nodes_per_sample of them) and scaled to the job; latencies
    follow from the arrival times of a single FIFO server."""
    import oracle_binding as O
    from cases import Cov, Grid

    O.set_threads(threads)
    service = []
    nodes_all = 0
    for job in jobs:
        g = Grid(*job["shape"])
        n_unk = int(np.sum(job["mask"] == 0))
        nodes = g.size * 3 if job["kind"] == "ik" else n_unk
        stride = max(1, g.size // nodes_per_sample)
        cov = Cov(job["ctype"], job["ranges"], sill=1.0, nugget=0.1)
        O.set_sample_stride(stride)
        t0 = time.perf_counter()
        if job["kind"] == "ok":
            O.ordinary_kriging((job["data"], job["mask"]), g, job["radii"], job["max_n"], cov)
        elif job["kind"] == "sk":
            O.simple_kriging((job["data"], job["mask"]), g, job["radii"], job["max_n"], cov)
        elif job["kind"] == "lvm":
            O.lvm_kriging((job["data"], job["mask"]), g, job["mean"], job["radii"], job["max_n"], cov)
        else:
            ikd = [{"cov_model": cov, "radiuses": job["radii"], "max_neighbours": job["max_n"]} for _ in range(3)]
            O.indicator_kriging((job["cat"], job["mask"], 3), g, ikd, job["marg"])
        dt = time.perf_counter() - t0
        O.set_sample_stride(1)
        service.append(dt * stride)
        nodes_all += nodes
    end = 0.0
    lat = []
    for job, sv in zip(jobs, service):
        start = max(job["t"], end)
        end = start + sv
        lat.append((end - job["t"]) * 1000.0)
    return {"p50_latency_ms": round(float(np.percentile(lat, 50)), 1),
            "p99_latency_ms": round(float(np.percentile(lat, 99)), 1),
            "nodes_per_s": nodes_all / end, "cores": threads, "kind": "port",
            "sample": "each job's cells sampled every s-th (about %d per job), service time scaled by s, "
                      "FIFO single server over the arrival trace" % nodes_per_sample}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--precision", choices=["mixed", "fp64"], default="mixed")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fp64-compare", action="store_true")
    ap.add_argument("--no-serving", action="store_true", help="skip the request-stream (p50/p99) measurement")
    ap.add_argument("--no-io", action="store_true", help="skip the property-file write/read measurement")
    ap.add_argument("--no-sis", action="store_true", help="skip the sequential indicator simulation measurement")
    ap.add_argument("--no-sgs", action="store_true", help="skip the batched SGS measurement")
    ap.add_argument("--sgs-only", action="store_true", help="run only the batched SGS measurement (for profiling)")
    ap.add_argument("--no-cokriging", action="store_true", help="skip the cokriging measurement")
    ap.add_argument("--no-variogram", action="store_true", help="skip the experimental variogram measurement")
    ap.add_argument("--variogram-only", action="store_true",
                    help="run only the experimental variogram measurement (for profiling)")
    ap.add_argument("--cokriging-only", action="store_true",
                    help="run only the cokriging measurement (for profiling)")
    ap.add_argument("--sis-only", action="store_true",
                    help="run only the sequential indicator simulation measurement (for profiling)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run only a plain 1 GiB device copy (x * 1 -> y, 3 times), the reference kernel for "
                         "scaling rocprofv3 FETCH_SIZE / WRITE_SIZE to bytes")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the kriging path has no multi-GPU mode in this build "
              "(--gpus %d requested)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import torch

    if args.copy_calibration:
        n = 256 * 1024 * 1024
        x = torch.ones(n, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        for _ in range(3):
            torch.mul(x, 1.0, out=y)
        torch.cuda.synchronize()
        print(json.dumps({"copy_kernel": "torch.mul(x, 1.0, out=y)", "bytes_read_per_launch": 4 * n,
                          "bytes_written_per_launch": 4 * n, "launches": 3}), flush=True)
        return

    from hpgl_amd.geo_bsd import hpgl_wrap
    from hpgl_amd.geo_bsd import _abi as abi

    lib = hpgl_wrap._hpgl_so
    vp = C.c_void_p
    lib.hpgl_amd_simple_kriging_dev.argtypes = [vp, vp, C.POINTER(abi._HPGL_SHAPE), C.POINTER(abi._HPGL_SK_PARAMS),
                                                vp, vp, vp]
    lib.hpgl_amd_indicator_kriging_dev.argtypes = [vp, vp, C.POINTER(abi._HPGL_SHAPE),
                                                   C.POINTER(abi._HPGL_IK_PARAMS), C.c_int, vp, vp, vp]

    if args.sgs_only:
        print(json.dumps({"sgs": sgs_bench(torch, lib, abi, hpgl_wrap, cpu=not args.no_cpu_baseline)}), flush=True)
        return
    if args.variogram_only:
        print(json.dumps({"variogram": variogram_bench(cpu=not args.no_cpu_baseline)}), flush=True)
        return
    if args.cokriging_only:
        print(json.dumps({"cokriging": cokriging_bench(torch, lib, abi, hpgl_wrap, steps=args.steps,
                                                       warmup=args.warmup, cpu=not args.no_cpu_baseline)}),
              flush=True)
        return
    if args.sis_only:
        print(json.dumps({"sis": sis_bench(torch, lib, abi, hpgl_wrap, steps=args.steps, warmup=args.warmup,
                                           cpu=not args.no_cpu_baseline)}), flush=True)
        return

    g, data, mask, cat, marg, sk, ik = make_cfg2()
    n_unk = int(np.sum(mask == 0))
    nodes_per_step = n_unk * (1 + len(ik))

    dev = torch.device("cuda:0")
    d_in = torch.from_numpy(data.ravel(order="F").copy()).to(dev)
    d_mask = torch.from_numpy(mask.ravel(order="F").copy()).to(dev)
    d_cat = torch.from_numpy(cat.ravel(order="F").copy()).to(dev)
    d_out = torch.empty_like(d_in)
    d_out_mask = torch.empty_like(d_mask)
    d_iout = torch.empty_like(d_cat)
    d_iout_mask = torch.empty_like(d_mask)
    shape = abi._HPGL_SHAPE(data=(C.c_int * 3)(*g.shape), strides=(C.c_int * 3)(1, g.x, g.x * g.y))
    skp, ikp = _params(abi, sk, ik)
    stream = torch.cuda.current_stream(dev)
    sp = vp(stream.cuda_stream)

    kernel_keys = ["ms_k_template", "ms_k_cluster", "ms_k_solve_mixed", "ms_k_solve_fp64"]
    count_keys = ["n_k_template", "n_k_cluster", "n_k_solve_mixed", "n_k_solve_fp64"]
    acc = {}

    def step(record):
        d_out.copy_(d_in)
        d_out_mask.copy_(d_mask)
        lib.hpgl_amd_simple_kriging_dev(d_in.data_ptr(), d_mask.data_ptr(), C.byref(shape), C.byref(skp),
                                        d_out.data_ptr(), d_out_mask.data_ptr(), sp)
        hpgl_wrap.check_last_error()
        s1 = hpgl_wrap.last_stats()
        d_iout.copy_(d_cat)
        d_iout_mask.copy_(d_mask)
        lib.hpgl_amd_indicator_kriging_dev(d_cat.data_ptr(), d_mask.data_ptr(), C.byref(shape), ikp, len(ik),
                                           d_iout.data_ptr(), d_iout_mask.data_ptr(), sp)
        hpgl_wrap.check_last_error()
        s2 = hpgl_wrap.last_stats()
        if record:
            for s in (s1, s2):
                for k, v in s.items():
                    acc[k] = acc.get(k, 0) + v

    hpgl_wrap.set_precision(args.precision)
    for _ in range(args.warmup):
        step(False)
    torch.cuda.synchronize()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    ev0.record(stream)
    for _ in range(args.steps):
        step(True)
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_per_step = wall * 1000.0 / args.steps
    value = nodes_per_step * args.steps / wall

    # dominant kernel and its roofline.  Algorithmic bytes per kriged node:
    # 5 B read (centre value + mask) + 5 B written + 8 B per neighbour
    # (value + index); flops per SK system of n neighbours:
    # n^3/3 + n^2/2 + n/6 (Cholesky) + 2n^2 (two triangular solves)
    # + 2n + 2n (sums) + 20 (n(n+1)/2 + n + 1) (covariance entries).
    dom = max(range(4), key=lambda i: acc.get(kernel_keys[i], 0.0))
    dom_name = ["k_template_scan", "k_cluster", "k_solve_mixed", "k_solve_fp64"][dom]
    launches = max(1, acc.get(count_keys[dom], 0))
    avg_ms = acc.get(kernel_keys[dom], 0.0) / launches
    n_avg = acc["neighbours_total"] / max(1, acc["nodes_target"])
    bytes_per_node = 10.0 + 8.0 * n_avg
    dom_nodes = {0: acc["nodes_template_path"], 1: acc["nodes_cluster_path"]}.get(
        dom, acc["solves_fp32"] + acc["solves_fp64"])
    alg_bytes = bytes_per_node * dom_nodes / launches
    achieved = alg_bytes / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else 0.0
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic_sk_ik_256x256x128.md")
    if os.path.exists(tpath):
        for line in open(tpath):
            if line.startswith("bench-traffic %s bytes_per_launch=" % dom_name):
                traffic = int(line.split("=", 1)[1])
    roofline = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 6), "traffic": traffic, "kernel": dom_name,
                "avg_launch_ms": round(avg_ms, 4), "algorithmic_bytes_per_launch": int(alg_bytes),
                "bytes_per_node": round(bytes_per_node, 2)}
    n = n_avg
    flops_per_system = n ** 3 / 3 + n ** 2 / 2 + n / 6 + 2 * n ** 2 + 4 * n + 20 * (n * (n + 1) / 2 + n + 1)
    solve_ms = acc.get("ms_k_solve_mixed", 0.0) + acc.get("ms_k_solve_fp64", 0.0)
    systems = acc["solves_fp32"] + acc["solves_fp64"] + acc["nodes_singular"]
    solve_tflops = flops_per_system * systems / (solve_ms * 1e-3) / 1e12 if solve_ms > 0 else 0.0
    roofline_solve = {"bound": "fp32-valu" if args.precision == "mixed" else "fp64-valu",
                      "achieved": round(solve_tflops, 3),
                      "peak": FP32_PEAK_TFLOPS if args.precision == "mixed" else FP64_PEAK_TFLOPS,
                      "unit": "TFLOP/s", "flops_per_system": round(flops_per_system, 1)}
    roofline_solve["frac"] = round(solve_tflops / roofline_solve["peak"], 6)

    out = {
        "metric": "kriged nodes/s, mixed-prec vs plain FP64 at stated err; p50/p99 job latency",
        "value": value,
        "unit": "kriged nodes/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64" if args.precision == "mixed" else "f64",
        "data": "synthetic",
        "config": {"workload": "SK + 5-category IK, 256x256x128 grid, 1% informed, max_neighbours 24 "
                               "(radii = ranges)", "cells": g.size, "kriged_nodes_per_step": nodes_per_step,
                   "precision": args.precision},
        "roofline": roofline,
        "roofline_solve": roofline_solve,
        "kernel_ms_per_step": {k: round(acc.get(k, 0.0) / args.steps, 3) for k in kernel_keys},
        "solves": {"fp32_accepted": int(acc.get("solves_fp32", 0)), "fp64": int(acc.get("solves_fp64", 0)),
                   "singular": int(acc.get("nodes_singular", 0)),
                   "no_neighbours": int(acc.get("nodes_no_neighbours", 0))},
        "paths": {"template": int(acc.get("nodes_template_path", 0)), "cluster": int(acc.get("nodes_cluster_path", 0)),
                  "cluster_candidates": int(acc.get("cluster_candidates", 0)),
                  "cluster_lists_spilled": int(acc.get("cluster_lists_spilled", 0))},
        "host_setup_ms_per_step": round(acc.get("ms_setup_host", 0.0) / args.steps, 2),
        "device_ms_per_step": round(float(ev0.elapsed_time(ev1)) / args.steps, 3),
    }

    if not args.no_fp64_compare and args.precision == "mixed":
        mixed_sk = d_out.cpu().numpy()
        mixed_ik = d_iout.cpu().numpy()
        hpgl_wrap.set_precision("fp64")
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        step(False)
        torch.cuda.synchronize()
        t64 = time.perf_counter() - t0
        hpgl_wrap.set_precision("mixed")
        ref = d_out.cpu().numpy().astype(np.float64)
        m = mask.ravel(order="F") == 0
        rel = np.abs(mixed_sk[m].astype(np.float64) - ref[m]) / np.maximum(np.abs(ref[m]), 1e-3)
        out["fp64_compare"] = {"fp64_value": nodes_per_step / t64, "fp64_ms_per_step": t64 * 1000,
                               "sk_max_rel_err_vs_fp64": float(rel.max()),
                               "sk_rms_rel_err_vs_fp64": float(np.sqrt(np.mean(rel ** 2))),
                               "ik_category_agreement": float(np.mean(mixed_ik == d_iout.cpu().numpy()))}

    if not args.no_sgs:
        out["sgs"] = sgs_bench(torch, lib, abi, hpgl_wrap, cpu=not args.no_cpu_baseline)
    if not args.no_sis:
        out["sis"] = sis_bench(torch, lib, abi, hpgl_wrap, cpu=not args.no_cpu_baseline)
    if not args.no_cokriging:
        out["cokriging"] = cokriging_bench(torch, lib, abi, hpgl_wrap, cpu=not args.no_cpu_baseline)
    if not args.no_variogram:
        out["variogram"] = variogram_bench(cpu=not args.no_cpu_baseline)
    if not args.no_io:
        import hpgl_amd.geo_bsd as geo_api

        shape = (g.x, g.y, g.z)
        out["property_io"] = property_io_bench(geo_api, d_out.cpu().numpy().reshape(shape, order="F"),
                                               d_out_mask.cpu().numpy().reshape(shape, order="F"))
    if not args.no_serving:
        trace, out["serving"] = serving_bench()
        if not args.no_cpu_baseline:
            out["serving"]["cpu_baseline"] = serving_cpu_baseline(trace)
        del trace
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(g, data, mask, cat, marg, sk, ik)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
