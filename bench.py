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
CU_COUNT = 256
CLOCK_HZ = 2.4e9  # peak engine clock
FP32_PEAK_TFLOPS = 157.3
FP64_PEAK_TFLOPS = 78.6


def cpu_setup(build_native=True):
    """Host set-up of the CPU baseline, before the oracle library is loaded:
    one OpenMP thread per physical core this process may run on (capped by
    the OMP_NUM_THREADS share when one is set) and the oracle rebuilt with
    -march=native on this machine (the reference's release flags,
    SConstruct:24) when the compiler is there.  The legs run in child
    processes with OMP_PROC_BIND=close and OMP_PLACES=cores (cpu_leg).
    Returns what was used, with lscpu's totals."""
    import subprocess

    try:
        out = subprocess.run(["lscpu", "-p=CPU,CORE,SOCKET"], capture_output=True, text=True, check=True).stdout
        rows = [tuple(int(v) for v in line.split(",")[:3]) for line in out.splitlines()
                if line and not line.startswith("#")]
    except Exception:
        rows = [(c, c, 0) for c in range(os.cpu_count() or 1)]
    allowed = os.sched_getaffinity(0)
    phys_all = {(core, sock) for _, core, sock in rows}
    phys_allowed = {(core, sock) for cpu, core, sock in rows if cpu in allowed} or phys_all
    share = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    threads = min(len(phys_allowed), share) if share > 0 else len(phys_allowed)
    flags = "-O2 -fopenmp -ffp-contract=off"
    so = None
    if build_native:
        r = subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "native"], capture_output=True,
                           text=True)
        if r.returncode == 0:
            so = os.path.join(ROOT, "oracle", "_build", "liboracle_native.so")
            os.environ["ORACLE_SO"] = so  # the legs that run the oracle in this process load it too
            flags += " -march=native"
    return {"threads": int(threads), "oracle_so": so, "lscpu_physical_cores": len(phys_all),
            "lscpu_sockets": len({s for _, _, s in rows}), "physical_cores_allowed": len(phys_allowed),
            "omp_num_threads_share": share or None, "omp_proc_bind": "close", "omp_places": "cores",
            "oracle_flags": flags}


CPU = {"threads": 16}  # filled by cpu_setup() when a CPU baseline runs


def progress(msg):
    """A progress line on stderr (stdout carries only the JSON line)."""
    print("[bench %s] %s" % (time.strftime("%H:%M:%S"), msg), file=sys.stderr, flush=True)


def median_time(fn, runs=3):
    """Median wall time of `runs` calls of fn (and the last call's value)."""
    ts = []
    out = None
    for _ in range(runs):
        t0 = time.perf_counter()
        out = fn()
        ts.append(time.perf_counter() - t0)
    return float(np.median(ts)), [round(t, 3) for t in ts], out


def err_stats(got, want, floor):
    """max and RMS of |got - want| / max(|want|, floor) over the cells given."""
    got = np.asarray(got, np.float64).ravel()
    want = np.asarray(want, np.float64).ravel()
    if got.size == 0:
        return {"max_rel": 0.0, "rms_rel": 0.0, "cells": 0, "floor": floor}
    rel = np.abs(got - want) / np.maximum(np.abs(want), floor)
    return {"max_rel": float(rel.max()), "rms_rel": float(np.sqrt(np.mean(rel ** 2))), "cells": int(rel.size),
            "floor": floor}


def output_rounding(got, want):
    """How far float32 outputs of the mixed path sit from the FP64 path's:
    cells whose floats differ, the largest difference in absolute terms, in
    units of the data's standard deviation (1 for the standard-normal
    workloads) and in float32 spacings of the FP64 path's value.  Weights
    that differ in their last digits flip the rounding of single float32
    terms of the combine sum (as the reference forms it), so a difference of
    a few spacings is rounding, not solve error; relative to a small floor
    such a flip on an estimate near zero reads large."""
    got = np.asarray(got, np.float32).ravel()
    want = np.asarray(want, np.float32).ravel()
    d = np.abs(got.astype(np.float64) - want.astype(np.float64))
    sp = np.spacing(np.maximum(np.abs(want), np.float32(1e-30))).astype(np.float64)
    return {"cells_differing": int(np.count_nonzero(got != want)), "max_abs_err": float(d.max()) if d.size else 0.0,
            "max_err_in_float_spacings": float((d / sp).max()) if d.size else 0.0}


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


def cpu_leg(spec, timeout=900):
    """Runs one CPU-baseline leg of this script in a child process, so that
    the oracle's OpenMP runtime starts with this leg's environment
    (OMP_PROC_BIND=close, OMP_PLACES=cores, one thread per physical core)
    and binds none of this process's threads; returns the child's summary
    and the arrays it wrote."""
    import subprocess
    import tempfile

    progress("CPU leg %s (oracle, %d threads)" % (spec["leg"], CPU["threads"]))
    env = dict(os.environ)
    env.update({"OMP_PROC_BIND": "close", "OMP_PLACES": "cores", "OMP_NUM_THREADS": str(CPU["threads"])})
    if CPU.get("oracle_so"):
        env["ORACLE_SO"] = CPU["oracle_so"]
    with tempfile.TemporaryDirectory() as tmp:
        spec = dict(spec, out=os.path.join(tmp, "leg.npz"), threads=CPU["threads"])
        r = subprocess.run([sys.executable, os.path.abspath(__file__), "--cpu-worker", json.dumps(spec)], env=env,
                           capture_output=True, text=True, timeout=timeout)
        if r.returncode != 0:
            raise RuntimeError("CPU leg %s failed: %s" % (spec["leg"], r.stderr[-2000:]))
        with np.load(spec["out"]) as z:
            arrays = {k: z[k] for k in z.files}
    return json.loads(r.stdout.strip().splitlines()[-1]), arrays


def _sample_stride(run, cells, budget_s, pilot):
    """Stride s such that run(s) (which estimates every s-th cell) takes about
    budget_s: a pilot run at stride `pilot`, then a second one at the stride
    it suggests (capped at ten times the pilot's samples), whose time per
    sample has the run's fixed set-up cost spread over more cells."""
    stride = pilot
    for _ in range(2):
        t0 = time.perf_counter()
        run(stride)
        per_sample = (time.perf_counter() - t0) / (cells / stride)
        nxt = max(1, int(cells * per_sample / budget_s))
        if nxt >= stride:
            return nxt
        stride = max(nxt, stride // 10)
    return stride


def _worker_headline(O, spec, arrays):
    g, data, mask, cat, marg, sk, ik = make_cfg2()
    ikd = [{"cov_model": p["cov"], "radiuses": p["radii"], "max_neighbours": p["max_n"]} for p in ik]
    res = {}

    def run(stride):
        O.set_sample_stride(stride)
        t0 = time.perf_counter()
        res["sk"] = O.simple_kriging((data, mask), g, sk["radii"], sk["max_n"], sk["cov"], mean=sk["mean"])
        res["t_sk"] = time.perf_counter() - t0
        res["ik"] = O.indicator_kriging((cat, mask, 5), g, ikd, marg)
        O.set_sample_stride(1)

    stride = _sample_stride(run, g.size, spec["budget_s"], 4096)
    med, all_t, _ = median_time(lambda: run(stride), spec["runs"])
    arrays.update(sk_d=res["sk"][0].ravel(order="F"), sk_m=res["sk"][1].ravel(order="F"),
                  ik_c=res["ik"][0].ravel(order="F"))
    n_unk = int(np.sum(mask == 0))
    return {"stride": stride, "seconds": med, "runs_s": all_t, "kriged_nodes": n_unk / stride * 6}


def _worker_cfg1(O, spec, arrays):
    g, data, mask, cov, radii, max_n = make_cfg1()
    res = {}

    def run():
        res["ok"] = O.ordinary_kriging((data, mask), g, radii, max_n, cov)

    med, all_t, _ = median_time(run, spec["runs"])
    arrays.update(d=res["ok"][0].ravel(order="F"), m=res["ok"][1].ravel(order="F"))
    return {"seconds": med, "runs_s": all_t}


def _worker_cfg3(O, spec, arrays):
    g, data, mask, twins, cov, radii, max_n = make_cfg3()
    res = {}

    def run(stride):
        O.set_sample_stride(stride)
        res["ok"] = O.ordinary_kriging((data, mask), g, radii, max_n, cov)
        O.set_sample_stride(1)

    stride = _sample_stride(run, g.size, spec["budget_s"], 997)
    med, all_t, _ = median_time(lambda: run(stride), spec["runs"])
    arrays.update(d=res["ok"][0].ravel(order="F"), m=res["ok"][1].ravel(order="F"))
    return {"stride": stride, "seconds": med, "runs_s": all_t}


def _worker_sgs(O, spec, arrays):
    g, data, mask, cov, radii, max_n = make_sgs_workload()
    P = spec["threads"]
    t0 = time.perf_counter()
    bd, _ = O.sgs_simulation_batch((data, mask), g, radii, max_n, cov, [1000 + r for r in range(P)])
    t_batch = time.perf_counter() - t0
    arrays.update(batch=bd)
    return {"batch_s": t_batch, "batch_realisations": P}


def _worker_serving(O, spec, arrays):
    from cases import Cov, Grid

    jobs = make_serving_trace()
    service = []
    nodes_all = 0
    for i, job in enumerate(jobs):
        g = Grid(*job["shape"])
        n_unk = int(np.sum(job["mask"] == 0))
        nodes = g.size * 3 if job["kind"] == "ik" else n_unk
        stride = max(1, g.size // spec["nodes_per_sample"])
        cov = Cov(job["ctype"], job["ranges"], sill=1.0, nugget=0.1)
        O.set_sample_stride(stride)
        t0 = time.perf_counter()
        if job["kind"] == "ok":
            od, om = O.ordinary_kriging((job["data"], job["mask"]), g, job["radii"], job["max_n"], cov)
        elif job["kind"] == "sk":
            od, om = O.simple_kriging((job["data"], job["mask"]), g, job["radii"], job["max_n"], cov)
        elif job["kind"] == "lvm":
            od, om = O.lvm_kriging((job["data"], job["mask"]), g, job["mean"], job["radii"], job["max_n"], cov)
        else:
            ikd = [{"cov_model": cov, "radiuses": job["radii"], "max_neighbours": job["max_n"]} for _ in range(3)]
            od, om = O.indicator_kriging((job["cat"], job["mask"], 3), g, ikd, job["marg"])
        dt = time.perf_counter() - t0
        O.set_sample_stride(1)
        sel = np.arange(g.size) % stride == 0
        arrays["d%d" % i] = od.ravel(order="F")[sel]
        arrays["m%d" % i] = om.ravel(order="F")[sel]
        service.append(dt * stride)
        nodes_all += nodes
    end = 0.0
    lat = []
    for job, sv in zip(jobs, service):
        start = max(job["t"], end)
        end = start + sv
        lat.append((end - job["t"]) * 1000.0)
    return {"p50_latency_ms": round(float(np.percentile(lat, 50)), 1),
            "p99_latency_ms": round(float(np.percentile(lat, 99)), 1), "nodes_per_s": nodes_all / end}


def cpu_worker(spec):
    """Child-process side of cpu_leg: loads the oracle, runs the leg and
    writes its arrays; prints the leg's summary as the last line."""
    import oracle_binding as O

    O.set_threads(spec["threads"])
    arrays = {}
    fn = {"headline": _worker_headline, "cfg1": _worker_cfg1, "cfg3": _worker_cfg3, "sgs": _worker_sgs,
          "serving": _worker_serving}[spec["leg"]]
    out = fn(O, spec, arrays)
    out["omp_max_threads"] = O.lib.oracle_get_threads()
    np.savez(spec["out"], **arrays)
    print(json.dumps(out), flush=True)


def cpu_baseline(g, mask, gpu=None, budget_s=5.0, runs=3):
    """The CPU oracle (a C++/OpenMP restatement of the reference's CPU path)
    on a bounded sample of the same SK + IK workload, in a child process:
    every s-th cell of the grid is estimated, s chosen from a short pilot
    so that one pass takes about budget_s; the median of `runs` passes is
    reported.  The GPU outputs in `gpu` (SK values in mixed and FP64 mode,
    IK categories) are compared with the oracle's on the sampled cells."""
    info, arr = cpu_leg({"leg": "headline", "budget_s": budget_s, "runs": runs})
    stride = info["stride"]
    nodes = info["kriged_nodes"]
    out = {"value": nodes / info["seconds"], "unit": "kriged nodes/s", "cores": CPU["threads"], "kind": "port",
           "sample": "every %d-th cell of the same SK+IK workload (%.0f kriged nodes), median of %d runs %s s"
                     % (stride, nodes, runs, info["runs_s"]), "host": dict(CPU, omp_max_threads=info["omp_max_threads"])}
    if gpu is not None:
        samp = np.arange(g.size) % stride == 0
        sel = samp & (mask.ravel(order="F") == 0)
        od, om, ic = arr["sk_d"], arr["sk_m"], arr["ik_c"]
        cmp = {"sampled_cells": int(sel.sum())}
        for key in ("sk_mixed", "sk_fp64"):
            if key in gpu:
                cmp[key + "_vs_oracle"] = err_stats(gpu[key][sel], od[sel], 1e-3)
                cmp[key + "_masks_equal"] = bool(np.array_equal(gpu[key + "_mask"][sel], om[sel]))
        for key in ("ik_mixed", "ik_fp64"):
            if key in gpu:
                cmp[key + "_category_agreement_vs_oracle"] = float(np.mean(gpu[key][samp] == ic[samp]))
        out["parity"] = cmp
    return out


def make_cfg3(shape=(300, 300, 100), seed=1003):
    """Precision-stress workload: informed density varying smoothly over the
    grid from 0.1% to 5% (logistic of a Gaussian-filtered noise field,
    sigma 15 cells), values N(0, 1), 0.5% of the samples with a twin in an
    x-adjacent cell (value + N(0, 1e-3)); ordinary kriging with a Gaussian
    covariance of range 25, sill 1, nugget 1e-4, radii 25/25/25 (radii are
    not given by the workload; chosen equal to the range), max_neighbours 64."""
    from cases import Cov, Grid, uneven_twins

    g = Grid(*shape)
    data, mask, twins = uneven_twins(shape, seed)
    return g, data, mask, twins, Cov(2, (25, 25, 25), sill=1.0, nugget=1e-4), (25, 25, 25), 64


def cfg3_bench(torch, lib, abi, hpgl_wrap, steps=1, cpu=True, budget_s=5.0, runs=3):
    """Batched uneven OK on the precision-stress workload, inputs resident in
    HBM: kriged nodes/s in mixed mode and in FP64 mode, the FP32-accepted /
    FP64 / singular counts of each, the neighbour-count histogram, and the
    mixed result's error against FP64.  The CPU leg runs the oracle on every
    s-th cell (median of `runs` passes) and compares both GPU results with
    it on those cells (masks = the dpotrf decisions, values)."""
    g, data, mask, twins, cov, radii, max_n = make_cfg3()
    p = abi._HPGL_OK_PARAMS(radiuses=(C.c_int * 3)(*radii), max_neighbours=max_n, covariance_type=cov.type,
                            ranges=(C.c_double * 3)(*cov.ranges), angles=(C.c_double * 3)(*cov.angles),
                            sill=cov.sill, nugget=cov.nugget)
    dev = torch.device("cuda:0")
    d_in = torch.from_numpy(data.ravel(order="F").copy()).to(dev)
    d_m = torch.from_numpy(mask.ravel(order="F").copy()).to(dev)
    d_out = torch.empty_like(d_in)
    d_om = torch.empty_like(d_m)
    shape = abi._HPGL_SHAPE(data=(C.c_int * 3)(*g.shape), strides=(C.c_int * 3)(1, g.x, g.x * g.y))
    stream = torch.cuda.current_stream(dev)
    sp = C.c_void_p(stream.cuda_stream)
    n_unk = int(np.sum(mask == 0))

    def run():
        d_out.copy_(d_in)
        d_om.copy_(d_m)
        lib.hpgl_amd_ordinary_kriging_dev(d_in.data_ptr(), d_m.data_ptr(), C.byref(shape), C.byref(p),
                                          d_out.data_ptr(), d_om.data_ptr(), sp)
        hpgl_wrap.check_last_error()
        return hpgl_wrap.last_stats()

    res = {"workload": "OK, 300x300x100 grid, informed density 0.1-5%% (logistic field), %d near-duplicate "
                       "twins, Gaussian range 25, nugget 1e-4, radii 25/25/25, max_neighbours 64" % twins,
           "cells": g.size, "informed": int(mask.sum()), "kriged_nodes_per_job": n_unk}
    outs = {}
    for prec in ("mixed", "fp64"):
        hpgl_wrap.set_precision(prec)
        run()  # warm: assets, code objects
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        sts = [run() for _ in range(steps)]
        torch.cuda.synchronize()
        wall = (time.perf_counter() - t0) / steps
        st = sts[-1]
        res[prec] = {"value": n_unk / wall, "unit": "kriged nodes/s", "ms_per_job": wall * 1e3,
                     "solves": {"fp32_accepted": int(st["solves_fp32"]), "fp32_two_steps": int(st["solves_fp32_two_steps"]),
                                "fp64": int(st["solves_fp64"]),
                                "singular": int(st["nodes_singular"]),
                                "no_neighbours": int(st["nodes_no_neighbours"])},
                     "kernel_ms": {k: round(st[k], 3) for k in ("ms_k_template", "ms_k_cluster", "ms_k_solve_mixed",
                                                                "ms_k_solve_fp64")},
                     "paths": {"template": int(st["nodes_template_path"]), "cluster": int(st["nodes_cluster_path"])},
                     "mean_neighbours": st["neighbours_total"] / max(1, st["nodes_target"])}
        outs[prec] = (d_out.cpu().numpy(), d_om.cpu().numpy())
    hpgl_wrap.set_precision("mixed")
    unk = mask.ravel(order="F") == 0
    ok = unk & (outs["fp64"][1] == 1)
    res["mixed_vs_fp64"] = err_stats(outs["mixed"][0][ok], outs["fp64"][0][ok], 1e-3)
    res["mixed_vs_fp64"]["masks_equal"] = bool(np.array_equal(outs["mixed"][1], outs["fp64"][1]))
    if cpu:
        info, arr = cpu_leg({"leg": "cfg3", "budget_s": budget_s, "runs": runs})
        stride = info["stride"]
        od, om = arr["d"], arr["m"]
        sel = unk & (np.arange(g.size) % stride == 0)
        n_s = int(sel.sum())
        res["cpu_baseline"] = {"value": n_s / info["seconds"], "unit": "kriged nodes/s", "cores": CPU["threads"],
                               "kind": "port", "sample": "every %d-th cell (%d kriged nodes), median of %d runs %s s"
                                                         % (stride, n_s, runs, info["runs_s"])}
        par = {"sampled_cells": n_s, "oracle_failures": int(np.sum(om[sel] == 0))}
        for prec in ("mixed", "fp64"):
            d, m = outs[prec]
            par[prec + "_masks_equal"] = bool(np.array_equal(m[sel], om[sel]))
            both = sel & (om == 1)
            par[prec + "_vs_oracle"] = err_stats(d[both], od[both], 1e-3)
            par[prec + "_bit_exact_cells"] = float(np.mean(d[both] == od[both])) if both.any() else 1.0
        res["parity"] = par
    return res


def make_cfg1(seed=1001):
    """Ordinary kriging on 100x100x50 (500,000 cells), 2% of cells informed
    uniformly, values lognormal(0, 1); spherical covariance, ranges
    20/15/5, sill 1, nugget 0.1; search radii 20/20/10; 12 neighbours."""
    from cases import Cov, Grid, scattered

    g = Grid(100, 100, 50)
    data, mask = scattered(g.shape, 0.02, seed)
    return g, data, mask, Cov(0, (20, 15, 5), sill=1.0, nugget=0.1), (20, 20, 10), 12


def cfg1_bench(torch, lib, abi, hpgl_wrap, steps=5, cpu=True, runs=3):
    """The CPU-sized ordinary-kriging workload, inputs resident in HBM:
    kriged nodes/s in mixed and FP64 mode.  The CPU leg runs the oracle on
    the whole grid (median of `runs`) and compares both GPU results with it
    cell for cell."""
    g, data, mask, cov, radii, max_n = make_cfg1()
    p = abi._HPGL_OK_PARAMS(radiuses=(C.c_int * 3)(*radii), max_neighbours=max_n, covariance_type=cov.type,
                            ranges=(C.c_double * 3)(*cov.ranges), angles=(C.c_double * 3)(*cov.angles),
                            sill=cov.sill, nugget=cov.nugget)
    dev = torch.device("cuda:0")
    d_in = torch.from_numpy(data.ravel(order="F").copy()).to(dev)
    d_m = torch.from_numpy(mask.ravel(order="F").copy()).to(dev)
    d_out = torch.empty_like(d_in)
    d_om = torch.empty_like(d_m)
    shape = abi._HPGL_SHAPE(data=(C.c_int * 3)(*g.shape), strides=(C.c_int * 3)(1, g.x, g.x * g.y))
    sp = C.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
    n_unk = int(np.sum(mask == 0))

    def run():
        d_out.copy_(d_in)
        d_om.copy_(d_m)
        lib.hpgl_amd_ordinary_kriging_dev(d_in.data_ptr(), d_m.data_ptr(), C.byref(shape), C.byref(p),
                                          d_out.data_ptr(), d_om.data_ptr(), sp)
        hpgl_wrap.check_last_error()

    res = {"workload": "OK, 100x100x50 grid, 2% informed (lognormal), spherical 20/15/5, nugget 0.1, "
                       "radii 20/20/10, max_neighbours 12", "kriged_nodes_per_job": n_unk}
    outs = {}
    for prec in ("mixed", "fp64"):
        hpgl_wrap.set_precision(prec)
        run()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            run()
        torch.cuda.synchronize()
        wall = (time.perf_counter() - t0) / steps
        st = hpgl_wrap.last_stats()
        res[prec] = {"value": n_unk / wall, "unit": "kriged nodes/s", "ms_per_job": wall * 1e3,
                     "solves": {"fp32_accepted": int(st["solves_fp32"]), "fp32_two_steps": int(st["solves_fp32_two_steps"]),
                                "fp64": int(st["solves_fp64"]),
                                "singular": int(st["nodes_singular"])}}
        outs[prec] = (d_out.cpu().numpy(), d_om.cpu().numpy())
    hpgl_wrap.set_precision("mixed")
    if cpu:
        info, arr = cpu_leg({"leg": "cfg1", "runs": runs})
        od, om = arr["d"], arr["m"]
        res["cpu_baseline"] = {"value": n_unk / info["seconds"], "unit": "kriged nodes/s", "cores": CPU["threads"],
                               "kind": "port", "sample": "the whole grid, median of %d runs %s s"
                                                         % (runs, info["runs_s"])}
        unk = mask.ravel(order="F") == 0
        par = {}
        for prec in ("mixed", "fp64"):
            d, m = outs[prec]
            both = unk & (om == 1)
            par[prec + "_masks_equal"] = bool(np.array_equal(m, om))
            par[prec + "_vs_oracle"] = err_stats(d[both], od[both], 1e-3)
            par[prec + "_bit_exact_cells"] = float(np.mean(d[both] == od[both]))
        res["parity"] = par
    return res


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
        O.set_threads(CPU["threads"])
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
        res["cpu_baseline"] = {"value": sampled / dt, "unit": "estimated cells/s", "cores": CPU["threads"], "kind": "port",
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


def make_sgs_workload():
    """200x200x100 grid, 1% conditioning data (normal scores, seed 1008),
    spherical ranges 30/30/6, sill 1, nugget 0, radii 30/30/6 (radii are not
    given by the workload; chosen equal to the ranges), max_neighbours 32."""
    from cases import Cov, Grid, scattered

    g = Grid(200, 200, 100)
    data, mask = scattered(g.shape, 0.01, 1008, law="normal")
    return g, data, mask, Cov(0, (30, 30, 6), sill=1.0, nugget=0.0), (30, 30, 6), 32


def sgs_bench(torch, lib, abi, hpgl_wrap, R=64, cpu=True):
    """Sequential Gaussian simulation, R realisations batched in one call on
    the SGS workload (SK, no transform, seeds 1000..1000+R-1), inputs
    resident in HBM; simulated cells/s over all realisations.  The CPU leg
    runs the oracle on the same workload: one realisation on one thread
    (the reference's simulation is serial), then one realisation per core
    side by side (seeds 1000..1000+P-1); the GPU realisations of the same
    seeds are compared with them cell for cell."""
    g, data, mask, cov, radii, max_n = make_sgs_workload()
    p = abi._HPGL_SGS_PARAMS(covariance_type=cov.type, ranges=(C.c_double * 3)(*cov.ranges),
                             angles=(C.c_double * 3)(*cov.angles), sill=cov.sill, nugget=cov.nugget,
                             radiuses=(C.c_int * 3)(*radii), max_neighbours=max_n, kriging_kind=1, seed=1000,
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
    n_sim1 = int(np.sum(mask == 0))
    n_sim = n_sim1 * R
    res = {"workload": "SGS, %d realisations in one call, 200x200x100 grid, 1%% conditioning (normal scores), "
                       "spherical ranges 30/30/6, radii 30/30/6, max_neighbours 32, SK" % R,
           "simulated_cells": n_sim, "value": n_sim / wall, "unit": "simulated cells/s",
           "ms_per_call": wall * 1e3, "ms_per_realisation": wall * 1e3 / R,
           "ms_template_scan": st["ms_k_template"], "ms_solve": st["ms_solve"],
           "solves": {"fp32_accepted": int(st["solves_fp32"]), "fp32_two_steps": int(st["solves_fp32_two_steps"]),
                      "fp64": int(st["solves_fp64"])}}
    if cpu:
        info, arr = cpu_leg({"leg": "sgs"})
        P = info["batch_realisations"]
        res["cpu_baseline"] = {
            "value": P * n_sim1 / info["batch_s"], "unit": "simulated cells/s", "cores": CPU["threads"],
            "kind": "port",
            "sample": "%d realisations of the same workload side by side, one per core (seeds 1000..%d), %.1f s; "
                      "64 realisations take %d such rounds" % (P, 999 + P, info["batch_s"], -(-64 // P)),
            "serial_one_realisation": {"value": n_sim1 / info["batch_s"], "unit": "simulated cells/s",
                                       "cores": 1, "seconds": round(info["batch_s"], 2),
                                       "note": "each realisation runs serially on one core while the other "
                                               "cores run theirs"}}
        out = d_out.view(R, g.size)[:P].cpu().numpy()
        om = d_om.view(R, g.size)[:P].cpu().numpy()
        simc = mask.ravel(order="F") == 0
        want = arr["batch"][:, simc]
        got = out[:, simc]
        res["parity"] = {"realisations": P, "vs_oracle": err_stats(got, want, 1.0),
                         "bit_exact_cells": float(np.mean(got == want)),
                         "masks_all_set": bool(np.all(om[:, simc] == 1))}
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
    outputs = [(h.out, h.out_mask) for h in handles]
    return jobs, outputs, {"jobs": len(jobs), "arrival_rate_per_s": 20.0, "slots": slots,
                  "p50_latency_ms": round(float(np.percentile(lat, 50)), 2),
                  "p99_latency_ms": round(float(np.percentile(lat, 99)), 2),
                  "mean_queue_ms": round(float(np.mean([r["queue_ms"] for r in results])), 2),
                  "kriged_nodes": int(nodes), "nodes_per_s": nodes / makespan, "makespan_s": round(makespan, 3),
                  "trace": "Poisson 20 jobs/s, OK/SK/IK/LVM, grids 32..256 x 32..256 x 32..128, seed 1005"}


def serving_cpu_baseline(jobs, outputs, nodes_per_sample=20000):
    """The same trace processed first-in first-out by the CPU oracle (in a
    child process): each job's service time is measured on a sample of its
    cells (every s-th cell, about nodes_per_sample of them) and scaled to
    the job; latencies follow from the arrival times of a single FIFO
    server.  Every job's GPU result from the request loop is compared with
    the oracle's on the sampled cells."""
    info, arr = cpu_leg({"leg": "serving", "nodes_per_sample": nodes_per_sample})
    worst_cont = {"max_rel": 0.0, "rms_sq": 0.0, "cells": 0}
    mask_mismatch = 0
    ik_agree = []
    for i, (job, (out, out_mask)) in enumerate(zip(jobs, outputs)):
        n = int(np.prod(job["shape"]))
        stride = max(1, n // nodes_per_sample)
        sel = np.arange(n) % stride == 0
        got = out.ravel(order="F")[sel]
        gm = out_mask.ravel(order="F")[sel]
        want, wm = arr["d%d" % i], arr["m%d" % i]
        mask_mismatch += int(np.sum(gm != wm))
        if job["kind"] == "ik":
            ik_agree.append(float(np.mean(got == want)))
            continue
        unk = job["mask"].ravel(order="F")[sel] == 0
        both = unk & (wm == 1)
        e = err_stats(got[both], want[both], 1e-3)
        worst_cont["max_rel"] = max(worst_cont["max_rel"], e["max_rel"])
        worst_cont["rms_sq"] += e["rms_rel"] ** 2 * e["cells"]
        worst_cont["cells"] += e["cells"]
    parity = {"ok_sk_lvm_vs_oracle": {"max_rel": worst_cont["max_rel"],
                                      "rms_rel": float(np.sqrt(worst_cont["rms_sq"] / max(1, worst_cont["cells"]))),
                                      "cells": worst_cont["cells"], "floor": 1e-3},
              "mask_mismatches": mask_mismatch,
              "ik_category_agreement_min": min(ik_agree) if ik_agree else None}
    return {"p50_latency_ms": info["p50_latency_ms"], "p99_latency_ms": info["p99_latency_ms"],
            "nodes_per_s": info["nodes_per_s"], "cores": CPU["threads"], "kind": "port",
            "sample": "each job's cells sampled every s-th (about %d per job), service time scaled by s, "
                      "FIFO single server over the arrival trace" % nodes_per_sample, "parity": parity}


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
    ap.add_argument("--cpu-worker", default=None, help=argparse.SUPPRESS)
    ap.add_argument("--no-host-path", action="store_true", help="skip the host-pointer geo_bsd timing")
    ap.add_argument("--no-cfg1", action="store_true", help="skip the 100x100x50 ordinary-kriging measurement")
    ap.add_argument("--cfg1-only", action="store_true", help="run only the 100x100x50 ordinary-kriging measurement")
    ap.add_argument("--no-cfg3", action="store_true", help="skip the precision-stress (uneven OK) measurement")
    ap.add_argument("--cfg3-only", action="store_true", help="run only the precision-stress OK measurement")
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

    if args.cpu_worker:
        cpu_worker(json.loads(args.cpu_worker))
        return

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
    lib.hpgl_amd_ordinary_kriging_dev.argtypes = [vp, vp, C.POINTER(abi._HPGL_SHAPE), C.POINTER(abi._HPGL_OK_PARAMS),
                                                  vp, vp, vp]
    if not args.no_cpu_baseline:
        CPU.update(cpu_setup())

    if args.cfg3_only:
        print(json.dumps({"cfg3": cfg3_bench(torch, lib, abi, hpgl_wrap, steps=max(1, args.steps),
                                             cpu=not args.no_cpu_baseline)}), flush=True)
        return
    if args.cfg1_only:
        print(json.dumps({"cfg1": cfg1_bench(torch, lib, abi, hpgl_wrap, steps=max(1, args.steps),
                                             cpu=not args.no_cpu_baseline)}), flush=True)
        return

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

    progress("headline SK + IK workload (256x256x128)")
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
    # the first step builds the covariance tables, templates and rank tables
    # (then resident) and loads the code objects: the cold-call cost
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    step(True)
    torch.cuda.synchronize()
    cold = {"first_step_ms": round((time.perf_counter() - t0) * 1e3, 2),
            "host_setup_ms": round(acc.get("ms_setup_host", 0.0), 2)}
    acc.clear()
    for _ in range(max(0, args.warmup - 1)):
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
    # the cell-binned search is limited by the vector L1 / address unit, not
    # by HBM: line accesses per target and the address unit's busy fraction
    # come from the committed counter summary, the launch time from this run
    roofline_l1 = None
    lpath = os.path.join(ROOT, "profiles", "l1_sk_ik_256x256x128.md")
    if dom_name == "k_cluster" and os.path.exists(lpath):
        for line in open(lpath):
            if line.startswith("bench-l1 k_cluster "):
                kv = dict(f.split("=", 1) for f in line.split()[2:])
                lines_per_target = float(kv["line_accesses_per_target"])
                per_launch = lines_per_target * dom_nodes / launches
                achieved_l1 = per_launch / (avg_ms * 1e-3) / (CU_COUNT * CLOCK_HZ) if avg_ms > 0 else 0.0
                roofline_l1 = {"bound": "vector L1 / address unit (TA)", "unit": "cache-line accesses per CU-cycle",
                               "line_accesses_per_target": lines_per_target,
                               "achieved": round(achieved_l1, 4),
                               "ta_busy_frac_measured": float(kv["ta_busy_frac"]),
                               "source": "profiles/l1_sk_ik_256x256x128.md (counters), launch time from this run"}
    roofline = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 6), "traffic": traffic,
                "traffic_source": "profiles/traffic_sk_ik_256x256x128.md (FETCH_SIZE + WRITE_SIZE, copy-calibrated)",
                "kernel": dom_name,
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
        "roofline_l1": roofline_l1,
        "kernel_ms_per_step": {k: round(acc.get(k, 0.0) / args.steps, 3) for k in kernel_keys},
        "solves": {"fp32_accepted": int(acc.get("solves_fp32", 0)),
                   "fp32_two_steps": int(acc.get("solves_fp32_two_steps", 0)), "fp64": int(acc.get("solves_fp64", 0)),
                   "singular": int(acc.get("nodes_singular", 0)),
                   "no_neighbours": int(acc.get("nodes_no_neighbours", 0))},
        "paths": {"template": int(acc.get("nodes_template_path", 0)), "cluster": int(acc.get("nodes_cluster_path", 0)),
                  "cluster_candidates": int(acc.get("cluster_candidates", 0)),
                  "cluster_lists_spilled": int(acc.get("cluster_lists_spilled", 0))},
        "host_setup_ms_per_step": round(acc.get("ms_setup_host", 0.0) / args.steps, 2),
        "cold_call": cold,
        "device_ms_per_step": round(float(ev0.elapsed_time(ev1)) / args.steps, 3),
    }

    gpu_out = {}
    if args.precision == "mixed":
        gpu_out.update(sk_mixed=d_out.cpu().numpy(), sk_mixed_mask=d_out_mask.cpu().numpy(),
                       ik_mixed=d_iout.cpu().numpy())
    if not args.no_fp64_compare and args.precision == "mixed":
        hpgl_wrap.set_precision("fp64")
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        step(False)
        torch.cuda.synchronize()
        t64 = time.perf_counter() - t0
        hpgl_wrap.set_precision("mixed")
        gpu_out.update(sk_fp64=d_out.cpu().numpy(), sk_fp64_mask=d_out_mask.cpu().numpy(),
                       ik_fp64=d_iout.cpu().numpy())
        m = mask.ravel(order="F") == 0
        e = err_stats(gpu_out["sk_mixed"][m], gpu_out["sk_fp64"][m], 1e-3)
        out["fp64_compare"] = {"fp64_value": nodes_per_step / t64, "fp64_ms_per_step": t64 * 1000,
                               "sk_max_rel_err_vs_fp64": e["max_rel"], "sk_rms_rel_err_vs_fp64": e["rms_rel"],
                               "rel_err_floor": 1e-3,
                               "sk_masks_equal": bool(np.array_equal(gpu_out["sk_mixed_mask"],
                                                                     gpu_out["sk_fp64_mask"])),
                               "ik_category_agreement": float(np.mean(gpu_out["ik_mixed"] == gpu_out["ik_fp64"]))}
        out["fp64_compare"].update(output_rounding(gpu_out["sk_mixed"][m], gpu_out["sk_fp64"][m]))
    if not args.no_host_path:
        # the drop-in path a geo_bsd user calls: host arrays, PCIe copies,
        # the front end's clone of the input
        import hpgl_amd.geo_bsd as geo_api

        ikd = [{"cov_model": p["cov"], "radiuses": p["radii"], "max_neighbours": p["max_n"]} for p in ik]
        hpgl_wrap.set_precision(args.precision)
        t0 = time.perf_counter()
        geo_api.simple_kriging((data, mask), g, sk["radii"], sk["max_n"], sk["cov"], mean=sk["mean"])
        geo_api.indicator_kriging((cat, mask, 5), g, ikd, list(marg))
        th = time.perf_counter() - t0
        out["host_pointer_path"] = {"value": nodes_per_step / th, "unit": "kriged nodes/s", "ms_per_step": th * 1e3,
                                    "note": "geo_bsd.simple_kriging + indicator_kriging on numpy arrays: the same "
                                            "workload with host-device copies and the input clone"}

    if not args.no_cfg1:
        progress("ordinary kriging (100x100x50)")
        out["cfg1"] = cfg1_bench(torch, lib, abi, hpgl_wrap, cpu=not args.no_cpu_baseline)
    if not args.no_cfg3:
        progress("precision-stress OK (300x300x100)")
        out["cfg3"] = cfg3_bench(torch, lib, abi, hpgl_wrap, cpu=not args.no_cpu_baseline)
    if not args.no_sgs:
        progress("SGS, 64 realisations")
        out["sgs"] = sgs_bench(torch, lib, abi, hpgl_wrap, cpu=not args.no_cpu_baseline)
    if not args.no_sis:
        progress("SIS")
        out["sis"] = sis_bench(torch, lib, abi, hpgl_wrap, cpu=not args.no_cpu_baseline)
    if not args.no_cokriging:
        progress("cokriging")
        out["cokriging"] = cokriging_bench(torch, lib, abi, hpgl_wrap, cpu=not args.no_cpu_baseline)
    if not args.no_variogram:
        progress("variogram")
        out["variogram"] = variogram_bench(cpu=not args.no_cpu_baseline)
    if not args.no_io:
        import hpgl_amd.geo_bsd as geo_api

        shape = (g.x, g.y, g.z)
        out["property_io"] = property_io_bench(geo_api, d_out.cpu().numpy().reshape(shape, order="F"),
                                               d_out_mask.cpu().numpy().reshape(shape, order="F"))
    if not args.no_serving:
        progress("serving stream")
        trace, outputs, out["serving"] = serving_bench()
        if not args.no_cpu_baseline:
            progress("CPU baseline of the serving stream")
            out["serving"]["cpu_baseline"] = serving_cpu_baseline(trace, outputs)
        del trace, outputs
    if not args.no_cpu_baseline:
        progress("CPU baseline of the headline workload")
        out["cpu_baseline"] = cpu_baseline(g, mask, gpu=gpu_out)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
