"""Writes the golden fixtures of tests/golden/ from the reference's own files.

Run in the container that mounts the reference (it reads /root/reference);
the JSON it writes is committed and the tests read only the JSON.  Nothing
of the reference is imported or executed: the script parses printed result
files and the literal data in them.

Sources:
  * sample-scripts/solved_problems/7.2. Conditioning by Kriging/
      transect.txt (input) and Result/problem_7.2.txt (printed outputs of
      problem_7_2.py: SK, SGS, second SK);
  * GsTL-1.3/test/kriging/test_kriging.cpp (8 data points, lines 150-158),
      test_kriging.result (SK and OK weights, combined estimates),
      gslib/kriging/OK_kt3d.out and SK_kt3d.out (GSLIB estimates);
  * dist/misc/test/test_data/NEW_TEST_PROP.INC, NEW_TEST_PROP_01.INC and
    mp_0_8_52_55_1.inc (55 x 52 x 1 property files the reference's test
    scripts load), copied unchanged to inc/;
  * sample-scripts/solved_problems/5.2. Variogram Calculation/
      3welldata.txt and allwelldata.txt (inputs, copied to variogram/) and
      Result/2d/1_2d_variogram.txt, Result/3d/1_3d_variogram.txt (printed
      semivariograms of 1_2d_variogram.py and 1_3d_variogram.py);
  * andrey/test_data/BIG_SOFT_DATA_160_141_20.INC and
    BIG_SOFT_DATA_CON_160_141_20.INC (the 166 x 141 x 20 inputs of
    geo_testing/test_compare.py), parsed into big_soft_data.npz.
"""
import json
import os
import re
import shutil

REF = "/root/reference/src"
HERE = os.path.dirname(os.path.abspath(__file__))


def _floats(text):
    return [float(t) for t in re.findall(r"[-+]?\d*\.\d+(?:[eE][-+]?\d+)?|[-+]?\d+(?:[eE][-+]?\d+)?", text)]


def problem_7_2():
    d = os.path.join(REF, "sample-scripts", "solved_problems", "7.2. Conditioning by Kriging")
    lines = open(os.path.join(d, "transect.txt")).read().split("\n")
    n_props = int(lines[1])
    pts = [list(map(float, ln.split())) for ln in lines[2 + n_props:] if ln.strip()]
    # problem_7_2.py:25-60: 20 cells, x0 = min(x) - 5, cell size 1 + (max-min)/20,
    # cell i = int((x - x0) / size).
    xs = [p[0] for p in pts]
    x0 = min(xs) - 5
    size = 1 + (max(xs) - min(xs)) / 20
    cells = [int((p[0] - x0) / size) for p in pts]
    vals = [p[1] for p in pts]
    res = open(os.path.join(d, "Result", "problem_7.2.txt")).read()
    blocks = re.split(r"(SK result:|SGS result:|Simple Kriging result:|Error Field:|Cond\.Sim:)", res)
    out = {}
    for k in range(1, len(blocks), 2):
        out[blocks[k].rstrip(":")] = _floats(blocks[k + 1])
    return {
        "source": "sample-scripts/solved_problems/7.2. Conditioning by Kriging (transect.txt, Result/problem_7.2.txt)",
        "grid": [20, 1, 1],
        "cells": cells,
        "values": vals,
        "sk": {"cov_type": 1, "ranges": [5, 1, 1], "sill": 1.0, "nugget": 0.0, "radiuses": [5, 1, 1],
               "max_neighbours": 2, "mean": None},
        "sgs": {"cov_type": 1, "ranges": [5, 1, 1], "sill": 1.0, "nugget": 0.0, "radiuses": [2, 1, 1],
                "max_neighbours": 2, "seed": 3244759, "kriging_type": "sk",
                "cdf_from": "informed values of the SK input (calc_cdf)"},
        "expected_sk": out["SK result"],
        "expected_sgs": out["SGS result"],
        "expected_sk2": out["Simple Kriging result"],
    }


def gstl_parta():
    d = os.path.join(REF, "GsTL-1.3", "test", "kriging")
    src = open(os.path.join(d, "test_kriging.cpp")).read()
    pts = {}
    for name, x, y, v in re.findall(r"Node (P\d)\((\d+),(\d+),([-\d.]+)\)", src):
        pts[name] = (int(x), int(y), float(v))
    order = ["P1", "P2", "P3", "P4", "P5", "P6", "P7", "P8"]  # voisin.add_node order
    res = open(os.path.join(d, "test_kriging.result")).read()
    sk_block = res.split("Simple kriging")[1].split("Ordinary kriging")[0]
    ok_block = res.split("Ordinary kriging")[1].split("Kriging with Trend")[0]
    sk_w = _floats(sk_block.split("Here are the weights:")[1])[1:]
    ok_w = _floats(ok_block.split("Here are the weights:")[1])[1:]
    comb = res.split("Combining the weights")[1]
    sk_mean, sk_var = _floats(comb.split("SK mean:")[1])[:2]
    ok_mean, ok_var = _floats(comb.split("OK mean:")[1])[:2]
    gs_ok = _floats(open(os.path.join(d, "gslib", "kriging", "OK_kt3d.out")).read().split("EstimationVariance")[1])
    gs_sk = _floats(open(os.path.join(d, "gslib", "kriging", "SK_kt3d.out")).read().split("EstimationVariance")[1])
    return {
        "source": "GsTL-1.3/test/kriging (test_kriging.cpp, test_kriging.result, gslib/kriging/*_kt3d.out)",
        "covariance": "exponential, range 10, sill 1, isotropic, 2D",
        "center": [41, 29],
        "points": [[pts[n][0], pts[n][1]] for n in order],
        "values": [pts[n][2] for n in order],
        "sk_weights": sk_w[:8],
        "ok_weights": ok_w[:8],
        "ok_lagrange": ok_w[8],
        "sk_mean_for_combination": 9.0,
        "sk_estimate_mean9": sk_mean,
        "sk_variance": sk_var,
        "ok_estimate": ok_mean,
        "ok_variance": ok_var,
        "gslib_ok": gs_ok[:2],
        "gslib_sk_mean0": gs_sk[:2],
    }


def inc_files():
    src = os.path.join(REF, "dist", "misc", "test", "test_data")
    os.makedirs(os.path.join(HERE, "inc"), exist_ok=True)
    for name in ["NEW_TEST_PROP.INC", "NEW_TEST_PROP_01.INC", "mp_0_8_52_55_1.inc"]:
        shutil.copyfile(os.path.join(src, name), os.path.join(HERE, "inc", name))
        print("copied", name)


def problem_5_2():
    """The point-set semivariograms of solved problem 5.2 (printed results of
    its scripts, Result/2d/1_2d_variogram.txt and Result/3d/1_3d_variogram.txt)
    with the scripts' search templates (1_2d_variogram.py:36-79,
    1_3d_variogram.py:104-116); the two well-data files are copied beside."""
    d = os.path.join(REF, "sample-scripts", "solved_problems", "5.2. Variogram Calculation")
    out = os.path.join(HERE, "variogram")
    os.makedirs(out, exist_ok=True)
    for name in ["3welldata.txt", "allwelldata.txt"]:
        shutil.copyfile(os.path.join(d, name), os.path.join(out, name))
    r2 = open(os.path.join(d, "Result", "2d", "1_2d_variogram.txt")).read()
    v2, h2 = r2.split("Horizontal")
    r3 = open(os.path.join(d, "Result", "3d", "1_3d_variogram.txt")).read()
    ell = lambda R2, R3, az, dip: {"R1": 1, "R2": R2, "R3": R3, "azimuth": az, "dip": dip, "rotation": 0}
    return {
        "2d_vertical": {"file": "3welldata.txt", "x": "X(Easting)(m)", "y": None, "z": "Z(m)", "value": "Por",
                        "lag_width": 4, "lag_separation": 4, "tol_distance": 4, "num_lags": 8,
                        "ellipsoid": ell(0.1, 0.1, 0, 90), "expected": _floats(v2.split(":", 1)[1])},
        "2d_horizontal": {"file": "3welldata.txt", "x": "X(Easting)(m)", "y": None, "z": "Z(m)", "value": "Por",
                          "lag_width": 1000, "lag_separation": 1000, "tol_distance": 1000, "num_lags": 3,
                          "ellipsoid": ell(0.1, 4.5, 0, 0), "expected": _floats(h2.split(":", 1)[1])},
        "3d_vertical": {"file": "allwelldata.txt", "x": "X", "y": "Y", "z": "Z", "value": "Por",
                        "lag_width": 4, "lag_separation": 4, "tol_distance": 1.5, "num_lags": 11,
                        "ellipsoid": ell(0.1, 10, 0, 90), "expected": _floats(r3.split(":", 1)[1])},
    }


def _read_inc(path):
    """Values of a property file: after the name line (the first line that
    starts with a letter), white-space separated, "--" comments to the end of
    the line, "/" terminator."""
    vals = []
    with open(path) as f:
        for line in f:
            if line[:1].isalpha():
                break
        for line in f:
            t = line.split()
            if not t or t[0].startswith("--"):
                continue
            if t[0].startswith("/"):
                break
            vals.extend(float(v) for v in t)
    return vals


def big_soft_data():
    """The 166 x 141 x 20 property files of the reference's regression script
    (geo_testing/test_compare.py:30-34, andrey/test_data/): the indicator
    property (values 0/1, undefined -99) and the continuous one, stored as
    arrays in Fortran order (cell = x + 166 y + 166*141 z)."""
    import numpy as np

    d = os.path.join(REF, "andrey", "test_data")
    out = {}
    for key, name in (("ind", "BIG_SOFT_DATA_160_141_20.INC"), ("con", "BIG_SOFT_DATA_CON_160_141_20.INC")):
        v = np.array(_read_inc(os.path.join(d, name)), np.float64)
        assert v.size == 166 * 141 * 20, v.size
        mask = (v != -99).astype(np.uint8)
        out[key + "_mask"] = mask
        out[key] = np.where(mask == 1, v, -99).astype(np.float32)
    np.savez_compressed(os.path.join(HERE, "big_soft_data.npz"), **out)
    print("wrote big_soft_data.npz")


def main():
    for name, fn in [("problem_7_2.json", problem_7_2), ("gstl_parta.json", gstl_parta),
                     (os.path.join("variogram", "problem_5_2.json"), problem_5_2)]:
        with open(os.path.join(HERE, name), "w") as f:
            json.dump(fn(), f, indent=1)
        print("wrote", name)
    inc_files()
    big_soft_data()


if __name__ == "__main__":
    main()
