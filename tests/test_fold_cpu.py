"""The solve kernels read a covariance table's non-negative octant when
the table equals it bit for bit (BoxTable::fold, hpgl_amd/csrc/model.cpp).
tests/native/fold_check.cpp builds the tables of unrotated and rotated
models of all three types and checks that unrotated models fold with every
octant entry equal to the full table's, that rotated models do not fold,
and that a planted asymmetry is refused."""
import os
import subprocess
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


def test_covariance_table_octant_fold():
    with tempfile.TemporaryDirectory() as d:
        exe = os.path.join(d, "fold")
        subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-o", exe,
                        os.path.join(HERE, "native", "fold_check.cpp"),
                        os.path.join(ROOT, "hpgl_amd", "csrc", "model.cpp")], check=True)
        r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.startswith("ok"), r.stdout
    folded, refused = (int(v) for v in r.stdout.split()[1:3])
    assert folded == 12 and refused >= 6
