"""The FP64 triangular solves replace each division by a diagonal entry
with a product by its correctly rounded reciprocal plus one exact
remainder correction (hpgl_amd/csrc/divrn.h), which must equal the IEEE
division bit for bit for the weights to stay the reference's.
tests/native/divrn_check.c restates the sequence on the host and compares
it with a / d on 2 x 10^7 random and adversarial pairs (divisors with
all-ones significands, quotients next to 1 and to powers of two, signed
zeros)."""
import os
import subprocess
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))


def test_reciprocal_division_equals_ieee_division():
    with tempfile.TemporaryDirectory() as d:
        exe = os.path.join(d, "divrn")
        subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-o", exe, os.path.join(HERE, "native", "divrn_check.c"),
                        "-lm"], check=True)
        r = subprocess.run([exe, "20000000"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.startswith("ok"), r.stdout
