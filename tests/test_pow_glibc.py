"""The device's power function (se_amd/csrc/glibc_pow.h), compiled as host
code, against the C library's pow on the same inputs, bit for bit.

The reference's Legendre evaluation calls pow(x, i) with small integer i
(Jacobi.cpp:43,53,63,73); the GPU path computes the same powers with
se_pow::pow_int, so this pins every power the semi-Lagrangian kernels use.
The C library here is glibc 2.35 with its FMA variant of pow (what the
reference's golden vectors were made with)."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "se_amd", "csrc")


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_pow_int_matches_libm(tmp_path):
    exe = str(tmp_path / "pow_check")
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-Wno-unknown-pragmas", "-I" + CSRC,
                    os.path.join(HERE, "pow_check.cpp"), "-o", exe], check=True)
    r = subprocess.run([exe, "3000000"], capture_output=True, text=True)
    print(r.stdout)
    assert r.returncode == 0, r.stdout
    assert "mismatches 0 of" in r.stdout


def test_generated_tables_are_current():
    """glibc_pow_tables.h is what gen_pow_tables.py writes."""
    pytest.importorskip("mpmath")
    r = subprocess.run(["python3", os.path.join(CSRC, "gen_pow_tables.py")], capture_output=True, text=True,
                       check=True)
    with open(os.path.join(CSRC, "glibc_pow_tables.h")) as f:
        assert f.read() == r.stdout
