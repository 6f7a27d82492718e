"""The rounding identities behind the fluid neighbor terms' reciprocal divisions, reciprocal
refinement and square-root selection (cn_device.h fdiv_rcp / frcp_exact / fsqrt_nearzero_free),
checked on the CPU with the same binary32 operations: exhaustive over dividend significands for the
smoothing radii the tests and the bench use plus pseudo-random divisors, over every divisor
significand for the refined reciprocal from any estimate within two ulps, and over every
significand of the square-root argument in [2^-96, 2^101).  The GPU parity tests then compare
whole substeps bit for bit with the oracle."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "cpp", "quotient_identities.c")

# smoothing radii of the scenes under test: radius-0.02 particles (0.0667), the 4M random fill
# (0.0287), the dam break (0.1), 0.2, 0.3, the clamp 1e-4, the reciprocal-path limit 1024 and a
# radius above it
RADII = ["0.06666666", "0.02866256", "0.09999999", "0.2", "1e-4", "1024", "2000", "0.3"]


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
def test_reciprocal_division_and_sqrt_selection(tmp_path):
    exe = str(tmp_path / "qid")
    subprocess.run(["gcc", "-O2", "-mfma", "-ffp-contract=off", "-fopenmp", SRC, "-o", exe, "-lm"], check=True)
    r = subprocess.run([exe, "8", *RADII], capture_output=True, text=True, timeout=600)
    print(r.stdout)
    assert r.returncode == 0, r.stdout
    assert r.stdout.startswith("ok ")
