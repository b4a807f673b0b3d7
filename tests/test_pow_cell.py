"""pow_cell, the branch-free power the search kernel uses for the size effect
(se = (H_REF / h)^mc, src/cutting.js:197), equals pow_js bit for bit over its domain.

pow_js itself is pinned against Node's Math.pow by test_oracle_math.py; this test compiles
tests/native/pow_cell_check.cpp for the host with hipcc and runs it (no GPU involved).
"""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "..", "machinig-sim_amd", "csrc")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not installed")
def test_pow_cell_equals_pow_js(tmp_path):
    exe = str(tmp_path / "pow_cell_check")
    subprocess.run([HIPCC, "-O2", "-std=c++17", "-ffp-contract=off", "-I", CSRC,
                    os.path.join(HERE, "native", "pow_cell_check.cpp"), "-o", exe],
                   check=True, capture_output=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "mismatches 0" in r.stdout
