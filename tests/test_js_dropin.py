"""The JavaScript drop-in modules (machinig-sim_amd/js: stabtd.js, sld.js, stock.js,
stabworker.js over the N-API addon) return the reference's answers on the golden cases."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
NODE = shutil.which("node")
needs_node = pytest.mark.skipif(NODE is None, reason="node not installed")


def run(what, timeout):
    r = subprocess.run([NODE, os.path.join(HERE, "js_dropin_check.mjs"), what], capture_output=True,
                       text=True, timeout=timeout)
    print(r.stdout[-3000:], r.stderr[-3000:])
    return r


@needs_node
def test_lobes_dropin_matches_reference():
    r = run("lobes", 120)
    assert r.returncode == 0 and "all lobes checks passed" in r.stdout


@needs_node
@pytest.mark.gpu
def test_gpu_dropin_matches_reference():
    r = run("gpu", 280)
    assert r.returncode == 0 and "all gpu checks passed" in r.stdout
