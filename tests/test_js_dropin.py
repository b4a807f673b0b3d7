"""The JavaScript drop-in modules (machinig-sim_amd/js: stabtd.js, sld.js, stock.js,
cutting.js, path.js, simrun.js, surrogate.js, stabworker.js over the N-API addon) return the
reference's answers on the golden cases."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
NODE = shutil.which("node")
needs_node = pytest.mark.skipif(NODE is None, reason="node not installed")


def run(what, timeout, *extra):
    r = subprocess.run([NODE, os.path.join(HERE, "js_dropin_check.mjs"), what, *extra], capture_output=True,
                       text=True, timeout=timeout)
    print(r.stdout[-3000:], r.stderr[-3000:])
    return r


@needs_node
def test_lobes_dropin_matches_reference():
    r = run("lobes", 120)
    assert r.returncode == 0 and "all lobes checks passed" in r.stdout


@needs_node
@pytest.mark.gpu
def test_gpu_dropin_matches_reference(tmp_path):
    """the device entry points through the JS modules, and timeDomainLimitsGuided with a
    surrogate model file (an untrained network: the guided search's answers do not depend
    on the prior)"""
    from mcsim import lib
    model = str(tmp_path / "surrogate.bin")
    lib.Mlp(3).save(model, sigma=0.4)
    r = run("gpu", 280, model)
    assert r.returncode == 0 and "all gpu checks passed" in r.stdout
