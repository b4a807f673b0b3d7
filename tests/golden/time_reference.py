"""Time the reference's own timeDomainLimit (src/stabtd.js:24-32, run by Node from its source
tree through run_reference.mjs, as the fixtures are made) against the oracle's restatement
(oracle/stabtd.c) on the same cells of the 40-speed diagram, one thread each.  The ratio
relates the bench's CPU baseline (the oracle on host threads) to the reference's Node path.

Runs ONLY in the build container, where the reference's source tree is mounted.

usage: python tests/golden/time_reference.py [every]   (default: every 4th of the 40 cells)
       python tests/golden/time_reference.py sim   (the simulator's sub-step loop, per sub-step)
"""
import json
import os
import re
import subprocess
import sys
import tempfile
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "machinig-sim_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, HERE)

from make_golden import cell_json  # noqa: E402
from mcsim import workloads as W  # noqa: E402


def sim_substeps(reps=4):
    """the reference Simulator's sub-step loop (sim.step on golden_simx.json's first program,
    run reps times in one Node process; the first run warms the JIT up) against the oracle's
    restated loop (oracle/simloop.c) on the same sub-steps: microseconds per sub-step"""
    sys.path.insert(0, os.path.dirname(HERE))
    import golden_util as G
    import oracle
    g = G.load("simx")
    case, res = g["cases"][0], g["results"][0]
    n_sub = len(res["actual"]["type"])
    with tempfile.TemporaryDirectory() as work:
        cin, cout = os.path.join(work, "time_cases.json"), os.path.join(work, "time_out.json")
        with open(cin, "w") as f:
            json.dump({"simx": [case] * reps}, f)
        cmd = ["node", "--experimental-loader", os.path.join(HERE, "downlevel_loader.mjs"),
               os.path.join(HERE, "run_reference.mjs"), cin, cout]
        r = subprocess.run(cmd, cwd=HERE, capture_output=True, text=True, check=True)
        per = [float(x) for x in re.findall(r"^simx case \d+ ([0-9.]+) s$", r.stderr, re.M)]
    sub, n = G.simx_program(res)
    cell, fnMax = G.simx_cell(res)
    typ, lut, rc = G.simx_tool(case, res, oracle.tip_lut)
    st = case["stock"]
    ts = []
    for _ in range(reps):
        a, i = oracle.OracleStock(st["w"], st["l"], st["h"], case["cell"]), oracle.OracleStock(st["w"], st["l"], st["h"], case["cell"])
        i.set_wall_band(2.0)
        e = oracle.OracleEngine()
        info = e.configure(cell, fnMax)
        e.set_wear(res["wears"][0][1], res["wears"][0][2])
        t0 = time.perf_counter()
        oracle.sim_run(a, i, e, sub, n, res["toolR"], res["fluteLen"], 0.0, info.stepsPerTooth / 4, typ, lut=lut, cornerR=rc)
        ts.append(time.perf_counter() - t0)
    print(json.dumps(dict(program=0, substeps=n_sub, reference_us_per_substep=[1e6 * t / n_sub for t in per],
                          oracle_us_per_substep=[1e6 * t / n for t in ts])))


def main():
    if len(sys.argv) > 1 and sys.argv[1] == "sim":
        sim_substeps()
        return
    every = int(sys.argv[1]) if len(sys.argv) > 1 else 4
    cells, opts = W.sld_40rpm()
    cells = cells[::every]
    with tempfile.TemporaryDirectory() as work:
        cin, cout = os.path.join(work, "time_cases.json"), os.path.join(work, "time_out.json")
        with open(cin, "w") as f:
            json.dump({"tdl": [dict(cell=cell_json(c), opts=dict(opts)) for c in cells]}, f)
        cmd = ["node", "--experimental-loader", os.path.join(HERE, "downlevel_loader.mjs"),
               os.path.join(HERE, "run_reference.mjs"), cin, cout]
        r = subprocess.run(cmd, cwd=HERE, capture_output=True, text=True, check=True)
        m = re.search(r"^tdl \d+ ([0-9.]+) s$", r.stderr, re.M)
        t_ref = float(m.group(1))
        ref = json.load(open(cout))["tdl"]
    import oracle
    t0 = time.perf_counter()
    got = [oracle.time_domain_limit(c, **opts) for c in cells]
    t_or = time.perf_counter() - t0
    same = all(g.alim == r_["alim"] for g, r_ in zip(got, ref))
    print(json.dumps(dict(cells=len(cells), reference_node_s=t_ref, oracle_s=t_or, ratio=t_ref / t_or,
                          same_alim=same, node=subprocess.run(["node", "--version"], capture_output=True,
                                                              text=True).stdout.strip())))


if __name__ == "__main__":
    main()
