"""Golden limits of the whole 256-speed x 64-tool lobe grid from the reference itself.

Runs ONLY in the build container, where the reference's source tree is mounted: the
reference's timeDomainLimit (src/stabtd.js:24-32, loaded by Node through
downlevel_loader.mjs as for the other fixtures) resolves every one of the grid's 16,384
cells with {aMin 0.05, aMax 10, iters 10}, split over parallel Node processes.  The cells
are not stored: they are regenerated from the seeded generator (mcsim.workloads.sld_grid),
so the fixture holds only the answers, in grid order:

  golden_tdl_grid.json  {"workload", "opts", "n", "alim", "lo", "hi" (base64 f64, hi = +inf
                         for open cells), "open" (base64 u8), "node", "procs", "node_s"}

node_s is the wall time of the Node processes (their own clocks), which also serves as the
reference's CPU rate on this grid at `procs` processes.

usage: python tests/golden/make_golden_grid.py [procs]
"""
import base64
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "machinig-sim_amd"))
sys.path.insert(0, HERE)

from make_golden import cell_json  # noqa: E402
from mcsim import workloads as W  # noqa: E402


def main():
    procs = int(sys.argv[1]) if len(sys.argv) > 1 else 8
    cells, opts = W.sld_grid()
    n = len(cells)
    # interleaved chunks balance the long low-speed cells over the processes
    chunks = [list(range(p, n, procs)) for p in range(procs)]
    res = [None] * n
    with tempfile.TemporaryDirectory() as work:
        jobs = []
        t0 = time.perf_counter()
        for p, idx in enumerate(chunks):
            cin, cout = os.path.join(work, "in%d.json" % p), os.path.join(work, "out%d.json" % p)
            with open(cin, "w") as f:
                json.dump({"tdl": [dict(cell=cell_json(cells[i]), opts=dict(opts)) for i in idx]}, f)
            cmd = ["node", "--experimental-loader", os.path.join(HERE, "downlevel_loader.mjs"),
                   os.path.join(HERE, "run_reference.mjs"), cin, cout]
            jobs.append((subprocess.Popen(cmd, cwd=HERE, stderr=subprocess.PIPE, text=True), cout, idx))
        for pr, cout, idx in jobs:
            _, err = pr.communicate()
            if pr.returncode != 0:
                raise SystemExit("node failed:\n" + err)
            for i, r in zip(idx, json.load(open(cout))["tdl"]):
                res[i] = r
        wall = time.perf_counter() - t0
    alim = np.array([r["alim"] for r in res], dtype=np.float64)
    lo = np.array([r["lo"] for r in res], dtype=np.float64)
    hi = np.array([np.inf if r["hi"] == "inf" else r["hi"] for r in res], dtype=np.float64)
    op = np.isinf(hi).astype(np.uint8)
    b64 = lambda a: base64.b64encode(a.tobytes()).decode()
    rec = dict(workload="sld_grid_256speeds_x_64tools", opts=opts, n=n, alim=b64(alim), lo=b64(lo), hi=b64(hi),
               open=b64(op), procs=procs, node_s=wall,
               node=subprocess.run(["node", "--version"], capture_output=True, text=True).stdout.strip())
    with open(os.path.join(HERE, "golden_tdl_grid.json"), "w") as f:
        json.dump(rec, f, separators=(",", ":"))
    print(json.dumps(dict(cells=n, open=int(op.sum()), procs=procs, node_s=wall, cells_per_s=n / wall)))


if __name__ == "__main__":
    main()
