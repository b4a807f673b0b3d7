"""Golden answer of the full heightmap batch from the reference itself: the 5,000-segment
zigzag tool path (mcsim.workloads.zigzag_segments(5000), stamps every min(hx, 0.4) * 0.6 mm
along it, about 2.4 million flat stamps) on the 512 x 512 stock of 0.1 mm cells, applied by
the reference's Stock.stamp (src/stock.js:114-163) in order.

Runs ONLY in the build container, where the reference's source tree is mounted (Node loads
it through downlevel_loader.mjs as for the other fixtures).  The stamps are not stored: they
are regenerated from the seeded workload, so the fixture holds the answers only:

  golden_zigzag.json  {"workload", "nx", "ny", "n", "H", "err", "wallD", "lineCut",
                       "lineWall", "markS", "markF" (SHA-256 of each final map's bytes),
                       "removed", "node_ms" (the reference's time for the batch)}

usage: python tests/golden/make_golden_zigzag.py
"""
import json
import os
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "machinig-sim_amd"))
sys.path.insert(0, HERE)

from make_golden import run  # noqa: E402
from mcsim import path, workloads as W  # noqa: E402

STOCK = dict(w=51.1, l=51.1, h=20.0, cell=0.1)


def stamps():
    rec, _ = path.schedule(W.zigzag_segments(5000), STOCK["cell"])
    return rec


def main():
    rec = stamps()
    with tempfile.TemporaryDirectory() as work:
        sf = os.path.join(work, "stamps.bin")
        rec.tofile(sf)
        cin, cout = os.path.join(work, "cases.json"), os.path.join(work, "out.json")
        with open(cin, "w") as f:
            json.dump({"zigzag": [dict(STOCK, stampsFile=sf, n=len(rec))]}, f)
        run(cin, cout)
        res = json.load(open(cout))["zigzag"][0]
    res["workload"] = "zigzag 5000 segments, 512 x 512 stock, 0.1 mm cells"
    with open(os.path.join(HERE, "golden_zigzag.json"), "w") as f:
        json.dump(res, f, indent=1)
    print(res)


if __name__ == "__main__":
    main()
