"""Helpers to read the golden fixtures (tests/golden/*.json) written from the reference."""
import base64
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, os.path.join(ROOT, "machinig-sim_amd"))

from mcsim._abi import make_cell, Stamp, MC_TOOL_BALL, MC_TOOL_FLAT  # noqa: E402


def load(name):
    with open(os.path.join(HERE, "golden", "golden_%s.json" % name)) as f:
        return json.load(f)


def f64(s):
    return np.frombuffer(base64.b64decode(s), dtype="<f8")


def f32(s):
    return np.frombuffer(base64.b64decode(s), dtype="<f4")


def i32(s):
    return np.frombuffer(base64.b64decode(s), dtype="<i4")


def num(v):
    if v == "inf":
        return float("inf")
    if v == "-inf":
        return float("-inf")
    return float(v)


def cell_from_json(c):
    md = c["modal"]
    mx = [(m["k"], m["fn"], m["zeta"]) for m in (md.get("modesX") or [md])]
    my = [(m["k"], m["fn"], m["zeta"]) for m in (md.get("modesY") or [])] or None
    return make_cell(c["tool"], mx, my, c["mat"], c["S"], c["arc"], c["fz"], c["revs"],
                     c["nonlinear"])


def stamps_from_json(stamps):
    arr = (Stamp * len(stamps))()
    for i, s in enumerate(stamps):
        arr[i] = Stamp(s["cx"], s["cy"], s["tipZ"], s["R"],
                       MC_TOOL_BALL if s["type"] == "ball" else MC_TOOL_FLAT, s["line"], s["errVal"])
    return arr
