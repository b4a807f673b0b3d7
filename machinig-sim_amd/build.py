"""Build libmcsim.so (HIP kernels for gfx950 + the C ABI) in-tree with hipcc.

usage: python machinig-sim_amd/build.py
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libmcsim.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    # JavaScript never fuses multiply-add: bit parity with the reference needs no contraction
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]
SOURCES = ["chatter.hip", "search.hip", "stock.hip", "engine.hip", "simrun.hip", "mlp.hip", "lobes.hip", "tiplut.cpp", "jsmath.cpp", "path.cpp", "api.cpp"]


def build(verbose=False, jobs=None):
    """compile every source (in parallel, they are independent) and link libmcsim.so"""
    from concurrent.futures import ThreadPoolExecutor
    objs, cmds = [], []
    for src in SOURCES:
        obj = os.path.join(CSRC, os.path.splitext(src)[0] + ".o")
        cmds.append([HIPCC] + HIP_FLAGS + ["-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    jobs = jobs or min(len(cmds), max(1, min(8, os.cpu_count() or 1)))
    with ThreadPoolExecutor(jobs) as ex:
        for cmd, r in zip(cmds, ex.map(lambda c: subprocess.run(c, capture_output=not verbose), cmds)):
            if verbose:
                print(" ".join(cmd))
            if r.returncode != 0:
                sys.stderr.write((r.stderr or b"").decode(errors="replace"))
                raise subprocess.CalledProcessError(r.returncode, cmd)
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", LIB] + objs
    subprocess.run(cmd, check=True)
    for o in objs:
        os.remove(o)
    build_addon(verbose)
    return LIB


NODE_INCLUDE = "/usr/include/node"
ADDON = os.path.join(HERE, "mcsim.node")


def build_addon(verbose=False):
    """N-API addon binding libmcsim.so for the JS host modules (skipped without Node headers)."""
    if not os.path.exists(os.path.join(NODE_INCLUDE, "node_api.h")):
        print("node headers not found: N-API addon not built", file=sys.stderr)
        return None
    cmd = ["g++", "-shared", "-fPIC", "-O2", "-std=c++17", "-DNODE_GYP_MODULE_NAME=mcsim",
           "-I", NODE_INCLUDE, os.path.join(CSRC, "napi_addon.cc"), "-o", ADDON,
           "-L", HERE, "-lmcsim", "-Wl,-rpath,$ORIGIN"]
    if verbose:
        print(" ".join(cmd))
    subprocess.run(cmd, check=True)
    return ADDON


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
