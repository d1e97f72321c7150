"""Builds libgeon_hip.so (the gfx950 backend) in-tree with hipcc.

Each .hip file is compiled on its own (in parallel) and linked into one shared
library next to this file.  Every kernel is compiled for gfx950 only, with the
packed-FP32 instructions disabled.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libgeon_hip.so")
SOURCES = ["geon_api.hip", "geon_batch.hip", "k_tree.hip", "k_force.hip", "k_step.hip", "k_field.hip", "k_learn.hip",
           "k_delay.hip", "k_merge.hip", "k_rad.hip", "k_ghost.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-Wno-unused-value",
]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError("command failed: %s\n%s" % (" ".join(cmd), out))
    return out


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    cmds = []
    for src in SOURCES:
        obj = os.path.join(BUILD, src.replace(".hip", ".o"))
        objs.append(obj)
        cmds.append([HIPCC] + FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj])
    with ThreadPoolExecutor(max_workers=min(8, len(cmds))) as ex:
        for out in ex.map(_run, cmds):
            if verbose and out.strip():
                print(out)
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    build_napi()
    return LIB


NAPI_DIR = os.path.join(HERE, "napi")
NAPI_ADDON = os.path.join(NAPI_DIR, "geon_napi.node")
NODE_INCLUDE = "/usr/include/node"


def build_napi():
    """The Node-API addon over the C ABI (needs only the node headers)."""
    if not os.path.exists(os.path.join(NODE_INCLUDE, "node_api.h")):
        return None
    tmp = NAPI_ADDON + ".tmp"
    _run(["gcc", "-O2", "-shared", "-fPIC", "-I" + NODE_INCLUDE, "-DNODE_GYP_MODULE_NAME=geon_napi",
          os.path.join(NAPI_DIR, "geon_napi.c"), "-L" + HERE, "-lgeon_hip", "-Wl,-rpath,$ORIGIN/..", "-o", tmp])
    os.replace(tmp, NAPI_ADDON)
    return NAPI_ADDON


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
