"""Build the N-API addon hackathon_amd/napi/ddm_napi.node (gcc, node_api.h).

It links libddm.so from hackathon_amd/ with an $ORIGIN/.. rpath, so the pair
travels together.  Skipped (returns None) when the Node headers are absent.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
OUT = os.path.join(HERE, "ddm_napi.node")
NODE_INC = "/usr/include/node"


def build(verbose: bool = False):
    if not os.path.exists(os.path.join(NODE_INC, "node_api.h")):
        if verbose:
            print("node_api.h not found; N-API addon not built")
        return None
    cmd = ["gcc", "-O2", "-fPIC", "-shared", "-std=c11", "-Wall", "-I" + NODE_INC,
           "-DNODE_GYP_MODULE_NAME=ddm_napi", "-pthread", "-o", OUT + ".tmp", os.path.join(HERE, "ddm_napi.c"),
           "-L" + PKG, "-l:libddm.so", "-Wl,-rpath,$ORIGIN/.."]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout)
        raise RuntimeError("N-API addon build failed")
    os.replace(OUT + ".tmp", OUT)
    if verbose:
        print("built", OUT)
    return OUT


if __name__ == "__main__":
    build(verbose=True)
