"""Builds libexoplasim_mi355x.so in-tree: every HIP and C++ source under
csrc/ compiled for gfx950 with hipcc, then linked into one shared library
whose exported symbols are the C-ABI of include/exoplasim_mi355x.h.

When Node's headers are present (/usr/include/node/node_api.h), it also
builds the Node-API binding node/exoplasim_napi.node over that library.

Usage: python web-exoplasim_amd/build.py [--jobs N]
"""
import argparse
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libexoplasim_mi355x.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

# Packed FP32 arithmetic stays off in device code; contraction to FMA is
# kept (hipcc's default for HIP).
DEVICE_FLAGS = [
    f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]
COMMON = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-result",
          "-Wno-unused-value", "-Wno-unused-function"]


def sources():
    out = []
    for name in sorted(os.listdir(CSRC)):
        if name.endswith(".hip") or name.endswith(".cpp"):
            out.append(os.path.join(CSRC, name))
    return out


def compile_one(src):
    obj = os.path.join(OBJ, os.path.basename(src) + ".o")
    cmd = [HIPCC] + COMMON
    if src.endswith(".hip"):
        cmd += DEVICE_FLAGS + ["-x", "hip"]
    cmd += ["-c", src, "-o", obj]
    proc = subprocess.run(cmd, capture_output=True, text=True)
    err = "\n".join(l for l in proc.stderr.splitlines()
                    if "packed-fp32-ops" not in l and l.strip())
    if proc.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{err}")
    return obj


def build(jobs=8, verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    srcs = sources()
    with ThreadPoolExecutor(max_workers=max(1, jobs)) as pool:
        objs = list(pool.map(compile_one, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError("link failed:\n" + proc.stderr)
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB} from {len(srcs)} sources")
    addon = build_node_addon()
    if verbose and addon:
        print(f"built {addon}")
    return LIB


NODE_INCLUDE = "/usr/include/node"
ADDON_SRC = os.path.join(HERE, "node", "exoplasim_napi.cc")
ADDON = os.path.join(HERE, "node", "exoplasim_napi.node")


def build_node_addon():
    """The Node-API binding, linked against the library with an rpath to
    it; skipped (returns None) where Node's headers are absent."""
    if not os.path.exists(os.path.join(NODE_INCLUDE, "node_api.h")):
        return None
    tmp = ADDON + ".tmp"
    cmd = ["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-Wall",
           "-DNODE_GYP_MODULE_NAME=exoplasim_napi", f"-I{NODE_INCLUDE}", ADDON_SRC,
           "-o", tmp, f"-L{HERE}", "-lexoplasim_mi355x", "-Wl,-rpath,$ORIGIN/.."]
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError("Node addon build failed:\n" + proc.stderr)
    os.replace(tmp, ADDON)
    return ADDON


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=8)
    args = ap.parse_args()
    try:
        build(args.jobs, verbose=True)
    except RuntimeError as exc:
        print(exc, file=sys.stderr)
        sys.exit(1)
