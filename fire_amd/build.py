"""Builds libfirex_hip.so (gfx950 kernels + C ABI) in-tree with hipcc.

Usage: python -m fire_amd.build   (or fire_amd.build.build_library()).

All device and host code is compiled with -ffp-contract=off so that no fused
multiply-add changes the f32 results, and with packed-FP32 instructions
disabled (-Xclang -target-feature -Xclang -packed-fp32-ops).
"""
from __future__ import annotations

import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIBRARY = os.path.join(HERE, "libfirex_hip.so")
ARCH = os.environ.get("FIREX_OFFLOAD_ARCH", "gfx950")

SOURCES = ["firex_kernels.hip", "firex_jacobi.hip", "firex_droplets.hip", "firex_audit.hip", "firex_render.hip",
           "firex_solver.hip", "firex_server.hip", "firex_host.cpp"]

COMMON = [
    "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math",
    "-Wall", "-Wno-unused-function", "-Wno-unused-command-line-argument",
]
DEVICE = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def _hipcc() -> str:
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    path = os.path.join(rocm, "bin", "hipcc")
    return path if os.path.exists(path) else "hipcc"


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [_hipcc()] + COMMON
    if src.endswith(".hip"):
        cmd += DEVICE + ["-x", "hip"]
    cmd += ["-c", os.path.join(CSRC, src), "-o", obj]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{res.stdout}\n{res.stderr}")
    return obj


def build_library(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    jobs = min(len(SOURCES), max(1, os.cpu_count() or 1))
    with concurrent.futures.ThreadPoolExecutor(max_workers=jobs) as pool:
        objects = list(pool.map(_compile, SOURCES))
    tmp = LIBRARY + ".tmp"
    cmd = [_hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objects
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, LIBRARY)
    if verbose:
        print(f"built {LIBRARY}")
    build_napi_addon(verbose)
    return LIBRARY


NAPI_SOURCE = os.path.join(HERE, "napi", "firex_napi.c")
NAPI_ADDON = os.path.join(HERE, "firex_napi.node")
NODE_INCLUDE = os.environ.get("NODE_INCLUDE", "/usr/include/node")


def build_napi_addon(verbose: bool = False) -> str:
    """Node-API addon over the C ABI (plain C, no node-gyp): firex_napi.node.

    Skipped (returns "") when the Node-API headers are not installed."""
    if not os.path.exists(os.path.join(NODE_INCLUDE, "node_api.h")):
        if verbose:
            print("node_api.h not found: N-API addon not built")
        return ""
    tmp = NAPI_ADDON + ".tmp"
    cmd = ["gcc", "-O2", "-std=c11", "-shared", "-fPIC", "-Wall", f"-I{NODE_INCLUDE}",
           "-DNODE_GYP_MODULE_NAME=firex_napi", NAPI_SOURCE, "-L" + HERE, "-lfirex_hip",
           "-Wl,-rpath,$ORIGIN", "-o", tmp]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"N-API addon build failed:\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, NAPI_ADDON)
    if verbose:
        print(f"built {NAPI_ADDON}")
    return NAPI_ADDON


if __name__ == "__main__":
    build_library(verbose=True)
    sys.exit(0)
