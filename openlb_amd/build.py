"""Build the HIP backend library openlb_amd/_lib/liblbmh.so for gfx950.

    python -m openlb_amd.build

Compiles every HIP source with hipcc for gfx950 only. It always compiles,
even when a built library exists. The library is built in-tree so that it
travels with the source tree to the GPU machine.
"""
from __future__ import annotations

import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
LIB_DIR = os.path.join(PKG, "_lib")
LIB_PATH = os.path.join(LIB_DIR, "liblbmh.so")
SOURCES = [os.path.join(PKG, "csrc", "lbmh.hip"), os.path.join(PKG, "csrc", "lbmh_ensemble.hip")]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    # packed FP32 instructions are kept out of every kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
    f"-I{os.path.join(ROOT, 'include')}",
    f"-I{os.path.join(PKG, 'csrc')}",
]


HOST_DIR = os.path.join(PKG, "host")
EXAMPLES = {"cavity3d_benchmark": os.path.join(HOST_DIR, "examples", "cavity3d_benchmark.cpp"),
            "uq_cavity3d": os.path.join(HOST_DIR, "examples", "uq_cavity3d.cpp")}


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        sys.stderr.write(proc.stdout + proc.stderr)
        raise RuntimeError(f"{cmd[0]} failed with exit code {proc.returncode}")


def build_examples(verbose: bool = False):
    """C++ programs on the host layer (openlb_amd/host/olb_hip.h), linked to liblbmh.so."""
    for name, src in EXAMPLES.items():
        out = os.path.join(LIB_DIR, name)
        _run(["g++", "-O2", "-std=c++17", "-Wall", f"-I{os.path.join(ROOT, 'include')}", f"-I{HOST_DIR}",
              src, "-o", out + ".tmp", f"-L{LIB_DIR}", "-llbmh", "-Wl,-rpath,$ORIGIN",
              "-L/opt/rocm/lib", "-Wl,-rpath-link,/opt/rocm/lib", "-lz"], verbose)
        os.replace(out + ".tmp", out)


def build(verbose: bool = False) -> str:
    os.makedirs(LIB_DIR, exist_ok=True)
    tmp = LIB_PATH + ".tmp"
    _run([HIPCC, *FLAGS, *SOURCES, "-o", tmp], verbose)
    os.replace(tmp, LIB_PATH)
    build_examples(verbose)
    return LIB_PATH


if __name__ == "__main__":
    print(build(verbose=True))
