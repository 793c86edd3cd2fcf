"""Add Platform::GPU_HIP to a scratch copy of the OpenLB 1.8 source tree.

    python integration/openlb/apply.py --reference /root/reference --dest /tmp/olb_gpu_hip

Copies the reference tree to `dest` (the reference itself is never written),
then
  * installs this directory's gpu_hip/ headers as src/core/platform/gpu/hip.h
    and src/core/platform/gpu/hip/*.h, and
  * makes the edits an OpenLB maintainer would make to plug a platform in,
    each next to the CUDA platform's counterpart:
      platform.h        the enum value and isPlatformCPU
      dispatch.h        a GPU_HIP case in every platform switch
      blockLattice.h    the forward declarations
      blockLattice.hh   ConcreteBlockLattice::stream -> gpu::hip::stream
      superLattice.hh   device synchronisation before post-processing stages
      loadBalancer.h    GPU_HIP as the default platform when compiled in
      olb.h             the platform's definitions after every other header
Every edit is anchored on a short line of the reference file and refuses to
run if its anchor is missing or ambiguous, so a changed reference fails loudly.
On a destination that already has the platform only the headers are installed
again (--fresh starts from a new copy).
"""
from __future__ import annotations

import argparse
import os
import re
import shutil
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
HEADERS = os.path.join(HERE, "gpu_hip")


class AnchorError(RuntimeError):
    pass


def _read(root, rel):
    with open(os.path.join(root, rel)) as f:
        return f.read()


def _write(root, rel, text):
    with open(os.path.join(root, rel), "w") as f:
        f.write(text)


def insert_after(text, anchor, addition, rel):
    if text.count(anchor) != 1:
        raise AnchorError(f"{rel}: anchor {anchor!r} found {text.count(anchor)} times")
    i = text.index(anchor) + len(anchor)
    return text[:i] + addition + text[i:]


def insert_before(text, anchor, addition, rel):
    if text.count(anchor) != 1:
        raise AnchorError(f"{rel}: anchor {anchor!r} found {text.count(anchor)} times")
    i = text.index(anchor)
    return text[:i] + addition + text[i:]


def duplicate_cuda_cases(text, rel, expected):
    """Every `#ifdef PLATFORM_GPU_CUDA ... #endif` case block of a platform
    switch gets a GPU_HIP twin right after it."""
    pattern = re.compile(r"#ifdef PLATFORM_GPU_CUDA\n(?:(?!#endif).)*?case Platform::GPU_CUDA:(?:(?!#endif).)*?#endif\n",
                         re.S)
    blocks = pattern.findall(text)
    if len(blocks) != expected:
        raise AnchorError(f"{rel}: expected {expected} CUDA case blocks, found {len(blocks)}")
    return pattern.sub(lambda m: m.group(0) + m.group(0).replace("GPU_CUDA", "GPU_HIP"), text)


def install_headers(dest):
    gpu = os.path.join(dest, "src", "core", "platform", "gpu")
    os.makedirs(os.path.join(gpu, "hip"), exist_ok=True)
    for name in sorted(os.listdir(HEADERS)):
        if not name.endswith(".h"):
            continue
        target = os.path.join(gpu, "hip.h") if name == "hip.h" else os.path.join(gpu, "hip", name)
        shutil.copyfile(os.path.join(HEADERS, name), target)


def apply(reference, dest, fresh=False):
    if not os.path.isdir(os.path.join(reference, "src", "core", "platform")):
        raise SystemExit(f"{reference} is not an OpenLB source tree")
    if not fresh and os.path.exists(os.path.join(dest, "src", "core", "platform", "gpu", "hip.h")):
        install_headers(dest)
        return dest
    if os.path.exists(dest):
        shutil.rmtree(dest)
    shutil.copytree(reference, dest, ignore=shutil.ignore_patterns(".git", "*.o", "*.a", "build"))
    install_headers(dest)

    rel = "src/core/platform/platform.h"
    t = _read(dest, rel)
    t = insert_after(t, "  GPU_CUDA, /// GPU using CUDA\n", "  GPU_HIP,  /// GPU using HIP through liblbmh (AMD Instinct)\n", rel)
    t = insert_before(t, "  default:\n    throw std::invalid_argument(\"Invalid PLATFORM\");\n  }\n}\n\n}\n",
                      "#ifdef PLATFORM_GPU_HIP\n  case Platform::GPU_HIP:\n    return false;\n#endif\n", rel)
    _write(dest, rel, t)

    rel = "src/core/platform/dispatch.h"
    _write(dest, rel, duplicate_cuda_cases(_read(dest, rel), rel, expected=3))

    rel = "src/core/blockLattice.h"
    t = _read(dest, rel)
    t = insert_after(t, "#ifdef PLATFORM_GPU_CUDA\n#include \"platform/gpu/cuda.h\"\n#endif\n",
                     "\n#ifdef PLATFORM_GPU_HIP\n#include \"platform/gpu/hip/forward.h\"\n#endif\n", rel)
    _write(dest, rel, t)

    rel = "src/core/blockLattice.hh"
    t = _read(dest, rel)
    t = insert_after(t, "void ConcreteBlockLattice<T,DESCRIPTOR,PLATFORM>::stream()\n{\n",
                     "#ifdef PLATFORM_GPU_HIP\n  if constexpr (PLATFORM == Platform::GPU_HIP) {\n"
                     "    gpu::hip::stream(*this);\n    return;\n  }\n#endif\n", rel)
    _write(dest, rel, t)

    rel = "src/core/superLattice.hh"
    t = _read(dest, rel)
    t = insert_after(t, "void SuperLattice<T,DESCRIPTOR>::executePostProcessors(STAGE stage)\n{\n",
                     "#ifdef PLATFORM_GPU_HIP\n  gpu::hip::synchronize();\n#endif\n", rel)
    _write(dest, rel, t)

    rel = "src/communication/loadBalancer.h"
    t = _read(dest, rel)
    t = insert_before(t, "    #ifdef PLATFORM_GPU_CUDA\n      return Platform::GPU_CUDA;\n",
                      "    #ifdef PLATFORM_GPU_HIP\n      return Platform::GPU_HIP;\n    #endif\n", rel)
    _write(dest, rel, t)

    rel = "src/olb.h"
    t = _read(dest, rel)
    t = insert_after(t, "#include \"cse/cse.h\"\n",
                     "\n#ifdef PLATFORM_GPU_HIP\n#include \"core/platform/gpu/hip.h\"\n#endif\n", rel)
    _write(dest, rel, t)
    return dest


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--reference", default="/root/reference")
    ap.add_argument("--dest", default="/tmp/olb_gpu_hip")
    ap.add_argument("--fresh", action="store_true", help="start from a new copy of the reference")
    args = ap.parse_args()
    try:
        print(apply(args.reference, args.dest, args.fresh))
    except AnchorError as e:
        sys.exit(f"apply.py: {e}")


if __name__ == "__main__":
    main()
