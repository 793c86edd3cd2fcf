"""Time OpenLB's own CPU path (CPU_SIMD + OpenMP, oracle/_ref/olb_ref_simd_f32)
beside the oracle (oracle/liboracle.so) on the same cores and the same case.

TEST INFRASTRUCTURE ONLY. It shows how closely the oracle, timed as bench.py's
CPU baseline, stands in for the reference's CPU speed. Both programs run 10
warm-up steps, then the timed steps, and count MLUPS over cells with dynamics
(the reference harness's definition, src/utilities/timer.hh:71-76).

    python oracle/compare_cpu.py [--size 128] [--steps 20] [--threads 8]
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
# import the oracle package from the repository root, not oracle/oracle.py as a top-level module
sys.path[:] = [p for p in sys.path if os.path.abspath(p or ".") != HERE]
sys.path.insert(0, ROOT)

REF_SIMD = os.path.join(HERE, "_ref", "olb_ref_simd_f32")
REF_SIMD_AVX512 = os.path.join(HERE, "_ref", "olb_ref_simd_f32_avx512")
AVX512_FLAGS = ("avx512f", "avx512vl", "avx512bw", "avx512dq")


def cpu_flags():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("flags"):
                    return set(line.split(":", 1)[1].split())
    except OSError:
        pass
    return set()


def reference_binary():
    """(path, ISA) of the OpenLB CPU_SIMD build for this host: the AVX-512 one when
    the CPU reports every AVX-512 subset it was compiled for, else the AVX2 one."""
    if os.path.exists(REF_SIMD_AVX512) and all(f in cpu_flags() for f in AVX512_FLAGS):
        return REF_SIMD_AVX512, "AVX-512"
    return REF_SIMD, "AVX2"


def default_threads():
    """The host share this process may use: OMP_NUM_THREADS when set (the GPU
    machines set it to the cores allotted per GPU), else every core in the affinity mask."""
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        return int(env)
    return len(os.sched_getaffinity(0))


def host_info(threads):
    """What a CPU figure needs beside it: the core model, nproc, the affinity mask
    and the OpenMP settings the timed run used."""
    model = ""
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"cpu_model": model, "nproc": os.cpu_count(), "affinity_cores": len(os.sched_getaffinity(0)),
            "omp_num_threads": threads, "omp_places": "cores", "omp_proc_bind": "close",
            "isa": reference_binary()[1] if os.path.exists(REF_SIMD) else None}


def write_case(d, case, steps):
    import numpy as np
    case.ids.astype(np.uint8).tofile(os.path.join(d, "ids.bin"))
    np.ascontiguousarray(case.rho_u, dtype=np.float64).tofile(os.path.join(d, "rhou.bin"))
    lines = [f"q {case.q}", f"dyn {case.kind}"]
    lines += [f"n{a} {n}" for a, n in zip("xyz", case.shape)]
    lines += [f"p{a} {p}" for a, p in zip("xyz", case.periodic)]
    lines += [f"omega {case.omega!r}", f"magic {case.magic!r}", f"smago {case.smago!r}"]
    lines += [f"f{a} {v!r}" for a, v in zip("xyz", case.force)]
    lines.append(f"steps {steps}")
    for wid, (rho, u) in case.walls.items():
        lines.append(f"wall {wid} {rho!r} {u[0]!r} {u[1]!r} {u[2]!r}")
    for mid, kind in (getattr(case, "local", {}) or {}).items():
        lines.append(f"local {mid} {kind}")
    with open(os.path.join(d, "case.txt"), "w") as fh:
        fh.write("\n".join(lines) + "\n")


def time_reference(case, steps, threads):
    """OpenLB CPU_SIMD+OpenMP on the case; None if the binary was not built."""
    if not os.path.exists(REF_SIMD):
        return None
    binary, isa = reference_binary()
    env = dict(os.environ, OLB_REF_TIMING="1", OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="close",
               OMP_PLACES="cores")
    with tempfile.TemporaryDirectory() as d:
        write_case(d, case, steps)
        out = subprocess.run([binary, d], cwd=d, env=env, capture_output=True, text=True, check=True)
    line = [ln for ln in out.stdout.splitlines() if ln.startswith("{")][-1]
    res = json.loads(line)
    res["isa"] = isa
    return res


def time_oracle(case, steps, threads, storage="f32"):
    from oracle import oracle as O
    lat = O.OracleLattice(case, storage)
    lat.upload(O.initial_populations(case, storage))
    lat.step(10, threads)
    t0 = time.perf_counter()
    lat.step(steps, threads)
    s = time.perf_counter() - t0
    cells = case.n_active
    return {"cells": cells, "steps": steps, "seconds": s, "mlups": cells * steps / s / 1e6}


def main():
    from openlb_amd import workloads as W
    ap = argparse.ArgumentParser()
    ap.add_argument("--size", type=int, default=128)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--threads", type=int, default=default_threads())
    args = ap.parse_args()
    n = args.size
    for case in (W.porous_channel((n, n, n), tau=0.55, force=1e-6), W.taylor_green((n, n, n), u0=0.05, tau=0.6)):
        ref = time_reference(case, args.steps, args.threads)
        ora = time_oracle(case, args.steps, args.threads)
        row = {"case": f"{case.name} {n}^3 {case.kind} f32", **host_info(args.threads),
               "openlb_cpu_simd_mlups": ref and round(ref["mlups"], 2), "oracle_mlups": round(ora["mlups"], 2),
               "ratio_oracle_over_openlb": ref and round(ora["mlups"] / ref["mlups"], 3)}
        print(json.dumps(row), flush=True)


if __name__ == "__main__":
    main()
