"""Throughput of the block-lattice time step on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload porous_channel]
                    [--shape 256 256 256] [--storage f32] [--cpu-seconds 15]

The default workload is the D3Q19 porous channel at 256^3 in FP32. It is periodic in
x, with BounceBack walls at the y/z bounds and seeded BounceBack spheres (solid
fraction 0.1, radius 3-6). The bulk cells run Guo-forced BGK (F = 1e-6 along x,
tau = 0.55). One step is one ``collideAndStream()`` of the whole block: one launch of
the fused collide-stream kernel. The inputs are resident in HBM before the timed
region starts.

The output is one JSON line on stdout. ``value`` is MLUPS, i.e. updated cells
(cells with dynamics, OpenLB's getNvoxel()) times steps per second, over the timed
steps. That region is bracketed by device synchronisations. ``roofline`` prices
the kernel's algorithmic traffic, 2*Q*sizeof(stored population) bytes per cell
update, against its average per-launch duration measured with HIP events on the
launch stream. ``cpu_baseline`` is OpenLB's own CPU path (CPU_SIMD + OpenMP, built
from the reference sources into oracle/_ref, kind "reference") timed on the host
cores on a bounded sample of the same workload; where that binary is absent, the
oracle (the CPU restatement of OpenLB's time step, kind "port") is timed instead.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, vendor peak
STORE_BYTES = {"f64": 8, "f32": 4, "f16": 2}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


class HipEvents:
    """hipEvent timing on a given stream (libamdhip64 through ctypes)."""

    def __init__(self):
        self.hip = ctypes.CDLL("libamdhip64.so")
        self.hip.hipEventCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
        self.hip.hipEventRecord.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        self.hip.hipEventSynchronize.argtypes = [ctypes.c_void_p]
        self.hip.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), ctypes.c_void_p,
                                                 ctypes.c_void_p]
        self.hip.hipEventDestroy.argtypes = [ctypes.c_void_p]

    def create(self, n):
        evs = []
        for _ in range(n):
            e = ctypes.c_void_p()
            if self.hip.hipEventCreate(ctypes.byref(e)) != 0:
                raise RuntimeError("hipEventCreate failed")
            evs.append(e)
        return evs

    def record(self, ev, stream=None):
        if self.hip.hipEventRecord(ev, ctypes.c_void_p(stream or 0)) != 0:
            raise RuntimeError("hipEventRecord failed")

    def elapsed_ms(self, a, b):
        self.hip.hipEventSynchronize(b)
        ms = ctypes.c_float(0)
        if self.hip.hipEventElapsedTime(ctypes.byref(ms), a, b) != 0:
            raise RuntimeError("hipEventElapsedTime failed")
        return float(ms.value)

    def destroy(self, evs):
        for e in evs:
            self.hip.hipEventDestroy(e)


def pattern_ceiling(lat, ev, repeats=20):
    """Average launch time (ms) of lbmh_pattern_ceiling on this lattice: the step
    kernel's memory access pattern without its arithmetic, on the null stream the
    step kernel runs on. It leaves the lattice freely streamed."""
    from openlb_amd import abi
    lib = abi.load()
    abi.check(lib.lbmh_pattern_ceiling(lat.handle, 2, None))      # warm-up
    marks = ev.create(2 * repeats)
    for k in range(repeats):
        ev.record(marks[2 * k])
        abi.check(lib.lbmh_pattern_ceiling(lat.handle, 1, None))
        ev.record(marks[2 * k + 1])
    lat.synchronize()
    ms = [ev.elapsed_ms(marks[2 * k], marks[2 * k + 1]) for k in range(repeats)]
    ev.destroy(marks)
    return float(np.mean(ms))


def calibrate_copy(mib: int, repeats: int = 5):
    """Device-to-device copies of a known size: the reference point that scales the
    FETCH_SIZE / WRITE_SIZE counters of the step kernel to HBM bytes."""
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipMemset.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    nbytes = mib * 1024 * 1024
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    if hip.hipMalloc(ctypes.byref(a), nbytes) or hip.hipMalloc(ctypes.byref(b), nbytes):
        raise SystemExit("hipMalloc failed")
    hip.hipMemset(a, 1, nbytes)
    hip.hipDeviceSynchronize()
    t0 = time.perf_counter()
    for _ in range(repeats):
        if hip.hipMemcpy(b, a, nbytes, 3) != 0:  # hipMemcpyDeviceToDevice
            raise SystemExit("hipMemcpy failed")
    hip.hipDeviceSynchronize()
    dt = (time.perf_counter() - t0) / repeats
    hip.hipFree(a)
    hip.hipFree(b)
    print(json.dumps({"calibration": "d2d copy", "bytes_read": nbytes, "bytes_written": nbytes,
                      "repeats": repeats, "copy_gbs": round(2 * nbytes / dt / 1e9, 1)}), flush=True)


def kernel_profile(steps: int = 10):
    """A short fixed run for hardware-counter passes (rocprofv3 --pmc): the
    kernels whose counters are tabulated, each launched `steps` times, plus
    two torch kernels of known traffic that calibrate the byte counters (a
    read-only float32 sum of 1 GiB, and a float32 elementwise negation of
    1 GiB into another 1 GiB).  It prints one JSON line naming each kernel's
    algorithmic bytes per launch."""
    import torch

    from openlb_amd import workloads as W
    from openlb_amd.ensemble import Ensemble
    from openlb_amd.lattice import BlockLattice
    out = {}
    for label, case, storage in (("porous_channel 256^3", W.porous_channel((256, 256, 256)), "f32"),
                                 ("turbulent_channel 384x192x192", W.turbulent_channel((384, 192, 192)), "f16")):
        with BlockLattice.from_case(case, storage) as lat:
            lat.ini_equilibrium(case.rho_u)
            lat.collide_and_stream(steps)
            lat.synchronize()
        out[f"collide_stream_kernel<{case.q},{storage},{case.kind}> {label}"] = {
            "launches": steps, "cells": case.n_active,
            "algorithmic_bytes": 2 * case.q * STORE_BYTES[storage] * case.n_active}
    cases = W.cavity_ensemble(128)
    pop = sum(Ensemble.bytes_for(19, "f16", c.shape)[0] for c in cases)
    ids = sum(Ensemble.bytes_for(19, "f16", c.shape)[1] for c in cases)
    with Ensemble(19, "f16", "bgk", max_members=len(cases), population_bytes=pop + (1 << 20),
                  cell_capacity=ids + (1 << 20)) as ens:
        for c in cases:
            m = ens.add(c)
            ens.ini_equilibrium(m, c.rho_u)
        ens.step(steps)
        ens.status(m)
    cells = sum(int((c.ids != 0).sum()) for c in cases)
    out["ensemble_step_kernel<19,f16,bgk> cavity_ensemble x128"] = {
        "launches": steps, "cells": cells, "algorithmic_bytes": 2 * 19 * 2 * cells}
    n = 1 << 28
    x = torch.ones(n, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    total = float(x.sum().item())
    torch.neg(x, out=y)
    torch.cuda.synchronize()
    if total != float(n):
        raise SystemExit("calibration sum is wrong")
    out["torch sum (reduce_kernel) 1 GiB float32"] = {"launches": 1, "bytes_read": 4 * n, "bytes_written": 0}
    out["torch neg (elementwise) 1 GiB float32"] = {"launches": 1, "bytes_read": 4 * n, "bytes_written": 4 * n}
    print(json.dumps({"kernel_profile": out}), flush=True)


def copy_ceiling(bytes_moved: int, repeats: int = 10):
    """Plain copies moving the step kernel's algorithmic bytes per launch, half
    read and half written, timed with HIP events on the null stream: the
    runtime's device-to-device hipMemcpyAsync and a vectorised elementwise
    kernel. A plain copy is the most favourable access pattern for the HBM, so
    it prices what a kernel that must move these bytes can reach on this part,
    beside the vendor peak."""
    ev = HipEvents()
    hip = ev.hip
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipMemset.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t]
    hip.hipMemcpyAsync.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int,
                                   ctypes.c_void_p]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    half = bytes_moved // 2
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    if hip.hipMalloc(ctypes.byref(a), half) != 0:
        raise RuntimeError("hipMalloc failed")
    if hip.hipMalloc(ctypes.byref(b), half) != 0:
        hip.hipFree(a)
        raise RuntimeError("hipMalloc failed")
    try:
        if hip.hipMemset(a, 1, half) != 0:
            raise RuntimeError("hipMemset failed")
        evs = ev.create(2 * repeats + 2)
        times = []
        for r in range(repeats + 1):  # the first copy is a warm-up
            ev.record(evs[2 * r])
            if hip.hipMemcpyAsync(b, a, half, 3, None) != 0:  # hipMemcpyDeviceToDevice
                raise RuntimeError("hipMemcpyAsync failed")
            ev.record(evs[2 * r + 1])
        for r in range(1, repeats + 1):
            times.append(ev.elapsed_ms(evs[2 * r], evs[2 * r + 1]))
        ev.destroy(evs)
    finally:
        hip.hipDeviceSynchronize()
        hip.hipFree(a)
        hip.hipFree(b)
    ms = float(np.mean(times))
    memcpy = {"gbs": round(2 * half / (ms * 1e-3) / 1e9, 1), "avg_ms": round(ms, 5)}

    # a vectorised elementwise kernel (torch's neg over float32, 16-byte loads
    # and stores) moving the same bytes, on the same null stream
    import torch
    src = torch.ones(half // 4, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    evs = ev.create(2 * repeats + 2)
    for r in range(repeats + 1):
        ev.record(evs[2 * r])
        torch.neg(src, out=dst)
        ev.record(evs[2 * r + 1])
    times = [ev.elapsed_ms(evs[2 * r], evs[2 * r + 1]) for r in range(1, repeats + 1)]
    ev.destroy(evs)
    torch.cuda.synchronize()
    if not bool((dst == -1).all()):
        raise RuntimeError("elementwise copy gave wrong values")
    del src, dst
    torch.cuda.empty_cache()
    ms = float(np.mean(times))
    elementwise = {"gbs": round(2 * (half // 4) * 4 / (ms * 1e-3) / 1e9, 1), "avg_ms": round(ms, 5)}
    return {"gbs": max(memcpy["gbs"], elementwise["gbs"]), "hipMemcpyAsync": memcpy,
            "elementwise_f32": elementwise, "repeats": repeats,
            "how": f"{half} B read and written (the step kernel's bytes per launch) by hipMemcpyAsync "
                   f"device-to-device and by a vectorised float32 elementwise kernel, HIP events on the "
                   f"null stream; gbs is the faster"}


def ensemble_cpu_baseline(cases, steps):
    """OpenLB's CPU_SIMD + OpenMP path, the way OpenLB runs UQ samples: one
    lattice after the other.  Every 32nd cavity runs 200 steps; the aggregate
    rate over that sample prices the whole ensemble's cell updates."""
    from oracle import compare_cpu
    if not os.path.exists(compare_cpu.REF_SIMD):
        return None
    threads = compare_cpu.default_threads()
    sample = cases[::32]
    lups = secs = 0.0
    for c in sample:
        r = compare_cpu.time_reference(c, 200, threads)
        lups += r["cells"] * r["steps"]
        secs += r["seconds"]
    rate = lups / secs
    total = sum(c.ncore for c in cases) * steps / rate
    host = compare_cpu.host_info(threads)
    return {"value": round(len(cases) / total, 3), "unit": "solves/s", "mlups": round(rate / 1e6, 1),
            "cores": threads, "kind": "reference", **host,
            "sample": f"{len(sample)} of the {len(cases)} cavities, 200 steps each, OpenLB 1.8 CPU_SIMD "
                      f"({host['isa']}) + OpenMP, {threads} threads, one lattice after the other; solves/s = "
                      f"members / (all members' cell updates at {steps} steps / aggregate rate)"}


def guarded_ensemble(cases, steps, ref):
    """The library's default for FP16 ensembles (openlb_amd.ensemble.GuardedEnsemble):
    FP16 storage, except members with tau < FP16_TAU_GUARD, near BGK's stability
    limit, where weakly damped non-hydrodynamic modes amplify FP16 rounding;
    those run in an FP32 ensemble beside it.  One timed region steps both
    (each a HIP-graph replay).  Velocity errors against the FP64 ensemble
    (`ref`, the members' velocities in case order)."""
    from openlb_amd.ensemble import FP16_TAU_GUARD, GuardedEnsemble
    with GuardedEnsemble.for_cases(cases) as ens:
        members = [ens.add(c, steps=steps) for c in cases]
        for m, c in zip(members, cases):
            ens.ini_equilibrium(m, c.rho_u)
        launch = ens.autotune(trial_steps=5)
        ens.step(0, graph=True)
        ens.status(members[0])
        t0 = time.perf_counter()
        ens.step(steps, graph=True)
        done = [ens.status(m)[0] for m in (members[0], members[-1])]   # synchronises the device
        seconds = time.perf_counter() - t0
        if done != [steps, steps]:
            raise SystemExit("guarded ensemble member did not finish")
        fp32 = sum(ens.storage_of(m) == "f32" for m in members)
        moments = [ens.moments(m)[1:] for m in members]
    errs = []
    for k, c in enumerate(cases):
        fluid = c.ids == 1
        errs.append(float(np.abs(moments[k][:, fluid] - ref[k][:, fluid]).max() / c.walls[3][1][0]))
    return {"guard": f"tau < {FP16_TAU_GUARD} in FP32 storage", "fp32_members": fp32, "launch": launch,
            "seconds": round(seconds, 3), "solves_per_s": round(len(cases) / seconds, 2),
            "velocity_error_max_abs_over_lid_velocity": {"max": round(max(errs), 7),
                                                         "median": round(float(np.median(errs)), 7)},
            "bound": 1e-3, "members_over_1e-3_max_norm": int(sum(e > 1e-3 for e in errs))}


def ensemble_bench(count=512, steps=2000, seed=20261204, no_cpu=False):
    """Config 4: `count` independent D3Q19 BGK lid-driven cavities (edges log-uniform
    in [16, 96], tau in [0.52, 1.2], lid velocity in [0.02, 0.1]), each run for `steps`
    steps. The plain baseline gives every cavity its own lattice and launches its steps
    one after another, as the reference runs UQ samples one SuperLattice at a time. The
    ensemble advances all members with one launch per step and replays the step loop
    from a HIP graph, with the launch shape the ensemble autotuner picks on the
    members themselves (outside the timed region). Velocity error: the FP16- and FP32-stored ensembles against the FP64
    ensemble, relative to each cavity's lid velocity (the FP64 path is pinned to OpenLB at
    1e-12); the FP32 figure on the member with the largest FP16 error tells how sensitive
    that flow is to rounding at all."""
    from openlb_amd import workloads as W
    from openlb_amd.ensemble import Ensemble
    from openlb_amd.lattice import BlockLattice
    cases = W.cavity_ensemble(count, seed=seed)
    cells = sum(c.ncore for c in cases)
    lups = cells * steps
    out = {"workload": f"cavity_ensemble_d3q19_bgk x{count}", "members": count, "steps_each": steps,
           "cells": cells, "seed": seed}

    lats = []
    for c in cases:
        lat = BlockLattice.from_case(c, "f32")
        lat.ini_equilibrium(c.rho_u)
        lats.append(lat)
    lats[0].synchronize()
    t0 = time.perf_counter()
    for lat in lats:
        lat.collide_and_stream(steps)
    lats[0].synchronize()
    plain = time.perf_counter() - t0
    for lat in lats:
        lat.close()
    del lats
    out["plain_f32"] = {"seconds": round(plain, 3), "solves_per_s": round(count / plain, 2),
                        "mlups": round(lups / plain / 1e6, 1)}
    log(f"ensemble: plain per-lattice FP32 launches {plain:.2f} s")

    final = {}
    for storage in ("f32", "f16", "f64"):
        pop = sum(Ensemble.bytes_for(19, storage, c.shape)[0] for c in cases)
        ids = sum(Ensemble.bytes_for(19, storage, c.shape)[1] for c in cases)
        with Ensemble(19, storage, "bgk", max_members=count, population_bytes=pop + (1 << 20),
                      cell_capacity=ids + (1 << 20)) as ens:
            members = [ens.add(c, steps=steps) for c in cases]
            for m, c in zip(members, cases):
                ens.ini_equilibrium(m, c.rho_u)
            launch = ens.autotune(trial_steps=5)     # members restored afterwards
            ens.step(0, graph=True)
            t0 = time.perf_counter()
            ens.step(steps, graph=True)
            done0 = ens.status(members[0])[0]
            dt = time.perf_counter() - t0     # status() synchronises the device
            if done0 != steps:
                raise SystemExit(f"ensemble member ran {done0} steps, expected {steps}")
            out[f"ensemble_{storage}"] = {"seconds": round(dt, 3), "solves_per_s": round(count / dt, 2),
                                          "mlups": round(lups / dt / 1e6, 1), "launch": launch}
            log(f"ensemble {storage}: {dt:.2f} s")
            final[storage] = [ens.moments(m)[1:] for m in members]

    def velocity_errors(storage):
        errs, l2s = [], []
        for c, a, b in zip(cases, final[storage], final["f64"]):
            fluid = c.ids == 1
            d = a[:, fluid] - b[:, fluid]
            errs.append(float(np.abs(d).max() / c.walls[3][1][0]))
            l2s.append(float(np.sqrt((d * d).sum() / (b[:, fluid] ** 2).sum())))
        return errs, l2s

    e16, l16 = velocity_errors("f16")
    e32, l32 = velocity_errors("f32")
    worst = int(np.argmax(e16))
    for storage, errs, l2s in (("f16", e16, l16), ("f32", e32, l32)):
        out[f"{storage}_velocity_error_vs_f64"] = {
            "max_abs_over_lid_velocity": {"max": round(max(errs), 7), "median": round(float(np.median(errs)), 7)},
            "l2_relative": {"max": round(max(l2s), 7), "median": round(float(np.median(l2s)), 7)},
            "members_over_1e-3_max_norm": int(sum(e > 1e-3 for e in errs))}
    # the member with the largest FP16 error, and what FP32 storage gives on it
    out["f16_velocity_error_vs_f64"]["worst_member"] = {
        "shape": list(cases[worst].shape), "tau": round(1 / cases[worst].omega, 4),
        "lid_u": round(cases[worst].walls[3][1][0], 4), "f32_error_same_member": round(e32[worst], 7)}
    out["speedup_f16_ensemble_vs_plain_f32"] = round(out["ensemble_f16"]["solves_per_s"] /
                                                     out["plain_f32"]["solves_per_s"], 2)
    out["speedup_f32_ensemble_vs_plain_f32"] = round(out["ensemble_f32"]["solves_per_s"] /
                                                     out["plain_f32"]["solves_per_s"], 2)
    out["ensemble_f16_guarded"] = guarded_ensemble(cases, steps, final["f64"])
    out["default_leg"] = "ensemble_f16_guarded"
    out["speedup_default_vs_plain_f32"] = round(out["ensemble_f16_guarded"]["solves_per_s"] /
                                                out["plain_f32"]["solves_per_s"], 2)
    log(f"ensemble f16 guarded: {out['ensemble_f16_guarded']['seconds']:.2f} s")
    if not no_cpu:
        out["cpu_baseline"] = ensemble_cpu_baseline(cases, steps)
        if out["cpu_baseline"]:
            log(f"ensemble cpu baseline {out['cpu_baseline']['value']} solves/s")
    return out


def turbulent_bench(shape=(384, 192, 192), steps=5000, no_cpu=False, cpu_seconds=10.0):
    """Config 3: D3Q27 Smagorinsky-BGK channel (C = 0.14, tau = 0.505). It is periodic in x
    and z, with BounceBack walls at y, a laminar profile and seeded +/-1% velocity noise.
    It runs `steps` steps in FP64, FP32 and FP16 storage. It reports throughput and the
    max / L2 errors of rho and u of the FP32 and FP16 runs against the FP64 run, and
    OpenLB's own CPU path on a 128^3 sample beside it. The FP64
    device path is pinned to OpenLB at 1e-12; D3Q27 Smagorinsky is the generic path, whose
    golden case matches at 2e-15."""
    from openlb_amd import workloads as W
    from openlb_amd.lattice import BlockLattice
    case = W.turbulent_channel(tuple(shape))
    fluid = case.ids == 1
    out = {"workload": "turbulent_channel_d3q27_smagorinsky", "shape": list(shape), "steps": steps,
           "tau": round(1 / case.omega, 6), "smagorinsky": case.smago}
    mom = {}
    for storage in ("f64", "f32", "f16"):
        with BlockLattice.from_case(case, storage) as lat:
            lat.ini_equilibrium(case.rho_u)
            lat.synchronize()
            t0 = time.perf_counter()
            lat.collide_and_stream(steps)
            lat.synchronize()
            dt = time.perf_counter() - t0
            m = lat.moments()
        mom[storage] = m[:, fluid]
        out[storage] = {"seconds": round(dt, 3), "mlups": round(case.n_active * steps / dt / 1e6, 1),
                        "hbm_gbs": round(2 * 27 * STORE_BYTES[storage] * case.n_active * steps / dt / 1e9, 1)}
        log(f"turbulent channel {storage}: {dt:.2f} s")
        if not np.isfinite(m).all():
            out[storage]["non_finite"] = True
    ref = mom["f64"]
    umax = float(np.abs(ref[1:]).max())
    for storage in ("f32", "f16"):
        d = mom[storage] - ref
        out[storage]["errors_vs_f64"] = {
            "rho_max": float(np.abs(d[0]).max()), "rho_l2": float(np.sqrt((d[0] ** 2).mean())),
            "u_max_over_umax": float(np.abs(d[1:]).max() / umax),
            "u_l2_relative": float(np.sqrt((d[1:] ** 2).sum() / (ref[1:] ** 2).sum()))}
    out["u_max_f64"] = umax
    # the FP16 accuracy bound of the parity tests, and what the precision guard
    # makes of this case by default
    from openlb_amd.ensemble import guarded_storage
    out["f16"]["bound_u_max_over_umax"] = 1e-3
    out["f16"]["within_bound"] = bool(out["f16"]["errors_vs_f64"]["u_max_over_umax"] <= 1e-3)
    out["f16"]["guarded_storage"] = guarded_storage(case, "f16")
    if not no_cpu:
        # OpenLB's CPU_SIMD + OpenMP path on a 128^3 sample of the same channel (FP32)
        out["cpu_baseline_f32"] = cpu_baseline(case, "f32", cpu_seconds)
        log(f"turbulent channel cpu baseline {out['cpu_baseline_f32']['value']} MLUPS")
    return out


def porous_variants(steps=50):
    """Config 2's other sizes and precisions: the same seeded porous channel at
    512x256x256 in FP32, and at 256^3 and 512x256x256 in FP64 (library default
    launch shapes).  MLUPS over cells with dynamics; TB/s from the algorithmic
    bytes 2*19*(bytes per population) per cell update."""
    from openlb_amd import workloads as W
    from openlb_amd.lattice import BlockLattice
    out = []
    for shape, storage in (((512, 256, 256), "f32"), ((256, 256, 256), "f64"), ((512, 256, 256), "f64")):
        case = W.porous_channel(shape)
        cells = case.n_active
        with BlockLattice.from_case(case, storage) as lat:
            lat.ini_equilibrium(case.rho_u)
            lat.collide_and_stream(10)
            lat.synchronize()
            t0 = time.perf_counter()
            lat.collide_and_stream(steps)
            lat.synchronize()
            dt = (time.perf_counter() - t0) / steps
            launch = lat.launch_config()
        b = STORE_BYTES[storage]
        out.append({"shape": list(shape), "storage": storage, "steps": steps, "ms_per_step": round(dt * 1e3, 4),
                    "mlups": round(cells / dt / 1e6, 1),
                    "hbm_gbs": round(2 * 19 * b * cells / dt / 1e9, 1), "launch": launch})
        log(f"porous channel {shape} {storage}: {cells / dt / 1e6:.0f} MLUPS")
    return out


def cavity_bench(size=200, steps=200, cpu_size=100, cpu_seconds=10.0, no_cpu=False):
    """OpenLB's own benchmark harness case (examples/laminar/cavity3dBenchmark,
    cavity3d.cpp:154-250) in its default configuration: a (size+1)^3 lid-driven
    cavity, D3Q19 BGK, FP32, LocalVelocity boundaries on the whole box (the
    regularized face dynamics and the edge / corner post-processors), lid
    velocity 0.01, tau = 1/2 + 3e-5 size; `size` 200 is the resolution of the
    harness's cluster recipe (bwUniCluster2.0.sh:34-39).  10 warm-up steps, then
    `steps` timed steps, MLUPS over all cells.  The bounce-back-lid variant
    (LID_DRIVEN_BOUNCE_BACK) is timed beside it, and OpenLB's CPU_SIMD build on
    the same case at `cpu_size`."""
    from openlb_amd import workloads as W
    from openlb_amd.lattice import BlockLattice
    out = {"workload": "cavity3d_benchmark", "size": size, "steps": steps, "storage": "f32"}
    for label, case in (("local_velocity", W.cavity_benchmark_local(size)),
                        ("bounce_back", W.cavity_benchmark(size))):
        with BlockLattice.from_case(case, "f32") as lat:
            lat.ini_equilibrium(case.rho_u)
            lat.initialize()
            lat.collide_and_stream(10)
            lat.synchronize()
            t0 = time.perf_counter()
            lat.collide_and_stream(steps)
            lat.synchronize()
            dt = time.perf_counter() - t0
            st = lat.statistics()
        mlups = case.n_active * steps / dt / 1e6
        out[label] = {"seconds": round(dt, 3), "mlups": round(mlups, 1),
                      "hbm_frac": round(mlups * 1e6 * 2 * 19 * 4 / (HBM_PEAK_GBS * 1e9), 4),
                      "max_u": st["max_u"], "average_energy": st["average_energy"]}
        log(f"cavity {label} {size}: {mlups:.0f} MLUPS")
    # the same LocalVelocity cavity on 2x2x2 blocks of one device (OpenLB's
    # cuboids): halo copies, and post-processors joined across the blocks
    from openlb_amd.super_lattice import SuperLattice
    case = W.cavity_benchmark_local(size)
    with SuperLattice(case, "f32", (2, 2, 2)) as sup:
        sup.ini_equilibrium(case.rho_u)
        sup.initialize()
        sup.collide_and_stream(10)
        sup.synchronize()
        t0 = time.perf_counter()
        sup.collide_and_stream(steps)
        sup.synchronize()
        dt = time.perf_counter() - t0
    mlups = case.n_active * steps / dt / 1e6
    out["local_velocity_split_2x2x2"] = {"seconds": round(dt, 3), "mlups": round(mlups, 1),
                                         "hbm_frac": round(mlups * 1e6 * 2 * 19 * 4 / (HBM_PEAK_GBS * 1e9), 4)}
    log(f"cavity local_velocity {size} on 2x2x2 blocks: {mlups:.0f} MLUPS")
    if not no_cpu:
        from oracle import compare_cpu
        if os.path.exists(compare_cpu.REF_SIMD):
            threads = compare_cpu.default_threads()
            sample = W.cavity_benchmark_local(cpu_size)
            probe = compare_cpu.time_reference(sample, 5, threads)
            n = max(5, int(cpu_seconds * probe["mlups"] * 1e6 / sample.n_active))
            res = compare_cpu.time_reference(sample, n, threads)
            out["cpu_reference"] = {"mlups": round(res["mlups"], 2), "cores": threads,
                                    **compare_cpu.host_info(threads),
                                    "sample": f"cavity {cpu_size + 1}^3 local velocity f32, {n} steps, "
                                              f"OpenLB 1.8 CPU_SIMD ({res['isa']}) + OpenMP"}
    return out


def serving_bench(count=400, rate=200.0, budget_gb=200.0, tick_steps=64):
    """Config 5: a Poisson stream of D3Q19 channel solves at `rate` req/s (sizes
    log-uniform in [24^3, 128^3], 500-5000 steps, a LocalVelocity inlet with inflow
    velocity N(0.05, 0.01^2) and a LocalPressure outlet, FP32 or FP16 storage at
    random), served with continuous batching. The stream is finite
    (`count` arrivals): 200 req/s offers ~2.3e11 cell updates per second, several times
    what one GPU sustains, so latency is dominated by queueing and processor sharing
    and is reported with the aggregate MLUPS. The arenas are sized to the stream's
    total footprint, capped at the HBM budget."""
    from openlb_amd import serving
    from openlb_amd.ensemble import Ensemble
    reqs = serving.poisson_channel_stream(count=count, rate=rate)
    need = {"f32": 0, "f16": 0}
    for r in reqs:
        need[r.storage] += Ensemble.bytes_for(19, r.storage, r.case.shape)[0]
    budget = int(min(budget_gb * 1e9, 2.2 * max(need.values())))
    with serving.SolveServer(budget_bytes=budget, max_members=count, tick_steps=tick_steps) as srv:
        t0 = time.perf_counter()
        results = srv.run(reqs, realtime=True)
        span = time.perf_counter() - t0
        out = serving.summarize(results, span)
        out["max_concurrent"] = srv.max_active
    out.update({"workload": "poisson_channel_d3q19_bgk_velocity_inlet_pressure_outlet", "rate_per_s": rate,
                "arrival_window_s": round(reqs[-1].arrival, 3), "tick_steps": tick_steps,
                "arena_bytes": budget, "hbm_budget_bytes": int(budget_gb * 1e9),
                "offered_mlups": round(sum(r.lups for r in reqs) / reqs[-1].arrival / 1e6, 1),
                "mean_mlup_per_request": round(sum(r.lups for r in reqs) / count / 1e6, 2),
                "fp16_share": round(sum(r.storage == "f16" for r in reqs) / count, 3)})
    return out


def export_bench(lat):
    """Result output without the populations: the VTK buffers of density and
    velocity (lbmh_field_export over the block plus one overlap layer,
    evaluated on the device) against a download of all populations."""
    from openlb_amd import vtk
    fs = [vtk.SuperLatticeDensity3D(lat), vtk.SuperLatticeVelocity3D(lat)]
    for f in fs:
        f.evaluate(1)
    lat.synchronize()
    t0 = time.perf_counter()
    nbytes = sum(f.evaluate(1).nbytes for f in fs)
    t_export = time.perf_counter() - t0
    t0 = time.perf_counter()
    pops = lat.download_populations(lat.storage)
    t_pop = time.perf_counter() - t0
    return {"fields": "density+velocity (float32, VTK point order, overlap 1)",
            "field_bytes": int(nbytes), "field_export_s": round(t_export, 4),
            "population_bytes": int(pops.nbytes), "population_download_s": round(t_pop, 4)}


def measured_traffic(kernel_name, case, launch):
    """HBM bytes per launch of this kernel, workload and launch shape from the
    committed counter runs (profiles/traffic.json: FETCH_SIZE and WRITE_SIZE scaled
    by a 1 GiB device copy), or None if that combination was not profiled."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(path):
        return None
    shape = "x".join(str(n) for n in case.shape)
    key = f"{kernel_name}|{case.name}|{shape}|lds{launch['lds_bytes']}|xcd{launch['xcd_group']}"
    entry = json.load(open(path)).get(key)
    return None if entry is None else entry["traffic_bytes"]


def make_case(args):
    from openlb_amd import workloads as W
    if args.workload == "porous_channel":
        return W.porous_channel(tuple(args.shape), tau=0.55, force=1e-6)
    if args.workload == "taylor_green":
        return W.taylor_green(tuple(args.shape), u0=0.05, tau=0.6)
    if args.workload == "turbulent_channel":
        return W.turbulent_channel(tuple(args.shape))
    if args.workload == "lid_driven_cavity":
        return W.lid_driven_cavity(tuple(args.shape))
    raise SystemExit(f"unknown workload {args.workload}")


def cpu_sample(case):
    """The same generator at 128^3 at most: MLUPS is normalised per cell update."""
    from openlb_amd import workloads as W
    shape = tuple(min(n, 128) for n in case.shape)
    if case.name == "porous_channel":
        return W.porous_channel(shape, tau=0.55, force=case.force[0])
    if case.name == "taylor_green":
        return W.taylor_green(shape, u0=0.05, tau=1.0 / case.omega)
    if case.name == "turbulent_channel":
        return W.turbulent_channel(shape)
    return W.lid_driven_cavity(shape)


def cpu_baseline(case, storage, seconds):
    """OpenLB's own CPU path (CPU_SIMD + OpenMP, compiled from the reference sources
    into oracle/_ref/olb_ref_simd_f32) on the host cores, on a bounded sample of the same
    workload. If that binary is absent or the storage is not FP32, the oracle is timed
    instead (kind "port"; about 8x slower than OpenLB's SIMD path on the same cores)."""
    from oracle import compare_cpu
    threads = compare_cpu.default_threads()
    sample = cpu_sample(case)
    shape = sample.shape
    if storage == "f32" and os.path.exists(compare_cpu.REF_SIMD):
        probe = compare_cpu.time_reference(sample, 10, threads)
        steps = max(10, int(seconds * probe["mlups"] * 1e6 / sample.n_active))
        res = compare_cpu.time_reference(sample, steps, threads)
        return {"value": round(res["mlups"], 2), "unit": "MLUPS", "cores": threads, "kind": "reference",
                **compare_cpu.host_info(threads),
                "sample": f"{sample.name} {shape[0]}x{shape[1]}x{shape[2]} {storage}, {steps} steps in "
                          f"{res['seconds']:.1f} s, OpenLB 1.8 CPU_SIMD ({res['isa']}) + OpenMP, {threads} "
                          f"threads, {os.path.relpath(compare_cpu.reference_binary()[0], ROOT)}"}
    return cpu_baseline_port(sample, storage, seconds, threads)


def cpu_baseline_port(sample, storage, seconds, threads):
    from oracle import oracle as O
    shape = sample.shape
    lat = O.OracleLattice(sample, storage)
    lat.upload(O.initial_populations(sample, storage))
    lat.step(2, threads)  # warm-up
    steps = 0
    t0 = time.perf_counter()
    while True:
        lat.step(2, threads)
        steps += 2
        dt = time.perf_counter() - t0
        if dt >= seconds:
            break
    mlups = sample.n_active * steps / dt / 1e6
    from oracle import compare_cpu
    return {"value": round(mlups, 2), "unit": "MLUPS", "cores": threads, "kind": "port",
            **compare_cpu.host_info(threads),
            "sample": f"{sample.name} {shape[0]}x{shape[1]}x{shape[2]} {storage}, {steps} steps "
                      f"in {dt:.1f} s, oracle/liboracle.so (OpenMP, {threads} threads)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="porous_channel")
    ap.add_argument("--shape", type=int, nargs=3, default=[256, 256, 256])
    ap.add_argument("--storage", default="f32", choices=["f64", "f32", "f16"])
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--autotune", action="store_true",
                    help="time candidate launch shapes first and keep the fastest "
                         "(default: the library's measured default shape)")
    ap.add_argument("--no-autotune", dest="autotune", action="store_false")
    ap.add_argument("--no-ensemble", dest="ensemble", action="store_false",
                    help="skip the 512-lattice ensemble measurement (config 4)")
    ap.add_argument("--ensemble-members", type=int, default=512)
    ap.add_argument("--ensemble-steps", type=int, default=2000)
    ap.add_argument("--no-turbulent", dest="turbulent", action="store_false",
                    help="skip the D3Q27 Smagorinsky channel precision run (config 3)")
    ap.add_argument("--turbulent-steps", type=int, default=5000)
    ap.add_argument("--no-variants", dest="variants", action="store_false",
                    help="skip config 2's 512x256x256 FP32 and FP64 lines")
    ap.add_argument("--no-cavity", dest="cavity", action="store_false",
                    help="skip OpenLB's cavity benchmark case (LocalVelocity boundaries)")
    ap.add_argument("--cavity-size", type=int, default=200)
    ap.add_argument("--cavity-steps", type=int, default=200)
    ap.add_argument("--no-serving", dest="serving", action="store_false",
                    help="skip the continuous-batching serving measurement (config 5)")
    ap.add_argument("--serving-requests", type=int, default=400)
    ap.add_argument("--serving-rate", type=float, default=200.0)
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run device-to-device copies of MIB MiB (counter calibration)")
    ap.add_argument("--kernel-profile", action="store_true",
                    help="only run the short fixed kernel sequence the counter tables come from")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU: this build runs a lattice on a single MI355X and has no "
              "multi-GPU path (run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    from openlb_amd import abi
    from openlb_amd.lattice import BlockLattice

    if args.calibrate_copy:
        calibrate_copy(args.calibrate_copy)
        return
    if args.kernel_profile:
        kernel_profile()
        return

    case = make_case(args)
    log(f"workload {case.name} {case.shape} {args.storage}: {case.n_active} cells with dynamics")
    lat = BlockLattice.from_case(case, args.storage)
    lat.ini_equilibrium(case.rho_u)
    lat.synchronize()
    launch = lat.autotune(trial_steps=5) if args.autotune else lat.launch_config()
    log(f"launch config {launch}")
    ev = HipEvents()

    for _ in range(args.warmup):
        lat.collide_and_stream(1)
    lat.synchronize()

    # timed region: K steps, device-synchronised on both sides; per-launch
    # events on the launch stream (the null stream)
    marks = ev.create(2 * args.steps)
    lat.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        ev.record(marks[2 * k])
        lat.collide_and_stream(1)
        ev.record(marks[2 * k + 1])
    lat.synchronize()
    wall = time.perf_counter() - t0
    kernel_ms = [ev.elapsed_ms(marks[2 * k], marks[2 * k + 1]) for k in range(args.steps)]
    ev.destroy(marks)
    moments = lat.reduce_moments()
    export = export_bench(lat)
    log(f"export: density+velocity {export['field_export_s']} s vs populations {export['population_download_s']} s")
    if not np.isfinite(moments).all():
        raise SystemExit("non-finite populations after the timed steps")
    pattern_ms = pattern_ceiling(lat, ev, repeats=20)
    lat.close()

    cells = case.n_active
    mlups = cells * args.steps / wall / 1e6
    kernel_name = f"collide_stream_kernel<{case.q},{args.storage},{case.kind}>"
    traffic = measured_traffic(kernel_name, case, launch)
    avg_ms = float(np.mean(kernel_ms))
    bytes_per_launch = 2 * case.q * STORE_BYTES[args.storage] * cells
    achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9
    log(f"{mlups:.0f} MLUPS, kernel avg {avg_ms:.4f} ms (min {min(kernel_ms):.4f}), "
        f"{achieved:.0f} GB/s algorithmic")
    copy = copy_ceiling(bytes_per_launch)
    log(f"copies of the same bytes: hipMemcpyAsync {copy['hipMemcpyAsync']['gbs']} GB/s, "
        f"elementwise {copy['elementwise_f32']['gbs']} GB/s")
    pattern = {"gbs": round(bytes_per_launch / (pattern_ms * 1e-3) / 1e9, 1), "avg_ms": round(pattern_ms, 5),
               "how": "lbmh_pattern_ceiling: the step kernel's PS in-place read-modify-write of the Q columns "
                      "(same cells, shifts, shift evolution, launch shape and stream) with the arithmetic taken "
                      "out, 20 launches timed with HIP events after the timed steps"}
    log(f"pattern ceiling {pattern['gbs']} GB/s; kernel at {achieved / pattern['gbs']:.3f} of it")

    ensemble = None
    if args.ensemble:
        log("ensemble (config 4) ...")
        ensemble = ensemble_bench(args.ensemble_members, args.ensemble_steps, no_cpu=args.no_cpu)

    turbulent = None
    if args.turbulent:
        log("turbulent channel (config 3) ...")
        turbulent = turbulent_bench(steps=args.turbulent_steps, no_cpu=args.no_cpu)

    variants = None
    if args.variants:
        log("porous channel variants (config 2) ...")
        variants = porous_variants()

    cavity = None
    if args.cavity:
        log("cavity benchmark (LocalVelocity) ...")
        cavity = cavity_bench(args.cavity_size, args.cavity_steps, no_cpu=args.no_cpu)

    served = None
    if args.serving:
        log("serving (config 5) ...")
        served = serving_bench(args.serving_requests, args.serving_rate)
        log(f"serving: p50 {served['p50_s']} s, p99 {served['p99_s']} s, {served['aggregate_mlups']} MLUPS")
        # the same stream at a rate the device sustains: 60% of the capacity the
        # saturated run measured, in requests per second of this stream's mean size
        per_req = served["mean_mlup_per_request"]
        rate = round(0.6 * served["aggregate_mlups"] / per_req, 2)
        log(f"serving at a sustainable rate ({rate} req/s) ...")
        steady = serving_bench(args.serving_requests, rate)
        steady["rate_rule"] = "0.6 x the aggregate MLUPS of the saturated run / mean MLUP per request"
        log(f"serving at {rate} req/s: p50 {steady['p50_s']} s, p99 {steady['p99_s']} s")
        served = {"saturated": served, "sustainable": steady}

    cpu = None
    if not args.no_cpu:
        log("cpu baseline ...")
        cpu = cpu_baseline(case, args.storage, args.cpu_seconds)
        log(f"cpu baseline {cpu['value']} MLUPS")

    out = {
        "metric": "MLUPS (D3Q19 BGK fp32 256^3 porous channel) + % HBM roofline",
        "value": round(mlups, 1),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall * 1e3 / args.steps, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"f32": "f32", "f64": "f64", "f16": "f16-storage/f32-math"}[args.storage],
        "data": "synthetic (seeded)",
        "config": {
            "workload": f"{case.name}_d3q{case.q}_{case.kind}",
            "shape": list(case.shape),
            "storage": args.storage,
            "tau": round(1.0 / case.omega, 6),
            "force": list(case.force),
            "cells_with_dynamics": cells,
            "solid_cells": int(np.count_nonzero(case.ids == 2)),
            "parallelism": "single GPU",
            "device": abi.device_arch(0),
            "launch": launch,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "kernel": kernel_name,
            "kernel_avg_ms": round(avg_ms, 5),
            "algorithmic_bytes_per_launch": bytes_per_launch,
            "device_copy": copy,
            "frac_of_device_copy": round(achieved / copy["gbs"], 4),
            "pattern_ceiling": pattern,
            "frac_of_pattern_ceiling": round(achieved / pattern["gbs"], 4),
        },
        "cpu_baseline": cpu,
        "porous_channel_variants": variants,
        "turbulent_channel": turbulent,
        "ensemble": ensemble,
        "cavity_benchmark": cavity,
        "serving": served,
        "result_export": export,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
