"""Continuous-batching solve server on one GPU.

Requests are independent D3Q19 solves. Each has its own extent, step count,
driving velocity and storage type (FP32 or FP16). Two ensembles (one per storage
type) hold the running solves on the device. By default the precision guard runs
an FP16 request whose tau is below ``ensemble.FP16_TAU_GUARD`` in the FP32
ensemble, where FP16 rounding would exceed the velocity bound.

The loop repeats:

1. Admit the requests that have arrived, in arrival order, whose footprint fits
   the ensemble's arenas and for which a slot is free. Admission is checked against
   the HBM budget the arenas were sized to. A request that does not fit yet stays
   queued and later, smaller requests may pass it (skip-ahead admission), up to
   ``max_bypass`` of them: after that nothing behind it is admitted until it fits,
   so a large solve cannot starve. A request that could not fit even in an empty
   arena is rejected at once with an error result instead of waiting forever.
2. Advance every ensemble that has running solves by one tick of ``tick_steps`` steps
   (a HIP-graph replay). Each solve carries its own step budget, so it stops at
   exactly its requested step count even in mid-tick.
3. Retire the finished solves: optionally fetch their moments, free their slots and
   arena space, and record their latency.

Every solve in the mix runs the same per-cell code on the same data as it would
alone, so its result is bit-identical (tests/test_gpu_serving.py).
"""
from __future__ import annotations

import collections
import dataclasses
import math
import time

import numpy as np

from . import abi, workloads
from .ensemble import FP16_TAU_GUARD, Ensemble, guarded_storage


@dataclasses.dataclass
class Request:
    rid: int
    arrival: float          # seconds after the start of the stream
    case: workloads.Case
    steps: int
    storage: str            # "f32" or "f16"; the precision guard may run an FP16 request in FP32

    @property
    def lups(self) -> int:
        return self.case.ncore * self.steps


@dataclasses.dataclass
class Result:
    rid: int
    storage: str
    cells: int
    steps: int
    arrival: float
    admitted: float
    finished: float
    moments: np.ndarray | None = None
    populations: np.ndarray | None = None
    error: str | None = None        # set when the request was rejected and never ran
    promoted: bool = False          # FP16 request run in FP32 storage by the precision guard

    @property
    def latency(self) -> float:
        return self.finished - self.arrival


def channel_request(rid, arrival, size, steps, u_in, storage, tau=0.6):
    """A D3Q19 square duct of size^3 cells with BounceBack walls: a LocalVelocity
    inlet at x = 0 with inflow velocity u_in along x, a LocalPressure outlet at
    x = size-1 at density 1 (the regularized boundaries of
    src/boundary/localVelocity3D.h and localPressure3D.h), starting at rest."""
    case = workloads.velocity_channel((size, size, size), u_in=u_in, rho_out=1.0, tau=tau)
    case.name = f"channel_{rid}"
    case.rho_u = None     # rest state (rho 1, u 0); built at admission to keep the queue small
    return Request(rid, arrival, case, int(steps), storage)


def initial_moments(case) -> np.ndarray:
    if case.rho_u is not None:
        return case.rho_u
    rho_u = np.zeros((4,) + tuple(case.shape))
    rho_u[0] = 1.0
    return rho_u


def poisson_channel_stream(count=400, rate=200.0, seed=20261205, size=(24, 128), steps=(500, 5000),
                           u_mean=0.05, u_std=0.01, storages=("f32", "f16"), tau=0.6):
    """Config-5 request stream. Arrivals are Poisson at `rate` per second. Sizes are
    log-uniform in `size`, step counts uniform in `steps`, and inflow velocities
    normal(u_mean, u_std^2) as UQ samples. The storage type is drawn uniformly from `storages`."""
    rng = np.random.Generator(np.random.PCG64(seed))
    t = 0.0
    out = []
    for rid in range(count):
        t += float(rng.exponential(1.0 / rate))
        s = int(round(math.exp(rng.uniform(math.log(size[0]), math.log(size[1])))))
        n = int(rng.integers(steps[0], steps[1] + 1))
        u = float(rng.normal(u_mean, u_std))
        st = storages[int(rng.integers(0, len(storages)))]
        out.append(channel_request(rid, t, s, n, u, st, tau))
    return out


class SolveServer:
    def __init__(self, budget_bytes: int, storages=("f32", "f16"), max_members: int = 1024,
                 tick_steps: int = 64, keep_moments: bool = False, keep_populations: bool = False,
                 device: int = 0, max_bypass: int = 64, precision_guard: float | None = FP16_TAU_GUARD):
        """``precision_guard``: FP16 requests with tau below it run in FP32 storage
        (``ensemble.guarded_storage``) when the server runs an FP32 ensemble; None
        runs every request in the storage it asks for."""
        self.tick_steps = int(tick_steps)
        self.precision_guard = precision_guard
        self.max_bypass = int(max_bypass)
        self.keep_moments = keep_moments
        self.keep_populations = keep_populations
        share = budget_bytes // len(storages)
        self.ens = {}
        self.capacity = {}
        for st in storages:
            # id arena: one byte per cell; a member's populations take at least
            # 2*19 bytes per cell, so cells <= pop bytes / 38
            cells = share // (19 * {"f32": 4, "f16": 2, "f64": 8}[st]) + (1 << 20)
            self.ens[st] = Ensemble(19, st, "bgk", max_members=max_members,
                                    population_bytes=share, cell_capacity=cells, device=device)
            self.capacity[st] = (self.ens[st].usage()["bytes_capacity"], cells)
        self.max_active = 0

    def close(self):
        for e in self.ens.values():
            e.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def never_fits(self, req: Request) -> str | None:
        """Why the request could not run even on an empty server, or None."""
        if req.storage not in self.ens:
            return f"storage {req.storage} is not served (served: {sorted(self.ens)})"
        pop, ids = Ensemble.bytes_for(19, req.storage, req.case.shape)
        pop_cap, cell_cap = self.capacity[req.storage]
        if pop > pop_cap:
            return f"needs {pop} population bytes, the {req.storage} arena holds {pop_cap}"
        if ids > cell_cap:
            return f"needs {ids} id bytes, the {req.storage} id arena holds {cell_cap}"
        return None

    def guard(self, req: Request) -> Request:
        """The request as it runs: an FP16 request below the precision guard's tau
        goes to the FP32 ensemble."""
        if req.storage == "f16" and "f32" in self.ens:
            st = guarded_storage(req.case, "f16", self.precision_guard)
            if st != req.storage:
                return dataclasses.replace(req, storage=st)
        return req

    def _try_admit(self, req: Request):
        ens = self.ens[req.storage]
        try:
            m = ens.add(req.case, steps=req.steps)
        except abi.LbmhError as exc:
            if exc.code == abi.E_MEMORY:
                return None
            raise
        try:
            ens.ini_equilibrium(m, initial_moments(req.case))
            ens.initialize(m)
        except BaseException:
            ens.remove(m)
            raise
        return m

    def run(self, requests, realtime: bool = True) -> list[Result]:
        """Serve the stream. With realtime=True a request becomes visible at its
        arrival time on the wall clock. Otherwise every request is present at time zero.
        Rejected requests come back as results with ``error`` set and no steps run."""
        queue = sorted(requests, key=lambda r: r.arrival)
        head = 0                 # requests before `head` in `queue` have all arrived
        waiting = []             # arrived, not yet admitted: [request, times passed over]
        active = {}              # (storage, member) -> (request, admitted)
        promoted = set()         # rids the precision guard moved to FP32
        results = []
        t0 = time.perf_counter()

        def reject(req, why, at):
            results.append(Result(req.rid, req.storage, req.case.ncore, 0, req.arrival, at, at, error=why))

        while head < len(queue) or waiting or active:
            now = time.perf_counter() - t0
            while head < len(queue) and (not realtime or queue[head].arrival <= now):
                req = queue[head]
                head += 1
                run_as = self.guard(req)
                if run_as is not req:
                    promoted.add(req.rid)
                    req = run_as
                why = self.never_fits(req)
                if why is None:
                    waiting.append([req, 0])
                else:
                    reject(req, why, time.perf_counter() - t0)
            # skip-ahead admission in arrival order, bounded by max_bypass
            kept = []
            blocked = False
            for entry in waiting:
                req = entry[0]
                m = None if blocked else self._try_admit(req)
                if m is None:
                    if not active and not blocked:
                        # nothing runs, so nothing will free up: it cannot ever fit
                        reject(req, "does not fit in the empty server", time.perf_counter() - t0)
                        continue
                    kept.append(entry)
                    blocked = blocked or entry[1] >= self.max_bypass
                    continue
                active[(req.storage, m)] = (req, time.perf_counter() - t0)
                for older in kept:
                    older[1] += 1
                    blocked = blocked or older[1] >= self.max_bypass
            waiting = kept
            self.max_active = max(self.max_active, len(active))
            if not active:
                if head < len(queue) and not waiting:
                    time.sleep(max(0.0, queue[head].arrival - (time.perf_counter() - t0)))
                continue
            running = {st for st, _ in active}
            for st in running:
                self.ens[st].step(self.tick_steps, graph=True)
            for st in running:
                done, left = self.ens[st].status_all()     # waits for the tick
                finished = time.perf_counter() - t0
                for (s2, m), (req, adm) in list(active.items()):
                    if s2 != st or left[m] != 0:
                        continue
                    if done[m] != req.steps:
                        raise RuntimeError(f"request {req.rid} ran {done[m]} of {req.steps} steps")
                    res = Result(req.rid, st, req.case.ncore, req.steps, req.arrival, adm, finished,
                                 promoted=req.rid in promoted)
                    if self.keep_moments:
                        res.moments = self.ens[st].moments(m)
                    if self.keep_populations:
                        res.populations = self.ens[st].download(m)
                    self.ens[st].remove(m)
                    del active[(st, m)]
                    results.append(res)
        return results


def summarize(results: list[Result], makespan: float | None = None) -> dict:
    rejected = sum(r.error is not None for r in results)
    results = [r for r in results if r.error is None]
    lat = np.array([r.latency for r in results])
    lups = sum(r.cells * r.steps for r in results)
    span = makespan if makespan is not None else max(r.finished for r in results)
    return {"requests": len(results), "rejected": rejected, "p50_s": round(float(np.percentile(lat, 50)), 4),
            "p99_s": round(float(np.percentile(lat, 99)), 4), "max_s": round(float(lat.max()), 4),
            "makespan_s": round(span, 3), "aggregate_mlups": round(lups / span / 1e6, 1),
            "solves_per_s": round(len(results) / span, 2),
            "promoted_to_f32": sum(r.promoted for r in results)}
