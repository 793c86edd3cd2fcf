"""Python front end of a lattice ensemble (C ABI ``lbmh_ensemble_*``).

An ensemble holds many independent block lattices on one GPU. Examples are
uncertainty-quantification samples or parameter sweeps, which the reference
runs one ``SuperLattice`` at a time
(``examples/uncertaintyQuantification/cavity3d/cavity3d.cpp:110-148``). The
ensemble advances all of its members with one launch per step.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import abi


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


_NP = {abi.F64: np.float64, abi.F32: np.float32, abi.F16: np.float16}


class Ensemble:
    def __init__(self, q: int = 19, storage: str = "f32", kind: str = "bgk", max_members: int = 512,
                 population_bytes: int = 1 << 30, cell_capacity: int = 1 << 24, device: int = 0,
                 boundary_capacity: int = 0):
        d = abi.EnsembleDesc()
        d.q, d.storage, d.bulk_kind, d.device = q, abi.STORAGE[storage], abi.KINDS[kind], device
        d.max_members = max_members
        d.population_bytes, d.cell_capacity = int(population_bytes), int(cell_capacity)
        d.boundary_capacity = int(boundary_capacity)
        h = ctypes.c_void_p()
        abi.check(abi.load().lbmh_ensemble_create(ctypes.byref(d), ctypes.byref(h)))
        self._h = h
        self.q, self.storage, self.kind = q, storage, kind
        self.max_members = max_members
        self.shapes: dict[int, tuple] = {}

    @staticmethod
    def bytes_for(q: int, storage: str, shape) -> tuple[int, int]:
        """(population bytes, id bytes) a member of this extent takes. This matches
        lbmh_ensemble_member_bytes without needing a device."""
        npad = (shape[0] + 2) * (shape[1] + 2) * (shape[2] + 2)
        elem = {"f64": 8, "f32": 4, "f16": 2}[storage]
        stride = -(-npad * elem // 256) * 256 // elem
        pop = -(-stride * elem * q // 256) * 256
        ids = -(-int(np.prod(shape)) // 256) * 256
        return pop, ids

    def close(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            abi.load().lbmh_ensemble_destroy(h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def member_bytes(self, shape) -> tuple[int, int]:
        a, b = ctypes.c_size_t(), ctypes.c_size_t()
        abi.check(abi.load().lbmh_ensemble_member_bytes(self._h, (ctypes.c_int * 3)(*shape),
                                                        ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def add(self, case, steps: int = -1) -> int:
        """Add a workloads.Case as a member. It returns the member id. The walls come
        from case.walls, and the populations start at zero."""
        if getattr(case, "force_field", None) is not None:
            raise ValueError("ensemble members take a uniform FORCE; a per-cell FORCE field needs a BlockLattice")
        ids = np.ascontiguousarray(case.ids, dtype=np.uint8)
        p = abi.BulkParams()
        p.omega, p.magic, p.smago = case.omega, case.magic, case.smago
        for d in range(3):
            p.force[d] = case.force[d]
        m = ctypes.c_int(-1)
        abi.check(abi.load().lbmh_ensemble_add(self._h, (ctypes.c_int * 3)(*case.shape),
                                               (ctypes.c_int * 3)(*case.periodic), _ptr(ids),
                                               ctypes.byref(p), int(steps), ctypes.byref(m)))
        local = getattr(case, "local", {}) or {}
        try:
            for wid, (rho, u) in case.walls.items():
                if wid in local:
                    continue
                abi.check(abi.load().lbmh_ensemble_set_wall(self._h, m.value, int(wid), float(rho),
                                                            (ctypes.c_double * 3)(*map(float, u))))
            for mid, kind in local.items():
                rho, u = case.walls.get(mid, (1.0, (0.0, 0.0, 0.0)))
                abi.check(abi.load().lbmh_ensemble_set_local_boundary(
                    self._h, m.value, int(mid),
                    {"velocity": 0, "pressure": 1, "interp_velocity": 2, "interp_pressure": 3}[kind],
                    float(rho), (ctypes.c_double * 3)(*map(float, u))))
        except BaseException:
            # a half-configured member must not keep its slot and arena space
            abi.load().lbmh_ensemble_remove(self._h, m.value)
            raise
        self.shapes[m.value] = tuple(case.shape)
        return m.value

    def initialize(self, member: int):
        """SuperLattice::initialize() of one member: its local boundaries' edge and
        corner post-processors once."""
        abi.check(abi.load().lbmh_ensemble_initialize(self._h, int(member)))

    def remove(self, member: int):
        abi.check(abi.load().lbmh_ensemble_remove(self._h, int(member)))
        self.shapes.pop(member, None)

    def set_steps(self, member: int, steps: int):
        abi.check(abi.load().lbmh_ensemble_set_steps(self._h, int(member), int(steps)))

    def status(self, member: int) -> tuple[int, int]:
        done, left = ctypes.c_int64(), ctypes.c_int64()
        abi.check(abi.load().lbmh_ensemble_status(self._h, int(member), ctypes.byref(done),
                                                  ctypes.byref(left)))
        return done.value, left.value

    def status_all(self) -> tuple[np.ndarray, np.ndarray]:
        """(steps done, steps left) of every slot, one device round trip."""
        done = np.zeros(self.max_members, dtype=np.int64)
        left = np.zeros(self.max_members, dtype=np.int64)
        abi.check(abi.load().lbmh_ensemble_status_all(self._h, _ptr(done), _ptr(left)))
        return done, left

    def ini_equilibrium(self, member: int, rho_u: np.ndarray):
        a = np.ascontiguousarray(rho_u, dtype=np.float64)
        abi.check(abi.load().lbmh_ensemble_ini_equilibrium(self._h, int(member), _ptr(a)))

    def upload(self, member: int, f: np.ndarray, dtype: str = "f64"):
        code = abi.STORAGE[dtype]
        a = np.ascontiguousarray(f, dtype=_NP[code])
        abi.check(abi.load().lbmh_ensemble_upload(self._h, int(member), _ptr(a), code))

    def download(self, member: int, dtype: str = "f64") -> np.ndarray:
        code = abi.STORAGE[dtype]
        out = np.empty((self.q,) + self.shapes[member], dtype=_NP[code])
        abi.check(abi.load().lbmh_ensemble_download(self._h, int(member), _ptr(out), code))
        return out

    def checkpoint(self, member: int, real: str = "f64") -> bytes:
        """OpenLB's serialized bytes of one member, as BlockLattice.checkpoint."""
        code = {"f64": abi.F64, "f32": abi.F32}[real]
        n = ctypes.c_size_t(0)
        L = abi.load()
        abi.check(L.lbmh_ensemble_checkpoint_size(self._h, int(member), code, ctypes.byref(n)))
        buf = ctypes.create_string_buffer(n.value)
        abi.check(L.lbmh_ensemble_checkpoint_save(self._h, int(member), code, buf, n.value))
        return buf.raw

    def restore(self, member: int, data: bytes, real: str = "f64"):
        code = {"f64": abi.F64, "f32": abi.F32}[real]
        buf = ctypes.create_string_buffer(bytes(data), len(data))
        abi.check(abi.load().lbmh_ensemble_checkpoint_load(self._h, int(member), code, buf, len(data)))

    def moments(self, member: int) -> np.ndarray:
        out = np.empty((4,) + self.shapes[member])
        abi.check(abi.load().lbmh_ensemble_moments(self._h, int(member), _ptr(out)))
        return out

    def step(self, steps: int = 1, graph: bool = False, stream: int | None = None):
        fn = abi.load().lbmh_ensemble_step_graph if graph else abi.load().lbmh_ensemble_step
        abi.check(fn(self._h, int(steps), ctypes.c_void_p(stream or 0)))

    def set_launch_config(self, lds_bytes: int, xcd_group: int):
        cfg = abi.LaunchConfig(int(lds_bytes), int(xcd_group))
        abi.check(abi.load().lbmh_ensemble_set_launch_config(self._h, ctypes.byref(cfg)))

    def autotune(self, trial_steps: int = 5) -> dict:
        """Time the launch shapes on the members themselves and keep the fastest;
        the members' populations and step state are restored afterwards."""
        cfg = abi.LaunchConfig()
        abi.check(abi.load().lbmh_ensemble_autotune(self._h, int(trial_steps), ctypes.byref(cfg)))
        return {"lds_bytes": cfg.lds_bytes, "xcd_group": cfg.xcd_group}

    def usage(self) -> dict:
        m = ctypes.c_int()
        cells, used, cap = ctypes.c_size_t(), ctypes.c_size_t(), ctypes.c_size_t()
        abi.check(abi.load().lbmh_ensemble_usage(self._h, ctypes.byref(m), ctypes.byref(cells),
                                                 ctypes.byref(used), ctypes.byref(cap)))
        return {"members": m.value, "cells": cells.value, "bytes_used": used.value,
                "bytes_capacity": cap.value}


# FP16 population storage keeps the velocity error under 1e-3 of the driving
# velocity for the cavity ensembles down to tau ~ 0.53; closer to BGK's
# stability limit (tau -> 1/2) weakly damped non-hydrodynamic modes amplify the
# FP16 rounding past that bound.  Members below this tau run in FP32 storage.
FP16_TAU_GUARD = 0.53


def guarded_storage(case, storage: str, tau_guard: float | None = FP16_TAU_GUARD) -> str:
    """The storage a member asked to run in `storage` gets under the precision
    guard: FP16 requests with tau < tau_guard run in FP32; None turns the guard off."""
    if storage == "f16" and tau_guard is not None and 1.0 / case.omega < tau_guard:
        return "f32"
    return storage


class GuardedEnsemble:
    """An FP16 ensemble with the per-member precision guard: members whose tau is
    below ``tau_guard`` go to an FP32 ensemble beside it (``guarded_storage``).
    Member ids are the composite's own; every call goes to the member's
    ensemble, and ``step`` advances both ensembles, back to back on the stream.
    Each member runs the same per-cell code as it would in a plain ensemble of
    its storage, so its result is bit-identical to that."""

    def __init__(self, q: int = 19, kind: str = "bgk", max_members: int = 512,
                 population_bytes: int = 1 << 30, cell_capacity: int = 1 << 24, device: int = 0,
                 boundary_capacity: int = 0, tau_guard: float | None = FP16_TAU_GUARD,
                 fp32_population_bytes: int | None = None, fp32_cell_capacity: int | None = None,
                 fp32_max_members: int | None = None):
        self.q, self.kind, self.tau_guard = q, kind, tau_guard
        self.parts: dict[str, Ensemble] = {}
        self.parts["f16"] = Ensemble(q, "f16", kind, max_members, population_bytes, cell_capacity, device,
                                     boundary_capacity)
        try:
            self.parts["f32"] = Ensemble(
                q, "f32", kind, fp32_max_members or max_members,
                fp32_population_bytes if fp32_population_bytes is not None else max(population_bytes // 4, 1 << 20),
                fp32_cell_capacity if fp32_cell_capacity is not None else max(cell_capacity // 4, 1 << 16),
                device, boundary_capacity)
        except BaseException:
            self.parts["f16"].close()
            raise
        self._where: dict[int, tuple[str, int]] = {}
        self._count = {"f16": 0, "f32": 0}
        self._next = 0

    @classmethod
    def for_cases(cls, cases, q: int = 19, kind: str = "bgk", tau_guard: float | None = FP16_TAU_GUARD,
                  device: int = 0, slack: int = 1 << 20):
        """A guarded ensemble whose arenas hold exactly `cases` (plus `slack` bytes each)."""
        need = {"f16": [0, 0, 0], "f32": [0, 0, 0]}
        for c in cases:
            st = guarded_storage(c, "f16", tau_guard)
            pop, ids = Ensemble.bytes_for(q, st, c.shape)
            need[st][0] += pop
            need[st][1] += ids
            need[st][2] += 1
        return cls(q, kind, max(need["f16"][2], 1), need["f16"][0] + slack, need["f16"][1] + slack, device,
                   tau_guard=tau_guard, fp32_population_bytes=need["f32"][0] + slack,
                   fp32_cell_capacity=need["f32"][1] + slack, fp32_max_members=max(need["f32"][2], 1))

    def close(self):
        for e in self.parts.values():
            e.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def storage_of(self, member: int) -> str:
        return self._where[member][0]

    def _part(self, member: int) -> tuple[Ensemble, int]:
        st, m = self._where[member]
        return self.parts[st], m

    def add(self, case, steps: int = -1) -> int:
        st = guarded_storage(case, "f16", self.tau_guard)
        m = self.parts[st].add(case, steps)
        gid = self._next
        self._next += 1
        self._where[gid] = (st, m)
        self._count[st] += 1
        return gid

    def remove(self, member: int):
        ens, m = self._part(member)
        ens.remove(m)
        self._count[self._where.pop(member)[0]] -= 1

    def initialize(self, member: int):
        ens, m = self._part(member)
        ens.initialize(m)

    def set_steps(self, member: int, steps: int):
        ens, m = self._part(member)
        ens.set_steps(m, steps)

    def status(self, member: int) -> tuple[int, int]:
        ens, m = self._part(member)
        return ens.status(m)

    def ini_equilibrium(self, member: int, rho_u: np.ndarray):
        ens, m = self._part(member)
        ens.ini_equilibrium(m, rho_u)

    def upload(self, member: int, f: np.ndarray, dtype: str = "f64"):
        ens, m = self._part(member)
        ens.upload(m, f, dtype)

    def download(self, member: int, dtype: str = "f64") -> np.ndarray:
        ens, m = self._part(member)
        return ens.download(m, dtype)

    def moments(self, member: int) -> np.ndarray:
        ens, m = self._part(member)
        return ens.moments(m)

    def checkpoint(self, member: int, real: str = "f64") -> bytes:
        ens, m = self._part(member)
        return ens.checkpoint(m, real)

    def restore(self, member: int, data: bytes, real: str = "f64"):
        ens, m = self._part(member)
        ens.restore(m, data, real)

    def step(self, steps: int = 1, graph: bool = False, stream: int | None = None):
        for st, ens in self.parts.items():
            if self._count[st]:
                ens.step(steps, graph=graph, stream=stream)

    def autotune(self, trial_steps: int = 5) -> dict:
        return {st: ens.autotune(trial_steps) for st, ens in self.parts.items() if self._count[st]}

    def usage(self) -> dict:
        return {st: ens.usage() for st, ens in self.parts.items()}
