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
                    self._h, m.value, int(mid), {"velocity": 0, "pressure": 1, "interp_velocity": 2}[kind],
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
