"""ctypes front end of the oracle library (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY. Only tests/, ``__graft_entry__.smoke()`` and
bench.py's ``cpu_baseline`` leg import this module, and only as the checker.
The product path never does. See olb_oracle.h for what it restates.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

STORAGE = {"f64": 0, "f32": 1, "f16": 2}
KINDS = {"bgk": 0, "forced_bgk": 1, "trt": 2, "forced_trt": 3, "smagorinsky": 4}
LOCAL = {"velocity": 0, "pressure": 1, "interp_velocity": 2, "interp_pressure": 3}

_lib = None


def build(force: bool = False) -> str:
    """Compile liboracle.so with the committed Makefile."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE, "-B" if force else "liboracle.so"],
                       check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        vp, i32, i64, dbl = ctypes.c_void_p, ctypes.c_int, ctypes.c_int64, ctypes.c_double
        L.olo_create.restype = vp
        L.olo_create.argtypes = [i32] * 8
        L.olo_destroy.argtypes = [vp]
        L.olo_set_ids.argtypes = [vp, vp]
        L.olo_set_bulk.argtypes = [vp, i32, dbl, dbl, dbl, dbl, dbl, dbl]
        L.olo_set_wall.argtypes = [vp, i32, dbl, dbl, dbl, dbl]
        L.olo_upload.argtypes = [vp, vp]
        L.olo_download.argtypes = [vp, vp]
        L.olo_step.argtypes = [vp, i64, i32]
        L.olo_get_shifts.argtypes = [vp, vp]
        L.olo_moments.argtypes = [vp, vp]
        L.olo_field_export.argtypes = [vp, i32, i32, vp, vp]
        L.olo_equilibrium.argtypes = [i32, i32, i32, i64, vp, dbl, dbl, dbl, vp]
        L.olo_equilibrium_field.argtypes = [i32, i32, i32, i64, vp, dbl, dbl, dbl, vp, vp]
        L.olo_set_force_field.argtypes = [vp, vp]
        L.olo_round_to_storage.argtypes = [i32, i64, vp]
        L.olo_define_local.argtypes = [vp, vp, i32, i32, dbl, dbl, dbl, dbl]
        L.olo_classify.argtypes = [i32] * 6 + [vp, vp]
        L.olo_initialize.argtypes = [vp]
        L.olo_statistics.argtypes = [vp, vp]
        L.olo_last_error.restype = ctypes.c_char_p
        _lib = L
    return _lib


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


class OracleLattice:
    """One padded block with periodic-shift columns, stepped on the CPU."""

    def __init__(self, case, storage: str = "f64"):
        L = lib()
        self.case = case
        self.q = case.q
        self.shape = tuple(case.shape)
        self.storage = storage
        h = L.olo_create(case.q, *self.shape, *case.periodic, STORAGE[storage])
        if not h:
            raise RuntimeError(L.olo_last_error().decode())
        self._h = h
        ids = np.ascontiguousarray(case.ids, dtype=np.uint8)
        L.olo_set_ids(h, _ptr(ids))
        if L.olo_set_bulk(h, KINDS[case.kind], case.omega, case.magic, case.smago,
                          *case.force) != 0:
            raise RuntimeError(L.olo_last_error().decode())
        field = getattr(case, "force_field", None)
        if field is not None:
            self._force = np.ascontiguousarray(field, dtype=np.float64)
            if self._force.shape != (3,) + self.shape:
                raise ValueError("force_field must be [3][nx][ny][nz]")
            if L.olo_set_force_field(h, _ptr(self._force)) != 0:
                raise RuntimeError(L.olo_last_error().decode())
        local = getattr(case, "local", {}) or {}
        for wid, (rho, u) in case.walls.items():
            if wid in local:
                continue
            if L.olo_set_wall(h, int(wid), float(rho), *map(float, u)) != 0:
                raise RuntimeError(L.olo_last_error().decode())
        for mid, kind in local.items():
            rho, u = case.walls.get(mid, (1.0, (0.0, 0.0, 0.0)))
            if L.olo_define_local(h, _ptr(ids), int(mid), LOCAL[kind], float(rho), *map(float, u)) != 0:
                raise RuntimeError(L.olo_last_error().decode())

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().olo_destroy(h)
            self._h = None

    def upload(self, f: np.ndarray):
        f = np.ascontiguousarray(f, dtype=np.float64).reshape(self.q, -1)
        lib().olo_upload(self._h, _ptr(f))

    def download(self) -> np.ndarray:
        out = np.empty((self.q,) + self.shape)
        lib().olo_download(self._h, _ptr(out))
        return out

    def step(self, steps: int = 1, threads: int | None = None):
        if threads is None:
            threads = min(16, os.cpu_count() or 1)
        lib().olo_step(self._h, int(steps), int(threads))

    def shifts(self) -> np.ndarray:
        out = np.zeros(self.q, dtype=np.int64)
        lib().olo_get_shifts(self._h, _ptr(out))
        return out

    def moments(self) -> np.ndarray:
        out = np.empty((4,) + self.shape)
        lib().olo_moments(self._h, _ptr(out))
        return out

    def field_export(self, field: int, overlap: int = 1, conversion=None) -> np.ndarray:
        """float32 VTK point buffer of one functor (0 density, 1 velocity,
        2 physVelocity, 3 physPressure), as SuperVTMwriter3D fills it."""
        dim = 3 if field in (1, 2) else 1
        points = int(np.prod([n + 2 * overlap for n in self.shape]))
        out = np.empty(points * dim, dtype=np.float32)
        conv = np.zeros(3) if conversion is None else np.asarray(conversion, dtype=np.float64)
        if lib().olo_field_export(self._h, int(field), int(overlap), _ptr(conv), _ptr(out)) != 0:
            raise RuntimeError("olo_field_export failed")
        return out

    def initialize(self):
        """SuperLattice::initialize(): the post-stream post-processors once."""
        lib().olo_initialize(self._h)

    def statistics(self) -> np.ndarray:
        """{cells, average rho, average energy, max |u|} that the next collision
        reports, boundary cells included (latticeStatistics.hh:75-90)."""
        acc = np.zeros(4)
        lib().olo_statistics(self._h, _ptr(acc))
        n = acc[0]
        if n == 0:
            return np.zeros(4)
        return np.array([n, acc[1] / n, 0.5 * acc[2] / n, np.sqrt(acc[3])])


def classify(materials: np.ndarray, periodic=(0, 0, 0)) -> np.ndarray:
    """OpenLB's boundary type and discrete normal of every wall cell,
    int8 [4][nx][ny][nz] (type -1 on fluid and outside cells)."""
    m = np.ascontiguousarray(materials, dtype=np.uint8)
    out = np.empty((4,) + m.shape, dtype=np.int8)
    lib().olo_classify(*m.shape, *map(int, periodic), _ptr(m), _ptr(out))
    return out


def equilibrium(case, storage: str = "f64", rho_u: np.ndarray | None = None) -> np.ndarray:
    """Shifted equilibrium populations with iniEquilibrium semantics, (q,)+shape."""
    if rho_u is None:
        rho_u = case.rho_u
    rho_u = np.ascontiguousarray(rho_u, dtype=np.float64)
    n = int(np.prod(rho_u.shape[1:]))
    out = np.empty((case.q,) + rho_u.shape[1:])
    field = getattr(case, "force_field", None)
    if field is not None:
        field = np.ascontiguousarray(field, dtype=np.float64)
        lib().olo_equilibrium_field(case.q, KINDS[case.kind], STORAGE[storage], n, _ptr(rho_u),
                                    *case.force, _ptr(field), _ptr(out))
    else:
        lib().olo_equilibrium(case.q, KINDS[case.kind], STORAGE[storage], n, _ptr(rho_u),
                              *case.force, _ptr(out))
    return out


def round_to_storage(values: np.ndarray, storage: str) -> np.ndarray:
    v = np.array(values, dtype=np.float64, copy=True, order="C")
    lib().olo_round_to_storage(STORAGE[storage], v.size, _ptr(v))
    return v


def initial_populations(case, storage: str = "f64") -> np.ndarray:
    """Equilibrium at the case's moments on cells with dynamics; zero elsewhere
    (the zero-initialised column of the reference)."""
    f = equilibrium(case, storage)
    if case.kind.startswith("forced"):
        # only the bulk (forced) dynamics inverse-shift u; walls use plain momenta
        import dataclasses
        plain = equilibrium(dataclasses.replace(case, kind="bgk"), storage)
        f = np.where((case.ids == 1)[None], f, plain)
    f[:, case.ids == 0] = 0.0
    return f


def statistics(case, rho_u: np.ndarray) -> np.ndarray:
    """OpenLB's lattice statistics of one state: {cells, average rho, average
    energy 0.5*<|u|^2>, max |u|} (LatticeStatistics::Aggregatable and reset(),
    src/core/latticeStatistics.h:49-71, latticeStatistics.hh:75-90).

    ``rho_u`` holds the per-cell moments as OracleLattice.moments() returns
    them (bulk u with the F/2 shift when forced).  Bulk cells report their
    own rho and |u|^2; a moving wall reports the rho and velocity it was
    defined with (src/dynamics/collision.h:343-345); NoDynamics and
    BounceBack report nothing (collision.h:52, 309).  Sums in FP64."""
    ids = case.ids
    bulk = ids == 1
    rho = [rho_u[0][bulk]]
    usq = [(rho_u[1:, bulk] ** 2).sum(axis=0)]
    for wid, (wrho, wu) in case.walls.items():
        n = int((ids == wid).sum())
        rho.append(np.full(n, float(wrho)))
        usq.append(np.full(n, float(np.dot(wu, wu))))
    rho = np.concatenate(rho)
    usq = np.concatenate(usq)
    n = rho.size
    if n == 0:
        return np.zeros(4)
    return np.array([n, rho.sum() / n, 0.5 * usq.sum() / n, np.sqrt(usq.max())])
