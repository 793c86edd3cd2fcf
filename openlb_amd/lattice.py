"""Python front end of one device block lattice (thin layer over the C ABI).

The names follow OpenLB's block/super-lattice API so that tests read like the
reference's example programs: ``define_dynamics`` (``SuperLattice::defineDynamics``,
``src/core/superLattice.h:214-233``), ``set_parameters`` (``setParameter<OMEGA>``,
``:318-336``), ``ini_equilibrium`` (``:344-352``) and ``collide_and_stream`` (``:380``).
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import abi, checkpoint


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


_NP_DTYPE = {abi.F64: np.float64, abi.F32: np.float32, abi.F16: np.float16}


class BlockLattice:
    """A padded D3Q19/D3Q27 block with periodic-shift population columns in HBM."""

    def __init__(self, q: int, shape, periodic=(0, 0, 0), storage: str = "f32",
                 kind: str = "bgk", device: int = 0):
        L = abi.load()
        self.q = int(q)
        self.shape = tuple(int(n) for n in shape)
        self.periodic = tuple(int(bool(p)) for p in periodic)
        self.storage = storage
        self.kind = kind
        desc = abi.LatticeDesc()
        desc.q = self.q
        desc.nx, desc.ny, desc.nz = self.shape
        for d in range(3):
            desc.periodic[d] = self.periodic[d]
        desc.storage = abi.STORAGE[storage]
        desc.bulk_kind = abi.KINDS[kind]
        desc.device = device
        h = ctypes.c_void_p()
        abi.check(L.lbmh_lattice_create(ctypes.byref(desc), ctypes.byref(h)))
        self._h = h

    @classmethod
    def from_case(cls, case, storage: str = "f32", device: int = 0, overlap_ids=None):
        """The lattice of a case.  ``overlap_ids`` (uint8, shape + 2 per axis)
        are the materials of the first overlap layer of a block of a split
        lattice: its local boundaries are classified with them
        (``lbmh_set_overlap_ids``)."""
        lat = cls(case.q, case.shape, case.periodic, storage, case.kind, device)
        lat.define_dynamics(case.ids)
        if overlap_ids is not None:
            lat.set_overlap_ids(overlap_ids)
        lat.set_parameters(omega=case.omega, magic=case.magic, smago=case.smago, force=case.force)
        if getattr(case, "force_field", None) is not None:
            lat.set_force_field(case.force_field)
        local = getattr(case, "local", {}) or {}
        for wid, (rho, u) in case.walls.items():
            if wid not in local:
                lat.set_wall(int(wid), rho, u)
        for mid, kind in local.items():
            rho, u = case.walls.get(mid, (1.0, (0.0, 0.0, 0.0)))
            lat.set_local_boundary(int(mid), kind, rho, u)
        return lat

    def close(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            abi.load().lbmh_lattice_destroy(h)
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

    @property
    def handle(self):
        return self._h

    @property
    def ncore(self) -> int:
        return int(np.prod(self.shape))

    def define_dynamics(self, ids: np.ndarray):
        ids = np.ascontiguousarray(ids, dtype=np.uint8)
        if ids.shape != self.shape:
            raise ValueError(f"ids shape {ids.shape} != lattice shape {self.shape}")
        abi.check(abi.load().lbmh_set_dynamics_ids(self._h, _ptr(ids)))

    def set_overlap_ids(self, ring: np.ndarray):
        """Materials of the first overlap layer (padded map, core entries
        unused; 0 where the domain ends), before any local boundary."""
        ring = np.ascontiguousarray(ring, dtype=np.uint8)
        want = tuple(n + 2 for n in self.shape)
        if ring.shape != want:
            raise ValueError(f"overlap ids shape {ring.shape} != padded shape {want}")
        abi.check(abi.load().lbmh_set_overlap_ids(self._h, _ptr(ring)))

    def set_local_boundary(self, material: int, kind: str = "velocity", rho: float = 1.0,
                           u=(0.0, 0.0, 0.0)):
        """boundary::set<LocalVelocity | LocalPressure | InterpolatedVelocity |
        InterpolatedPressure>(lattice, geometry, material)
        followed by defineRhoU on the material (src/boundary/setBoundary.h:83-101):
        classify the cells carrying id ``material`` and give each the regularized
        boundary dynamics of its type."""
        kinds = {"velocity": 0, "pressure": 1, "interp_velocity": 2, "interp_pressure": 3}
        uu = np.ascontiguousarray(u, dtype=np.float64)
        abi.check(abi.load().lbmh_set_local_boundary(self._h, int(material), kinds[kind], float(rho),
                                                     _ptr(uu)))

    def initialize(self):
        """SuperLattice::initialize(): the edge and corner post-processors of the
        local boundaries run once (src/core/blockLattice.hh:56-67)."""
        abi.check(abi.load().lbmh_initialize(self._h))

    def set_parameters(self, omega: float, magic: float = 0.25, smago: float = 0.0,
                       force=(0.0, 0.0, 0.0)):
        p = abi.BulkParams()
        p.omega, p.magic, p.smago = float(omega), float(magic), float(smago)
        for d in range(3):
            p.force[d] = float(force[d])
        abi.check(abi.load().lbmh_set_bulk_params(self._h, ctypes.byref(p)))

    def set_force_field(self, force):
        """Per-cell FORCE of the forced bulk kinds, shape (3, nx, ny, nz), or None
        for the uniform force of set_parameters (the FORCE field ForcedBGKdynamics
        reads, src/dynamics/forcing.h:37-74)."""
        if force is None:
            abi.check(abi.load().lbmh_set_force_field(self._h, None))
            return
        a = np.ascontiguousarray(force, dtype=np.float64)
        if a.shape != (3,) + self.shape:
            raise ValueError(f"force field shape {a.shape} != {(3,) + self.shape}")
        abi.check(abi.load().lbmh_set_force_field(self._h, _ptr(a)))

    def set_wall(self, wall_id: int, rho: float, u):
        u3 = (ctypes.c_double * 3)(*map(float, u))
        abi.check(abi.load().lbmh_set_wall(self._h, int(wall_id), float(rho), u3))

    def upload_populations(self, f: np.ndarray, dtype: str = "f64"):
        code = abi.STORAGE[dtype]
        a = np.ascontiguousarray(f, dtype=_NP_DTYPE[code]).reshape(self.q, -1)
        if a.shape[1] != self.ncore:
            raise ValueError("population array does not match the lattice")
        abi.check(abi.load().lbmh_populations_upload(self._h, _ptr(a), code))

    def download_populations(self, dtype: str = "f64") -> np.ndarray:
        code = abi.STORAGE[dtype]
        out = np.empty((self.q,) + self.shape, dtype=_NP_DTYPE[code])
        abi.check(abi.load().lbmh_populations_download(self._h, _ptr(out), code))
        return out

    def upload_column(self, i: int, column: np.ndarray, shift: int, padding: int):
        """Population column i from a host block in OpenLB's storage order:
        (nx+2p)(ny+2p)(nz+2p) values, element k the padded cell (k - shift) mod count."""
        column = np.ascontiguousarray(column)
        dtype = {np.dtype(np.float64): abi.F64, np.dtype(np.float32): abi.F32}[column.dtype]
        abi.check(abi.load().lbmh_column_upload(self._h, int(i), _ptr(column), int(shift), int(padding), dtype))

    def download_column(self, i: int, column: np.ndarray, shift: int, padding: int):
        """Write the core cells and the first overlap layer of column i into
        `column` (OpenLB storage order at `shift`); other entries stay as they are."""
        assert column.flags.c_contiguous
        dtype = {np.dtype(np.float64): abi.F64, np.dtype(np.float32): abi.F32}[column.dtype]
        abi.check(abi.load().lbmh_column_download(self._h, int(i), _ptr(column), int(shift), int(padding), dtype))

    def ini_equilibrium(self, rho_u: np.ndarray):
        a = np.ascontiguousarray(rho_u, dtype=np.float64)
        if a.shape != (4,) + self.shape:
            raise ValueError("rho_u must have shape (4, nx, ny, nz)")
        abi.check(abi.load().lbmh_ini_equilibrium(self._h, _ptr(a)))

    def collide_and_stream(self, steps: int = 1, stream: int | None = None):
        abi.check(abi.load().lbmh_step(self._h, int(steps), ctypes.c_void_p(stream or 0)))

    def synchronize(self):
        abi.check(abi.load().lbmh_synchronize(self._h))

    def step_count(self) -> int:
        s = ctypes.c_int64(0)
        abi.check(abi.load().lbmh_get_step_count(self._h, ctypes.byref(s)))
        return s.value

    def shifts(self) -> np.ndarray:
        out = np.zeros(self.q, dtype=np.int64)
        abi.check(abi.load().lbmh_get_shifts(self._h, _ptr(out)))
        return out

    # -- checkpoints (Serializable::save / load, src/core/serializer.h:145-189) --

    def checkpoint(self, real: str = "f64") -> bytes:
        """OpenLB's serialized bytes of this block for value type `real` ("f64"
        for a double lattice, "f32" for float): Serializable::save(uint8_t*)."""
        code = {"f64": abi.F64, "f32": abi.F32}[real]
        n = ctypes.c_size_t(0)
        L = abi.load()
        abi.check(L.lbmh_checkpoint_size(self._h, code, ctypes.byref(n)))
        buf = ctypes.create_string_buffer(n.value)
        abi.check(L.lbmh_checkpoint_save(self._h, code, buf, n.value))
        return buf.raw

    def checkpoint_size(self, real: str = "f64") -> int:
        """Bytes of this block's checkpoint (Serializable::getSerializableSize)."""
        n = ctypes.c_size_t(0)
        abi.check(abi.load().lbmh_checkpoint_size(self._h, {"f64": abi.F64, "f32": abi.F32}[real], ctypes.byref(n)))
        return n.value

    def restore(self, data: bytes, real: str = "f64"):
        """Replace the populations and shifts by a checkpoint's (Serializable::
        load(const uint8_t*) followed by postLoad)."""
        code = {"f64": abi.F64, "f32": abi.F32}[real]
        buf = ctypes.create_string_buffer(bytes(data), len(data))
        abi.check(abi.load().lbmh_checkpoint_load(self._h, code, buf, len(data)))

    def save(self, file_name: str, real: str = "f64", enforce_uint: bool = False) -> str:
        """Write `<file_name>_rank0_size1.dat` as OpenLB's save(fileName) does; returns the path."""
        path = checkpoint.parallel_file_name(file_name)
        checkpoint.write_file(path, self.checkpoint(real), enforce_uint)
        return path

    def load(self, file_name: str, real: str = "f64", enforce_uint: bool = False):
        """Read `<file_name>_rank0_size1.dat`, as OpenLB's load(fileName) does."""
        self.restore(checkpoint.read_file(checkpoint.parallel_file_name(file_name), enforce_uint), real)

    REDUCE_FIELDS = {"density": 0, "velocity": 1, "physVelocity": 2, "physPressure": 3, "force": 4}
    REDUCE_OPS = {"sum": 0, "min": 1, "max": 2}

    def reduce_field(self, field: str = "density", op: str = "sum", materials=None, conversion=None):
        """(values, cells): the sum, minimum or maximum of a cell field over the
        core cells of the `materials` (the ids given to define_dynamics; cells a
        local boundary relabels keep their material; default: every material
        but 0), in
        FP64, as OpenLB's SuperLatticeFieldReductionO with SumO / MinO / MaxO
        (src/core/superLatticeFieldReductionO.h:41-160) or SuperSum3D /
        SuperMin3D / SuperMax3D over a material indicator compute it; the
        average of SuperAverage3D is values / cells.  `conversion` = the
        {velocity, pressure, charPhysPressure} factors of the physical fields
        (vtk.UnitConverter.factors())."""
        mask = None
        if materials is not None:
            mask = np.zeros(256, dtype=np.uint8)
            for m in materials:
                mask[int(m)] = 1
        conv = None if conversion is None else np.ascontiguousarray(conversion, dtype=np.float64)
        out = np.zeros(3)
        cells = ctypes.c_uint64(0)
        abi.check(abi.load().lbmh_reduce_field(self._h, self.REDUCE_FIELDS[field], self.REDUCE_OPS[op],
                                               None if mask is None else _ptr(mask),
                                               None if conv is None else _ptr(conv), _ptr(out),
                                               ctypes.byref(cells)))
        comps = 3 if field in ("velocity", "physVelocity", "force") else 1
        return out[:comps].copy(), int(cells.value)

    def moments(self) -> np.ndarray:
        out = np.empty((4,) + self.shape)
        abi.check(abi.load().lbmh_moments(self._h, _ptr(out)))
        return out

    def reduce_moments(self) -> np.ndarray:
        out = np.zeros(4)
        abi.check(abi.load().lbmh_reduce_moments(self._h, _ptr(out)))
        return out

    def statistics(self) -> dict:
        """OpenLB's lattice statistics on the current populations (src/core/
        latticeStatistics.hh:75-90): cells counted, average rho, average
        energy 0.5*<|u|^2> and max |u| over bulk cells and moving walls."""
        out = np.zeros(4)
        abi.check(abi.load().lbmh_statistics(self._h, _ptr(out)))
        return {"cells": int(out[0]), "average_rho": float(out[1]), "average_energy": float(out[2]),
                "max_u": float(out[3])}

    def set_launch_config(self, lds_bytes: int, xcd_group: int = 0):
        cfg = abi.LaunchConfig(int(lds_bytes), int(xcd_group))
        abi.check(abi.load().lbmh_set_launch_config(self._h, ctypes.byref(cfg)))

    def launch_config(self) -> dict:
        cfg = abi.LaunchConfig()
        abi.check(abi.load().lbmh_get_launch_config(self._h, ctypes.byref(cfg)))
        return {"lds_bytes": cfg.lds_bytes, "xcd_group": cfg.xcd_group}

    def autotune(self, trial_steps: int = 10) -> dict:
        """Pick the fastest launch shape on a scratch copy (populations untouched)."""
        cfg = abi.LaunchConfig()
        abi.check(abi.load().lbmh_autotune(self._h, int(trial_steps), ctypes.byref(cfg)))
        return {"lds_bytes": cfg.lds_bytes, "xcd_group": cfg.xcd_group}

    def footprint(self) -> int:
        b = ctypes.c_size_t(0)
        abi.check(abi.load().lbmh_lattice_footprint(self._h, ctypes.byref(b)))
        return b.value


def classify_boundary(ids: np.ndarray, periodic=(0, 0, 0), overlap_ids=None) -> np.ndarray:
    """Boundary type and discrete normal of every wall cell of an id map, as
    OpenLB's computeBoundaryTypeAndNormal gives them (host-side, no GPU):
    int8 [4][nx][ny][nz] = type (-1 on fluid and outside cells), n_x, n_y, n_z.
    With ``overlap_ids`` (padded map) the map is one block of a split lattice
    and its first overlap layer holds the neighbouring blocks' materials."""
    m = np.ascontiguousarray(ids, dtype=np.uint8)
    ext = np.array(m.shape, dtype=np.int32)
    per = np.array([int(bool(p)) for p in periodic], dtype=np.int32)
    out = np.empty((4,) + m.shape, dtype=np.int8)
    lib = abi.load()
    if overlap_ids is None:
        abi.check(lib.lbmh_classify_boundary(_ptr(ext), _ptr(per), _ptr(m), _ptr(out)))
    else:
        ring = np.ascontiguousarray(overlap_ids, dtype=np.uint8)
        if ring.shape != tuple(n + 2 for n in m.shape):
            raise ValueError(f"overlap ids shape {ring.shape} != padded shape of {m.shape}")
        abi.check(lib.lbmh_classify_boundary_overlap(_ptr(ext), _ptr(per), _ptr(m), _ptr(ring), _ptr(out)))
    return out
