"""Several blocks of one lattice on one GPU, with the overlap exchange between
them: the single-device part of OpenLB's SuperLattice.

OpenLB splits a domain into cuboids (``CuboidDecomposition``) and gives each
block an overlap layer.  Each time step collides every block, copies the
overlap cells each block reads from its neighbours (and, across periodic
faces, from the periodic image), then streams every block
(``src/core/superLattice.hh:601-663``; copy lists from
``src/communication/blockCommunicationNeighborhood.hh:89-107``, periodic images
resolved through ``src/geometry/cuboidDecomposition.hh:155-177``).

Here every block is a ``BlockLattice`` with its own PS columns.  A block that
spans a whole periodic axis keeps the fused periodic refresh of its step
kernel on that axis; every other overlap cell is filled by an
``lbmh_halo_copy`` from the block that owns the cell.  The arithmetic per cell
is the same as in a single block, so the result is bit-identical to the
undivided lattice.  Over several GPUs the same lists become the send/receive
buffers of the exchange.

Local boundaries: each block gets the materials of its first overlap layer
(``lbmh_set_overlap_ids``) before its boundaries are set, so it classifies its
wall cells as OpenLB does on the block geometry, and its post-processors read
neighbours beyond a block face from the block that owns them.  Joined blocks
run their post-processors together, stage by stage, once the last of them has
streamed, which keeps the split lattice bit-identical to the undivided one;
OpenLB's own split runs (overlap 3) give the same populations.
"""
from __future__ import annotations

import ctypes
import itertools

import numpy as np

from . import abi, checkpoint
from .lattice import BlockLattice


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


def split_axis(n: int, parts: int) -> list[tuple[int, int]]:
    """[begin, end) ranges of the `parts` pieces OpenLB's ``Cuboid::divide`` cuts
    0..n-1 into: piece k has (n + parts - k - 1) // parts cells, so the pieces
    differ by at most one cell and the larger ones come first
    (``src/geometry/cuboid.hh:184-211``)."""
    if parts < 1 or parts > n:
        raise ValueError(f"cannot split {n} cells into {parts} blocks")
    out, begin = [], 0
    for k in range(parts):
        size = (n + parts - k - 1) // parts
        out.append((begin, begin + size))
        begin += size
    return out


class SuperLattice:
    """A case split into blocks[0] x blocks[1] x blocks[2] cuboids on one device:
    the cuboids, and their order (x outermost, z innermost), of OpenLB's
    ``Cuboid::divide({blocks})`` (``src/geometry/cuboid.hh:184-211``), which
    ``CuboidDecomposition::splitRegular`` uses."""

    def __init__(self, case, storage: str = "f32", blocks=(2, 1, 1), device: int = 0):
        self.case = case
        self.storage = storage
        self.shape = tuple(case.shape)
        self.q = case.q
        self.ranges = [split_axis(n, b) for n, b in zip(self.shape, blocks)]
        self.blocks: dict[tuple[int, int, int], BlockLattice] = {}
        self.origin: dict[tuple[int, int, int], tuple[int, int, int]] = {}
        self.halos: list = []
        # a block keeps its fused periodic refresh on an axis only where it
        # spans the whole periodic axis (its own periodic image)
        self_periodic = tuple(int(case.periodic[d] and len(self.ranges[d]) == 1) for d in range(3))
        for key in itertools.product(*(range(len(r)) for r in self.ranges)):
            lo = tuple(self.ranges[d][key[d]][0] for d in range(3))
            hi = tuple(self.ranges[d][key[d]][1] for d in range(3))
            sub = _sub_case(case, lo, hi, self_periodic)
            ring = overlap_ids(case, lo, hi)
            self.blocks[key] = BlockLattice.from_case(sub, storage, device, overlap_ids=ring)
            self.origin[key] = lo
        self._build_halos(self_periodic)
        self._halo_table = (ctypes.c_void_p * max(1, len(self.halos)))(*[h.value for h in self.halos])

    # -- decomposition ---------------------------------------------------------

    def _owner(self, g):
        """Block key and local coordinates of global cell g (already wrapped)."""
        key, local = [], []
        for d in range(3):
            for k, (a, b) in enumerate(self.ranges[d]):
                if a <= g[d] < b:
                    key.append(k)
                    local.append(g[d] - a)
                    break
        return tuple(key), tuple(local)

    def _build_halos(self, self_periodic):
        lib = abi.load()
        for key, lat in self.blocks.items():
            lo = self.origin[key]
            n = lat.shape
            sy, sz = n[1] + 2, n[2] + 2
            pairs: dict[tuple, tuple[list, list]] = {}
            for x in range(-1, n[0] + 1):
                for y in range(-1, n[1] + 1):
                    for z in range(-1, n[2] + 1):
                        r = (x, y, z)
                        out = [d for d in range(3) if r[d] < 0 or r[d] >= n[d]]
                        if not out:
                            continue
                        # the step kernel's periodic refresh covers overlap
                        # cells whose every out-of-range axis is self-periodic
                        if all(self_periodic[d] for d in out):
                            continue
                        g = [lo[d] + r[d] for d in range(3)]
                        ok = True
                        for d in range(3):
                            if g[d] < 0 or g[d] >= self.shape[d]:
                                if not self.case.periodic[d]:
                                    ok = False
                                    break
                                g[d] %= self.shape[d]
                        if not ok:
                            continue
                        src_key, s = self._owner(g)
                        sn = self.blocks[src_key].shape
                        dst_cell = ((x + 1) * sy + (y + 1)) * sz + (z + 1)
                        src_cell = ((s[0] + 1) * (sn[1] + 2) + (s[1] + 1)) * (sn[2] + 2) + (s[2] + 1)
                        lists = pairs.setdefault(src_key, ([], []))
                        lists[0].append(dst_cell)
                        lists[1].append(src_cell)
            for src_key, (dst_cells, src_cells) in pairs.items():
                d = np.asarray(dst_cells, dtype=np.uint32)
                s = np.asarray(src_cells, dtype=np.uint32)
                h = ctypes.c_void_p()
                abi.check(lib.lbmh_halo_create(lat.handle, self.blocks[src_key].handle, _ptr(d), _ptr(s),
                                               len(d), ctypes.byref(h)))
                self.halos.append(h)

    # -- time step -------------------------------------------------------------

    def collide_and_stream(self, steps: int = 1):
        lib = abi.load()
        for _ in range(int(steps)):
            for lat in self.blocks.values():
                abi.check(lib.lbmh_collide(lat.handle, None))
            if self.halos:
                # every block's overlap copies in one launch
                abi.check(lib.lbmh_halo_copy_many(self._halo_table, len(self.halos), None))
            for lat in self.blocks.values():
                abi.check(lib.lbmh_stream(lat.handle, None))

    def initialize(self):
        """SuperLattice::initialize(): every block's post-processors once on the
        current populations (src/core/superLattice.hh:191-204); blocks whose
        post-processors read other blocks take those neighbours as they stand
        before any block's post-processors ran."""
        for lat in self.blocks.values():
            lat.initialize()

    # -- data ------------------------------------------------------------------

    def _scatter(self, full: np.ndarray, fn):
        for key, lat in self.blocks.items():
            lo = self.origin[key]
            sl = tuple(slice(lo[d], lo[d] + lat.shape[d]) for d in range(3))
            fn(lat, np.ascontiguousarray(full[(slice(None),) + sl]))

    def upload_populations(self, f: np.ndarray, dtype: str = "f64"):
        self._scatter(np.asarray(f), lambda lat, part: lat.upload_populations(part, dtype))

    def ini_equilibrium(self, rho_u: np.ndarray):
        self._scatter(np.asarray(rho_u), lambda lat, part: lat.ini_equilibrium(part))

    def download_populations(self, dtype: str = "f64") -> np.ndarray:
        out = None
        for key, lat in self.blocks.items():
            part = lat.download_populations(dtype)
            if out is None:
                out = np.empty((self.q,) + self.shape, dtype=part.dtype)
            lo = self.origin[key]
            out[(slice(None),) + tuple(slice(lo[d], lo[d] + lat.shape[d]) for d in range(3))] = part
        return out

    def synchronize(self):
        next(iter(self.blocks.values())).synchronize()

    # -- checkpoints -----------------------------------------------------------

    def checkpoint(self, real: str = "f64") -> bytes:
        """OpenLB's SuperLattice serialization (Serializable::save(uint8_t*)):
        every block's serialized bytes in load-balancer order, nothing between
        them (src/core/superLattice.hh:816-832, serializer.hh:106-110); on one
        process the load balancer holds the cuboids in their own order.  The
        overlap fields are those OpenLB's set-up leaves, with the walls of the
        neighbouring blocks seen through the halos."""
        return b"".join(lat.checkpoint(real) for lat in self.blocks.values())

    def restore(self, data: bytes, real: str = "f64"):
        """Serializable::load(const uint8_t*) + postLoad over the blocks.  The
        split must match the file's (the byte counts of every block); a block
        that refuses its part leaves every block as it was."""
        data = bytes(data)
        sizes = [lat.checkpoint_size(real) for lat in self.blocks.values()]
        if sum(sizes) != len(data):
            raise ValueError(f"checkpoint holds {len(data)} bytes, the {len(sizes)} blocks take {sum(sizes)}")
        before = [lat.checkpoint(real) for lat in self.blocks.values()]
        done = []
        off = 0
        try:
            for lat, n in zip(self.blocks.values(), sizes):
                lat.restore(data[off:off + n], real)
                done.append(lat)
                off += n
        except BaseException:
            for lat, old in zip(done, before):
                lat.restore(old, real)
            raise

    def save(self, file_name: str, real: str = "f64", enforce_uint: bool = False) -> str:
        """Write `<file_name>_rank0_size1.dat` as OpenLB's SuperLattice::save(fileName) does."""
        path = checkpoint.parallel_file_name(file_name)
        checkpoint.write_file(path, self.checkpoint(real), enforce_uint)
        return path

    def load(self, file_name: str, real: str = "f64", enforce_uint: bool = False):
        self.restore(checkpoint.read_file(checkpoint.parallel_file_name(file_name), enforce_uint), real)

    # -- evaluation ------------------------------------------------------------

    def cuboids(self):
        """[(block key, origin in cells)] in OpenLB's cuboid order."""
        return [(key, self.origin[key]) for key in self.blocks]

    def field_export(self, key, field: int, overlap: int, conversion=None) -> np.ndarray:
        """float32 points of one moment field (lbmh_field_export's fields) over
        cuboid `key` and `overlap` cells around it, in VTK point order, as
        OpenLB's SuperVTMwriter3D writes the cuboid's piece: its writer
        communicates the overlap first, so an overlap cell that copies a bulk
        cell of another cuboid reports that cell's moments; every other overlap
        cell follows the block's own rule (periodic images, else rho = 1, u = 0)."""
        lat = self.blocks[key]
        out = _export(lat, field, overlap, conversion)
        if overlap:
            dim = out.size // int(np.prod([n + 2 * overlap for n in lat.shape]))
            flat = out.reshape(-1, dim)
            cache = {}
            for src_key, (dst, src) in self._pieces()[key].items():
                if src_key not in cache:
                    cache[src_key] = _export(self.blocks[src_key], field, 0, conversion).reshape(-1, dim)
                flat[dst] = cache[src_key][src]
        return out

    def _pieces(self):
        if getattr(self, "_piece_map", None) is None:
            self._piece_map = {key: piece_sources(self.case, self.ranges, key) for key in self.blocks}
        return self._piece_map

    def close(self):
        lib = abi.load()
        for h in self.halos:
            lib.lbmh_halo_destroy(h)
        self.halos = []
        for lat in self.blocks.values():
            lat.close()
        self.blocks = {}

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def _export(lat, field: int, overlap: int, conversion) -> np.ndarray:
    dim = 3 if field in (1, 2) else 1
    points = int(np.prod([n + 2 * overlap for n in lat.shape]))
    out = np.empty(points * dim, dtype=np.float32)
    conv = None if conversion is None else np.ascontiguousarray(conversion, dtype=np.float64)
    abi.check(abi.load().lbmh_field_export(lat.handle, int(field), int(overlap),
                                           None if conv is None else _ptr(conv), _ptr(out), out.size))
    return out


def piece_sources(case, ranges, key):
    """The overlap points of cuboid `key`'s VTK piece (overlap 1) that hold a
    copy of a bulk cell of a cuboid on another axis split, after OpenLB's
    writer communicates the overlap (src/io/superVtmWriter3D.hh,
    SuperLatticeF functors on communicated blocks): {source key: (point
    indices in this piece, point indices in the source cuboid's core export)},
    both in VTK point order (x fastest).  Overlap cells across a face the
    block spans periodically by itself are its own images and not listed."""
    shape = case.shape
    lo = [ranges[d][key[d]][0] for d in range(3)]
    n = [ranges[d][key[d]][1] - lo[d] for d in range(3)]
    self_periodic = [bool(case.periodic[d] and len(ranges[d]) == 1) for d in range(3)]
    ex, ey = n[0] + 2, n[1] + 2
    out: dict = {}
    for x in range(-1, n[0] + 1):
        for y in range(-1, n[1] + 1):
            for z in range(-1, n[2] + 1):
                r = (x, y, z)
                away = [d for d in range(3) if r[d] < 0 or r[d] >= n[d]]
                if not away or all(self_periodic[d] for d in away):
                    continue
                g = [lo[d] + r[d] for d in range(3)]
                ok = True
                for d in range(3):
                    if not 0 <= g[d] < shape[d]:
                        if not case.periodic[d]:
                            ok = False
                            break
                        g[d] %= shape[d]
                if not ok or case.ids[tuple(g)] != 1:
                    continue
                src_key = tuple(next(k for k, (a, b) in enumerate(ranges[d]) if a <= g[d] < b) for d in range(3))
                s = [g[d] - ranges[d][src_key[d]][0] for d in range(3)]
                sn = [ranges[d][src_key[d]][1] - ranges[d][src_key[d]][0] for d in range(3)]
                lists = out.setdefault(src_key, ([], []))
                lists[0].append(((z + 1) * ey + (y + 1)) * ex + (x + 1))
                lists[1].append((s[2] * sn[1] + s[1]) * sn[0] + s[0])
    return {k: (np.asarray(a, dtype=np.int64), np.asarray(b, dtype=np.int64)) for k, (a, b) in out.items()}


def overlap_ids(case, lo, hi) -> np.ndarray:
    """Materials around the block [lo, hi) of a case: the padded map of the
    block's first overlap layer, read from the whole domain (periodic axes
    wrapped, 0 beyond a non-periodic face), as the block geometry's overlap
    holds them for OpenLB's boundary classification
    (src/boundary/setBoundary.h:83-101).  The core entries are the block's own."""
    shape = tuple(case.shape)
    idx = []
    inside = np.ones([hi[d] - lo[d] + 2 for d in range(3)], dtype=bool)
    for d in range(3):
        g = np.arange(lo[d] - 1, hi[d] + 1)
        ok = (g >= 0) & (g < shape[d])
        if case.periodic[d]:
            g = g % shape[d]
            ok[:] = True
        dims = [1, 1, 1]
        dims[d] = len(g)
        inside &= ok.reshape(dims)
        idx.append(np.clip(g, 0, shape[d] - 1))
    ids = np.asarray(case.ids, dtype=np.uint8)[np.ix_(*idx)]
    return np.where(inside, ids, 0).astype(np.uint8)


def _sub_case(case, lo, hi, periodic):
    import dataclasses
    sl = tuple(slice(lo[d], hi[d]) for d in range(3))
    ids = np.ascontiguousarray(case.ids[sl])
    rho_u = None if case.rho_u is None else np.ascontiguousarray(case.rho_u[(slice(None),) + sl])
    ff = getattr(case, "force_field", None)
    ff = None if ff is None else np.ascontiguousarray(ff[(slice(None),) + sl])
    return dataclasses.replace(case, shape=tuple(hi[d] - lo[d] for d in range(3)), periodic=tuple(periodic),
                               ids=ids, rho_u=rho_u, force_field=ff)
