"""VTK output of block lattices in OpenLB's format: SuperVTMwriter3D and the
lattice functors it writes, with the moment fields evaluated on the device.

The files are those OpenLB's writer produces (src/io/superVtmWriter3D.hh):

  <out_dir><name>.pvd                          collection, one DataSet per write
  <out_dir>data/<name>_iT<7 digits>.vtm        multiblock, one Block per cuboid
  <out_dir>data/<name>_iT<7>iC<5 digits>.vti   ImageData piece of one cuboid

Each piece covers the cuboid and `overlap` cells around it (WholeExtent
-overlap .. extent + overlap - 1) and holds one Float32 DataArray per functor,
base64 encoded after a uint32 size prefix, zlib-compressed by default
(compress2 at the default level, with the four-word header of
vtkZLibDataCompressor).  The values come from lbmh_field_export: only the
moment fields leave the device.
"""
from __future__ import annotations

import base64
import ctypes
import math
import os
import struct
import zlib

import numpy as np

from . import abi

DENSITY, VELOCITY, PHYS_VELOCITY, PHYS_PRESSURE = 0, 1, 2, 3


class UnitConverter:
    """OpenLB's UnitConverter(physDeltaX, physDeltaT, charPhysLength,
    charPhysVelocity, physViscosity, physDensity, charPhysPressure)
    (src/core/unitConverter.h:152-173): the conversion factors the physical
    functors use, computed in the lattice's value type."""

    def __init__(self, phys_delta_x, phys_delta_t, char_phys_length, char_phys_velocity, phys_viscosity,
                 phys_density, char_phys_pressure=0.0, real: str = "f64"):
        T = np.float64 if real == "f64" else np.float32
        dx, dt, rho = T(phys_delta_x), T(phys_delta_t), T(phys_density)
        self.conversion_velocity = T(dx / dt)
        # util::pow(T, int) computes in double (std::pow promotes), the rest in T
        mass = T(float(rho) * math.pow(float(dx), 3))
        force = T(T(mass * dx) / T(dt * dt))
        self.conversion_pressure = T(float(force) / math.pow(float(dx), 2))
        self.conversion_viscosity = T(T(dx * dx) / dt)
        self.char_phys_pressure = T(char_phys_pressure)
        self.phys_delta_x = dx
        self.lattice_relaxation_time = T(T(phys_viscosity) / self.conversion_viscosity * T(3) + T(0.5))

    def factors(self) -> np.ndarray:
        return np.array([self.conversion_velocity, self.conversion_pressure, self.char_phys_pressure],
                        dtype=np.float64)


class _Functor:
    field = DENSITY
    dim = 1
    default_name = "density"

    def __init__(self, lattice, converter: UnitConverter | None = None, member: int | None = None):
        """`lattice` is a BlockLattice, a SuperLattice, or an Ensemble with `member`."""
        self.lattice, self.converter, self.member = lattice, converter, member
        self.name = self.default_name

    def pieces(self):
        """[(origin in cells, functor of one cuboid)]: one per cuboid of a
        SuperLattice, in OpenLB's cuboid order; a block is cuboid 0."""
        if self.member is None and hasattr(getattr(self, "lattice", None), "cuboids"):
            return [(lo, _Piece(self, key)) for key, lo in self.lattice.cuboids()]
        return [((0, 0, 0), self)]

    def shape(self):
        if self.member is None:
            return tuple(self.lattice.shape)
        return tuple(self.lattice.shapes[self.member])

    def evaluate(self, overlap: int = 1) -> np.ndarray:
        """float32 [points * dim] in VTK point order over the extended extent."""
        if self.member is None and hasattr(self.lattice, "cuboids"):
            raise ValueError("a SuperLattice functor is evaluated per cuboid: use pieces()")
        n = self.shape()
        points = int(np.prod([v + 2 * overlap for v in n]))
        out = np.empty(points * self.dim, dtype=np.float32)
        conv = None
        if self.converter is not None:
            conv = self.converter.factors()
        cptr = conv.ctypes.data_as(ctypes.c_void_p) if conv is not None else None
        optr = out.ctypes.data_as(ctypes.c_void_p)
        L = abi.load()
        if self.member is None:
            abi.check(L.lbmh_field_export(self.lattice.handle, self.field, int(overlap), cptr, optr, out.size))
        else:
            abi.check(L.lbmh_ensemble_field_export(self.lattice._h, int(self.member), self.field, int(overlap),
                                                   cptr, optr, out.size))
        return out


class _Piece:
    """A functor restricted to one cuboid of a SuperLattice."""

    def __init__(self, functor: _Functor, key):
        self.functor, self.key = functor, key
        self.dim, self.field = functor.dim, functor.field

    @property
    def name(self):
        return self.functor.name

    def shape(self):
        return tuple(self.functor.lattice.blocks[self.key].shape)

    def evaluate(self, overlap: int = 1) -> np.ndarray:
        conv = None if self.functor.converter is None else self.functor.converter.factors()
        return self.functor.lattice.field_export(self.key, self.field, overlap, conv)


class SuperLatticeDensity3D(_Functor):
    field, dim, default_name = DENSITY, 1, "density"


class SuperLatticeVelocity3D(_Functor):
    field, dim, default_name = VELOCITY, 3, "velocity"


class SuperLatticePhysVelocity3D(_Functor):
    field, dim, default_name = PHYS_VELOCITY, 3, "physVelocity"

    def __init__(self, lattice, converter: UnitConverter, member: int | None = None):
        super().__init__(lattice, converter, member)


class SuperLatticePhysPressure3D(_Functor):
    field, dim, default_name = PHYS_PRESSURE, 1, "physPressure"

    def __init__(self, lattice, converter: UnitConverter, member: int | None = None):
        super().__init__(lattice, converter, member)


def create_file_name(name: str, it: int | None = None, ic: int | None = None) -> str:
    """createFileName (src/io/fileName.hh:34-74)."""
    if it is None:
        return name
    out = f"{name}_iT{it:07d}"
    if ic is not None:
        out += f"iC{ic:05d}"
    return out


def _g(v) -> str:
    """A number as std::ostream writes it by default (6 significant digits)."""
    return "%g" % float(v)


def _b64(data: bytes) -> str:
    return base64.b64encode(data).decode("ascii")


class SuperVTMwriter3D:
    """SuperVTMwriter3D<T, float> (src/io/superVtmWriter3D.h:50-144) for the
    blocks of one lattice (one cuboid each)."""

    def __init__(self, name: str, overlap: int = 1, binary: bool = True, compress: bool = True,
                 out_dir: str = "./tmp/vtkData/"):
        self.name, self.overlap, self.binary, self.compress = name, int(overlap), binary, compress
        self.out_dir = out_dir if out_dir.endswith("/") else out_dir + "/"
        os.makedirs(self.out_dir + "data/", exist_ok=True)
        self.functors: list[_Functor] = []

    def add_functor(self, f: _Functor, name: str | None = None):
        if name is not None:
            f.name = name
        self.functors.append(f)

    def clear_added_functors(self):
        self.functors = []

    def create_master_file(self):
        with open(self.out_dir + create_file_name(self.name) + ".pvd", "w") as fh:
            fh.write('<?xml version="1.0"?>\n<VTKFile type="Collection" version="0.1" '
                     'byte_order="LittleEndian">\n<Collection>\n')
            fh.write("</Collection>\n</VTKFile>\n")

    def write(self, it: int, origin=(0.0, 0.0, 0.0), delta: float = 1.0, cuboids=None):
        """One time step: the .pvd entry, the .vtm and a .vti per cuboid.
        `cuboids` lists (functors, origin) per cuboid; by default the added
        functors give one piece per cuboid of their lattice (a SuperLattice's
        cuboids at `origin` + `delta` x their origin in cells, a block as
        cuboid 0 at `origin`)."""
        if not self.functors and not cuboids:
            raise RuntimeError("No functor to write")
        if cuboids is None:
            pieces = [f.pieces() for f in self.functors]
            if len({len(p) for p in pieces}) != 1:
                raise ValueError("the added functors belong to lattices of different cuboids")
            cuboids = [([p[i][1] for p in pieces],
                        tuple(origin[d] + delta * pieces[0][i][0][d] for d in range(3)))
                       for i in range(len(pieces[0]))]
        pvd = self.out_dir + create_file_name(self.name) + ".pvd"
        with open(pvd) as fh:
            text = fh.read()
        text = text[:-25] + (f'<DataSet timestep="{it}" group="" part="" '
                             f'file="data/{create_file_name(self.name, it)}.vtm"/>\n')
        text += "</Collection>\n</VTKFile>\n"
        with open(pvd, "w") as fh:
            fh.write(text)
        vtm = [f'<?xml version="1.0"?>\n<VTKFile type="vtkMultiBlockDataSet" version="1.0" '
               f'byte_order="LittleEndian">\n<vtkMultiBlockDataSet>\n']
        for ic in range(len(cuboids)):
            vtm.append(f'<Block index="{ic}" >\n<DataSet index= "0" file="{create_file_name(self.name, it, ic)}.vti">\n'
                       "</DataSet>\n</Block>\n")
        vtm.append("</vtkMultiBlockDataSet>\n</VTKFile>\n")
        with open(self.out_dir + "data/" + create_file_name(self.name, it) + ".vtm", "w") as fh:
            fh.write("".join(vtm))
        for ic, (functors, org) in enumerate(cuboids):
            self._write_vti(self.out_dir + "data/" + create_file_name(self.name, it, ic) + ".vti", functors, org,
                            delta)

    def _write_vti(self, path, functors, origin, delta):
        o = self.overlap
        n = functors[0].shape()
        ext = " ".join(f"{-o} {n[d] + o - 1}" for d in range(3))
        parts = ['<?xml version="1.0"?>\n<VTKFile type="ImageData" version="0.1" ']
        parts.append('byte_order="LittleEndian" compressor="vtkZLibDataCompressor">\n' if self.compress
                     else 'byte_order="LittleEndian">\n')
        parts.append(f'<ImageData WholeExtent="{ext}" Origin="{_g(origin[0])} {_g(origin[1])} {_g(origin[2])}" '
                     f'Spacing="{_g(delta)} {_g(delta)} {_g(delta)}">\n')
        parts.append(f'<Piece Extent="{ext}">\n<PointData>\n')
        for f in functors:
            name = f.name.replace("<", "_").replace(">", "_")
            parts.append(f'<DataArray type="Float32" Name="{name}" NumberOfComponents="{f.dim}" ')
            parts.append('format="binary" encoding="base64">\n' if (self.compress or self.binary)
                         else 'format="ascii" >\n')
            data = f.evaluate(o)
            raw = data.tobytes()
            if self.compress:
                comp = zlib.compress(raw, -1)
                parts.append(_b64(struct.pack("<4I", 1, len(raw), len(raw), len(comp))))
                parts.append(_b64(comp))
            elif self.binary:
                parts.append(_b64(struct.pack("<I", len(raw))))
                parts.append(_b64(raw))
            else:
                parts.append("".join(_g(v) + " " for v in data))
            parts.append("\n</DataArray>\n")
        parts.append("</PointData>\n</Piece>\n</ImageData>\n</VTKFile>\n")
        with open(path, "w") as fh:
            fh.write("".join(parts))


def read_vti_arrays(path: str) -> dict:
    """{DataArray name: float32 array} of a piece this writer or OpenLB's wrote
    (binary or compressed base64; ascii arrays as text)."""
    with open(path) as fh:
        text = fh.read()
    compressed = 'compressor="vtkZLibDataCompressor"' in text
    out = {}
    pos = 0
    while True:
        i = text.find("<DataArray ", pos)
        if i < 0:
            break
        head_end = text.index(">\n", i)
        head = text[i:head_end]
        name = head.split('Name="')[1].split('"')[0]
        body_end = text.index("\n</DataArray>", head_end)
        body = text[head_end + 2:body_end]
        if 'format="ascii"' in head:
            out[name] = np.array([float(v) for v in body.split()], dtype=np.float32)
        elif compressed:
            words = struct.unpack("<4I", base64.b64decode(body[:24]))
            out[name] = np.frombuffer(zlib.decompress(base64.b64decode(body[24:])), dtype=np.float32)
            assert words[1] == out[name].nbytes
        else:
            size = struct.unpack("<I", base64.b64decode(body[:8]))[0]
            out[name] = np.frombuffer(base64.b64decode(body[8:]), dtype=np.float32)[: size // 4]
        pos = body_end
    return out
