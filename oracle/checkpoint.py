"""Restatement of OpenLB's checkpoint layout for one block lattice.

TEST INFRASTRUCTURE ONLY. Only tests/ import this module, as the checker of
the library's lbmh_checkpoint_* entry points. The product path never does.

What it restates (OpenLB 1.8 sources):
  * the byte stream of Serializable::save(std::uint8_t*): every block of the
    serializable, in order, with nothing between them
    (src/core/serializer.hh:106-110, src/io/serializerIO.hh:96-106);
  * a ConcreteBlockLattice's blocks: one entry per field marked for
    serialization, in the order of a std::map keyed by typeid(Array<FIELD>)
    .name() (src/core/data.h:516-526, data.hh:52-62), each field D columns
    (src/core/columnVector.h:155-176);
  * a CyclicColumn's blocks: ptrdiff_t shift, size_t count, T data[count]
    (src/core/platform/cpu/sisd/column.hh:66-89), element (n + shift) mod count
    holding logical cell n (sisd/column.h:154-186); rotate(offset) subtracts
    the neighbour distance and wraps into (-count, count) (sisd/column.h:193-205);
  * a Column's blocks: size_t count, T data[count] (sisd/column.hh:38-63);
  * the fields the dynamics of the reference driver's set-up allocate, and the
    values defineRhoU / the momenta's initialize() give them
    (src/dynamics/momenta/elements.h:237-262, 528-571, 578-634;
    dynamics/dynamics.h:473-483, 550-555; momenta/aliases.h:68-73, 126-171).

The field list and the per-cell values are pinned by OpenLB's own checkpoints
(tests/golden/checkpoints/), which this module must reproduce byte for byte
on every cell, overlap included.
"""
from __future__ import annotations

import base64
import struct

import numpy as np

from . import oracle

# discrete velocities of D3Q19 and D3Q27 in OpenLB's order
# (src/descriptor/definition/common.h:120-166, 313-366)
VELOCITIES = {
    19: [(0, 0, 0), (-1, 0, 0), (0, -1, 0), (0, 0, -1), (-1, -1, 0), (-1, 1, 0), (-1, 0, -1),
         (-1, 0, 1), (0, -1, -1), (0, -1, 1), (1, 0, 0), (0, 1, 0), (0, 0, 1), (1, 1, 0),
         (1, -1, 0), (1, 0, 1), (1, 0, -1), (0, 1, 1), (0, 1, -1)],
    27: [(0, 0, 0), (-1, 0, 0), (0, -1, 0), (0, 0, -1), (-1, -1, 0), (-1, 1, 0), (-1, 0, -1),
         (-1, 0, 1), (0, -1, -1), (0, -1, 1), (-1, -1, -1), (-1, -1, 1), (-1, 1, -1), (-1, 1, 1),
         (1, 0, 0), (0, 1, 0), (0, 0, 1), (1, 1, 0), (1, -1, 0), (1, 0, 1), (1, 0, -1),
         (0, 1, 1), (0, 1, -1), (1, 1, 1), (1, 1, -1), (1, -1, 1), (1, -1, -1)],
}

_PRE = "N3olb5ArrayINS_"


def real_dtype(real: str):
    return {"f64": np.float64, "f32": np.float32}[real]


def padded_shape(case):
    return tuple(int(n) + 2 for n in case.shape)


def neighbor_distance(case, c):
    P = padded_shape(case)
    return int(c[0]) * P[1] * P[2] + int(c[1]) * P[2] + int(c[2])


def rotate_shifts(case, steps: int) -> np.ndarray:
    """OpenLB's signed CyclicColumn shifts after `steps` rotations from zero,
    by the loop of CyclicColumn::rotate."""
    n = int(np.prod(padded_shape(case)))
    out = np.zeros(case.q, dtype=np.int64)
    for i, c in enumerate(VELOCITIES[case.q]):
        off = neighbor_distance(case, c)
        s = 0
        for _ in range(int(steps)):
            s -= off
            if s >= n:
                s -= n
            elif s <= -n:
                s += n
        out[i] = s
    return out


def _mangled_int(v: int) -> str:
    return f"n{-v}" if v < 0 else str(v)


def cell_roles(case):
    """Per core cell: role name and, for pressure faces, (direction, orientation).

    Roles: none, bulk, bounce_back, moving_wall, velocity, pressure."""
    role = np.full(case.shape, "none", dtype=object)
    face = np.zeros((2,) + tuple(case.shape), dtype=np.int64)
    role[case.ids == 1] = "bulk"
    cls = oracle.classify(case.ids, case.periodic) if case.local else None
    for m in np.unique(case.ids):
        m = int(m)
        if m in (0, 1):
            continue
        sel = case.ids == m
        kind = case.local.get(m)
        if kind is None:
            role[sel] = "bounce_back" if m == 2 else "moving_wall"
            continue
        for x, y, z in zip(*np.nonzero(sel)):
            t = int(cls[0, x, y, z])
            nrm = [int(v) for v in cls[1:, x, y, z]]
            flat = t == 0 and any(nrm)
            if kind in ("pressure", "interp_pressure"):
                if not flat:
                    role[x, y, z] = "none"   # interpolatedPressure3D.h: no dynamics off flat faces
                    continue
                d = next(a for a in range(3) if nrm[a] != 0)
                role[x, y, z] = "pressure"
                face[:, x, y, z] = (d, nrm[d])
            elif t == 0 and not flat:
                role[x, y, z] = "none"       # flat type without a normal: no dynamics
            else:
                role[x, y, z] = "velocity"
    return role, face


def fields(case, within=None):
    """[(typeid name, components, which, (dir, orient))] in OpenLB's order,
    POPULATION excluded (it sorts first).  `within` = (whole case, origin of
    this block in it) for one cuboid of a split lattice: OpenLB defines
    dynamics on the overlap too, so walls of the neighbouring cuboids seen in
    the overlap allocate their fields in this block as well."""
    role, face = cell_roles(case)
    present = set(np.unique(role))
    if within is not None:
        img = _image_roles(case, role, within)
        present |= set(np.unique(img)) & {"bounce_back", "moving_wall"}
    out = []
    if case.kind in ("forced_bgk", "forced_trt"):
        out.append((_PRE + "11descriptors5FORCEEEE", 3, "force", None))
    if "moving_wall" in present:
        out.append((_PRE + "11descriptors8VELOCITYEEE", 3, "wall_velocity", None))
    if present & {"bounce_back", "moving_wall", "pressure"}:
        out.append((_PRE + "7momenta12FixedDensity3RHOEEE", 1, "rho", None))
    sel = role == "pressure"
    faces = sorted({(int(d), int(o)) for d, o in zip(face[0][sel], face[1][sel])})
    for d, o in faces:
        out.append((_PRE + f"7momenta21FixedPressureMomentumILi{_mangled_int(d)}ELi{_mangled_int(o)}EE8VELOCITYEEE",
                    3, "pressure_velocity", (d, o)))
    if "velocity" in present:
        out.append((_PRE + "7momenta28FixedVelocityMomentumGeneric8VELOCITYEEE", 3, "velocity", None))
    out.sort(key=lambda f: f[0])
    return out


def _image_roles(case, role, within=None):
    """Role of the core cell every padded cell is, or is a periodic image of.
    With `within` = (whole case, origin of this block in it), of the cell of
    the whole lattice every padded cell copies (a neighbouring cuboid's cell,
    or a periodic image of the whole lattice)."""
    P = padded_shape(case)
    lo = (0, 0, 0)
    if within is not None:
        case, lo = within
        role, _ = cell_roles(case)
    out = np.full(P, "none", dtype=object)
    idx = []
    for a in range(3):
        g = np.arange(P[a]) - 1 + lo[a]
        ok = (g >= 0) & (g < case.shape[a])
        if case.periodic[a]:
            g = g % case.shape[a]
            ok[:] = True
        idx.append((g, ok))
    for x in range(P[0]):
        gx, okx = idx[0][0][x], idx[0][1][x]
        for y in range(P[1]):
            gy, oky = idx[1][0][y], idx[1][1][y]
            for z in range(P[2]):
                gz, okz = idx[2][0][z], idx[2][1][z]
                if okx and oky and okz:
                    out[x, y, z] = role[gx, gy, gz]
    return out


def field_values(case, field, within=None):
    """float64 [components][padded cells] of one field, as OpenLB's set-up
    leaves it: defineRhoU's values on the core, initialize()'s (RHO = 1) on
    overlap cells that are periodic images of BounceBack / BounceBackVelocity
    cells, zero elsewhere."""
    key, comps, which, fo = field
    role, face = cell_roles(case)
    P = padded_shape(case)
    out = np.zeros((comps,) + P)
    core = (slice(1, -1),) * 3
    ids = case.ids
    walls = case.walls
    vals = np.zeros((comps,) + tuple(case.shape))
    for m in np.unique(ids):
        m = int(m)
        sel = ids == m
        rho, u = walls.get(m, (1.0, (0.0, 0.0, 0.0)))
        if which == "force":
            sel = sel & (role == "bulk")
            for d in range(3):
                vals[d][sel] = case.force[d] if case.force_field is None else case.force_field[d][sel]
        elif which == "wall_velocity":
            sel = sel & (role == "moving_wall")
            for d in range(3):
                vals[d][sel] = u[d]
        elif which == "rho":
            vals[0][sel & (role == "bounce_back")] = 1.0
            vals[0][sel & ((role == "moving_wall") | (role == "pressure"))] = rho
        elif which == "pressure_velocity":
            sel = sel & (role == "pressure") & (face[0] == fo[0]) & (face[1] == fo[1])
            for d in range(3):
                vals[d][sel] = u[d]
        else:
            sel = sel & (role == "velocity")
            for d in range(3):
                vals[d][sel] = u[d]
    for d in range(comps):
        out[d][core] = vals[d]
    if which == "rho":
        img = _image_roles(case, role, within)
        pad = np.ones(P, dtype=bool)
        pad[core] = False
        # defineDynamics covers the overlap; boundary::set skips cells of
        # neighbourhood radius 0, the outermost layer (setBoundary.h:87-88)
        mask = pad & np.isin(img, ["bounce_back", "moving_wall"])
        out[0][mask] = 1.0
    return out.reshape(comps, -1)


def serializable_size(case, real: str, within=None) -> int:
    n = int(np.prod(padded_shape(case)))
    t = np.dtype(real_dtype(real)).itemsize
    cols = sum(f[1] for f in fields(case, within))
    return case.q * (16 + t * n) + cols * (8 + t * n)


def unpack(case, raw: bytes, real: str, within=None):
    """(signed shifts [q], populations [q][padded cells] in logical cell order,
    {field name: [components][padded cells]}) of a checkpoint."""
    dt = real_dtype(real)
    t = np.dtype(dt).itemsize
    n = int(np.prod(padded_shape(case)))
    if len(raw) != serializable_size(case, real, within):
        raise ValueError(f"checkpoint holds {len(raw)} bytes, expected {serializable_size(case, real, within)}")
    off = 0
    shifts = np.zeros(case.q, dtype=np.int64)
    pops = np.empty((case.q, n), dtype=dt)
    for i in range(case.q):
        s, cnt = struct.unpack_from("<qQ", raw, off)
        off += 16
        if cnt != n:
            raise ValueError("population column count")
        data = np.frombuffer(raw, dtype=dt, count=n, offset=off)
        off += t * n
        shifts[i] = s
        pops[i] = data[(np.arange(n) + s) % n]
    out = {}
    for f in fields(case, within):
        cols = np.empty((f[1], n), dtype=dt)
        for d in range(f[1]):
            (cnt,) = struct.unpack_from("<Q", raw, off)
            off += 8
            if cnt != n:
                raise ValueError("field column count")
            cols[d] = np.frombuffer(raw, dtype=dt, count=n, offset=off)
            off += t * n
        out[f[0]] = cols
    return shifts, pops, out


def pack(case, pops_padded: np.ndarray, shifts: np.ndarray, real: str, within=None) -> bytes:
    """Checkpoint bytes from populations [q][padded cells] in logical order and
    signed shifts; the other fields from the set-up."""
    dt = real_dtype(real)
    n = int(np.prod(padded_shape(case)))
    parts = []
    for i in range(case.q):
        s = int(shifts[i])
        data = np.empty(n, dtype=dt)
        data[(np.arange(n) + s) % n] = np.asarray(pops_padded[i], dtype=dt)
        parts.append(struct.pack("<qQ", s, n))
        parts.append(data.tobytes())
    for f in fields(case, within):
        vals = field_values(case, f, within)
        for d in range(f[1]):
            parts.append(struct.pack("<Q", n))
            parts.append(vals[d].astype(dt).tobytes())
    return b"".join(parts)


def core_of(case, padded: np.ndarray) -> np.ndarray:
    """[k][padded cells] -> [k][nx][ny][nz]."""
    P = padded_shape(case)
    a = np.asarray(padded).reshape((-1,) + P)
    return a[:, 1:-1, 1:-1, 1:-1]


def padded_of(case, core: np.ndarray) -> np.ndarray:
    """[k][nx][ny][nz] -> [k][padded cells], zero overlap."""
    P = padded_shape(case)
    core = np.asarray(core)
    out = np.zeros((core.shape[0],) + P, dtype=core.dtype)
    out[:, 1:-1, 1:-1, 1:-1] = core
    return out.reshape(core.shape[0], -1)


def decode_dat(text: str, enforce_uint: bool = False) -> bytes:
    """Bytes of an OpenLB .dat file: base64 of the binary size, then base64 of
    the data, each with its own padding (serializerIO.hh:40-64, base64.hh)."""
    head = 8 if enforce_uint else 12
    size = struct.unpack("<I" if enforce_uint else "<Q", base64.b64decode(text[:head]))[0]
    body = base64.b64decode("".join(text[head:].split()))
    if len(body) < size:
        raise ValueError("truncated checkpoint file")
    return body[:size]
