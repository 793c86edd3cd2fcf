"""Host-side checks of the C ABI (no GPU needed). The library loads, exports every
function include/lbmh.h declares, and rejects invalid arguments before touching the
device."""
import ctypes
import os
import re

import numpy as np
import pytest

from openlb_amd import abi

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "lbmh.h")


def declared_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(lbmh_\w+)\s*\(", text, re.M)))


def test_header_and_binding_agree():
    assert declared_functions() == sorted(abi.EXPORTS)


def test_library_exports_every_declared_symbol():
    lib = abi.load()
    for name in declared_functions():
        assert hasattr(lib, name), name


def test_abi_version():
    assert abi.load().lbmh_abi_version() == 1


@pytest.mark.parametrize("field,value", [("q", 20), ("nx", 0), ("storage", 7), ("bulk_kind", 9)])
def test_create_rejects_invalid_descriptor(field, value):
    lib = abi.load()
    desc = abi.LatticeDesc()
    desc.q, desc.nx, desc.ny, desc.nz = 19, 8, 8, 8
    desc.storage, desc.bulk_kind = abi.F32, 0
    setattr(desc, field, value)
    h = ctypes.c_void_p()
    rc = lib.lbmh_lattice_create(ctypes.byref(desc), ctypes.byref(h))
    assert rc == abi.E_INVALID
    assert not h.value
    assert lib.lbmh_last_error().decode()


def test_create_rejects_cell_index_overflow():
    """Padded cell count must stay below 2^31 (32-bit cell ids, OpenLB CellID uint32)."""
    lib = abi.load()
    desc = abi.LatticeDesc()
    desc.q, desc.nx, desc.ny, desc.nz = 19, 1300, 1300, 1300
    h = ctypes.c_void_p()
    assert lib.lbmh_lattice_create(ctypes.byref(desc), ctypes.byref(h)) == abi.E_INVALID
    assert "2^31" in lib.lbmh_last_error().decode()


@pytest.mark.parametrize("storage,n", [("F64", 660), ("F32", 820), ("F16", 1030)])
def test_create_rejects_columns_of_2_gib(storage, n):
    """A population column must stay below 2 GiB: the kernels address it with
    32-bit byte offsets (2^28 padded cells in FP64, 2^29 in FP32, 2^30 in FP16).
    The check comes before any device call."""
    lib = abi.load()
    desc = abi.LatticeDesc()
    desc.q, desc.nx, desc.ny, desc.nz = 19, n, n, n
    desc.storage = getattr(abi, storage)
    h = ctypes.c_void_p()
    assert lib.lbmh_lattice_create(ctypes.byref(desc), ctypes.byref(h)) == abi.E_INVALID
    assert not h.value
    assert "2 GiB" in lib.lbmh_last_error().decode()


def test_null_arguments_are_rejected():
    lib = abi.load()
    assert lib.lbmh_step(None, 1, None) == abi.E_INVALID
    assert lib.lbmh_set_dynamics_ids(None, None) == abi.E_INVALID
    assert lib.lbmh_lattice_destroy(None) == abi.OK
    with pytest.raises(abi.LbmhError):
        abi.check(lib.lbmh_step(None, 1, None))


@pytest.mark.parametrize("geometry", ["cavity", "channel", "step", "spheres", "spheres_closed"])
def test_library_classification_matches_reference(geometry):
    """The library's host-side boundary classification (used by
    lbmh_set_local_boundary) against OpenLB's computeBoundaryTypeAndNormal on
    the same materials.  Host code only: no device is touched."""
    import golden_cases as G
    from openlb_amd.lattice import classify_boundary
    z = np.load(f"{G.GOLDEN}/normals.npz", allow_pickle=False)
    got = classify_boundary(z[f"{geometry}_ids"], tuple(z[f"{geometry}_periodic"]))
    np.testing.assert_array_equal(got, z[f"{geometry}_normals"])


SPLITS_FOR_NORMALS = [(2, 1, 1), (1, 3, 2), (2, 2, 2), (3, 2, 3)]


@pytest.mark.parametrize("blocks", SPLITS_FOR_NORMALS, ids=["x".join(map(str, b)) for b in SPLITS_FOR_NORMALS])
@pytest.mark.parametrize("geometry", ["cavity", "channel", "step", "spheres", "spheres_closed"])
def test_split_block_classification_matches_reference(geometry, blocks):
    """Each block of a split lattice classifies its wall cells with the
    materials of its first overlap layer (lbmh_classify_boundary_overlap, as
    lbmh_set_local_boundary does after lbmh_set_overlap_ids): OpenLB classifies
    on the block geometry, whose overlap holds the neighbours' materials, so
    every block must reproduce OpenLB's normals of the whole domain on its own
    cells.  Without the overlap materials the block faces would read as the
    end of the domain."""
    import golden_cases as G
    from openlb_amd.lattice import classify_boundary
    from openlb_amd.super_lattice import overlap_ids, split_axis
    from openlb_amd import workloads as W

    z = np.load(f"{G.GOLDEN}/normals.npz", allow_pickle=False)
    ids = z[f"{geometry}_ids"]
    periodic = tuple(int(p) for p in z[f"{geometry}_periodic"])
    want = z[f"{geometry}_normals"]
    case = W.Case(geometry, 19, "bgk", ids.shape, periodic, ids, None, 1.0, 0.25, 0.1, (0.0, 0.0, 0.0), {}, {})
    ranges = [split_axis(n, min(b, n)) for n, b in zip(ids.shape, blocks)]
    self_periodic = [int(periodic[d] and len(ranges[d]) == 1) for d in range(3)]
    interfaces = 0
    for rx in ranges[0]:
        for ry in ranges[1]:
            for rz in ranges[2]:
                lo, hi = (rx[0], ry[0], rz[0]), (rx[1], ry[1], rz[1])
                sl = tuple(slice(lo[d], hi[d]) for d in range(3))
                got = classify_boundary(ids[sl], self_periodic, overlap_ids(case, lo, hi))
                np.testing.assert_array_equal(got, want[(slice(None),) + sl], err_msg=f"block {lo}-{hi}")
                alone = classify_boundary(ids[sl], self_periodic)
                interfaces += int(np.any(alone != want[(slice(None),) + sl]))
    if int(np.prod([len(r) for r in ranges])) > 1:
        assert interfaces > 0, "no block of this split has a wall cell on a block face"


def test_split_axis_cuts_as_openlb_divides_a_cuboid():
    """Cuboid::divide (src/geometry/cuboid.hh:184-211) gives piece k of n cells
    in p parts (n + p - k - 1) / p cells: 10 in 3 -> 4, 3, 3; 7 in 2 -> 4, 3."""
    from openlb_amd.super_lattice import split_axis

    assert split_axis(10, 3) == [(0, 4), (4, 7), (7, 10)]
    assert split_axis(7, 2) == [(0, 4), (4, 7)]
    assert split_axis(64, 3) == [(0, 22), (22, 43), (43, 64)]
    for n in (1, 7, 64, 257):
        for parts in range(1, min(n, 6) + 1):
            r = split_axis(n, parts)
            sizes = [b - a for a, b in r]
            assert r[0][0] == 0 and r[-1][1] == n
            assert all(r[k][1] == r[k + 1][0] for k in range(parts - 1))
            assert sizes == sorted(sizes, reverse=True) and max(sizes) - min(sizes) <= 1
    with pytest.raises(ValueError):
        split_axis(3, 4)
