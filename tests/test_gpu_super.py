"""GPU tests of several blocks of one lattice on one device (openlb_amd/super_lattice.py,
C ABI lbmh_collide / lbmh_halo_copy / lbmh_stream).

Splitting a domain into cuboids changes no arithmetic: every cell runs the same
collision on the same values, and the overlap copies deliver exactly the values
a single block's periodic refresh and streaming would read.  So a split lattice
must be bit-identical to the undivided one, must match the oracle's undivided
run within the single-block bounds, and must reach the populations OpenLB
itself computes on the same cuboids (tests/golden/multiblock/, OpenLB run with
CuboidDecomposition::splitRegular: the overlap exchange of
src/communication/superCommunicator.hh:189-218 with periodic images resolved
through src/geometry/cuboidDecomposition.hh:155-177)."""
import dataclasses

import numpy as np
import pytest

import golden_cases as G
from oracle import oracle as O
from openlb_amd import workloads as W
from openlb_amd.lattice import BlockLattice
from openlb_amd.super_lattice import SuperLattice

pytestmark = pytest.mark.gpu

CASES = [
    ("tgv_periodic", lambda: W.taylor_green((20, 18, 22), u0=0.05, tau=0.6), "f64"),
    ("porous_channel", lambda: W.porous_channel((24, 20, 18), seed=4, force=1e-5, rmin=1, rmax=3), "f32"),
    ("cavity", lambda: W.lid_driven_cavity((17, 15, 19), tau=0.6, lid_u=0.08), "f64"),
    ("channel27_smagorinsky", lambda: W.turbulent_channel((20, 14, 12), tau=0.51, noise=0.2, seed=2), "f32"),
    ("random_periodic_trt", lambda: W.random_periodic((12, 14, 10), seed=9, kind="trt", tau=0.8), "f16"),
]
SPLITS = [(2, 1, 1), (1, 2, 3), (2, 2, 2), (3, 1, 2)]


@pytest.mark.parametrize("blocks", SPLITS, ids=["x".join(map(str, b)) for b in SPLITS])
@pytest.mark.parametrize("label,make,storage", CASES, ids=[c[0] for c in CASES])
def test_split_lattice_equals_single_block(gpu_available, label, make, storage, blocks):
    case = make()
    f0 = O.initial_populations(case, storage)
    steps = 23
    with BlockLattice.from_case(case, storage) as lat:
        lat.upload_populations(f0, "f64")
        lat.collide_and_stream(steps)
        ref = lat.download_populations("f64")
    with SuperLattice(case, storage, blocks) as sup:
        sup.upload_populations(f0, "f64")
        sup.collide_and_stream(steps)
        got = sup.download_populations("f64")
    # wall cells hold values streamed in from outside the domain, which differ
    # with the padded extent and never reach a fluid cell: compare bulk cells
    mask = case.ids == 1
    np.testing.assert_array_equal(got[:, mask], ref[:, mask])
    if storage in ORACLE_TOL:
        # the oracle's undivided run, at the bounds of test_gpu_matches_oracle_ragged
        ora = O.OracleLattice(case, storage)
        ora.upload(f0)
        ora.step(steps, threads=8)
        err = G.rel_error(got, ora.download(), case.q, mask)
        print(f"{label} {storage} split {blocks} vs oracle: rel {err:.3e}", flush=True)
        assert err <= ORACLE_TOL[storage]


ORACLE_TOL = {"f64": 1e-13, "f32": 1e-5}
TOL = {"f64": 1e-12, "f32": 2e-5}

# Local boundaries on split lattices: each block classifies its wall cells
# with its neighbours' materials, and post-processors read the neighbours that
# lie in other blocks as they stand after the stream and before any
# post-processor.  With block interfaces more than two cells from every edge
# and corner, that is what a single block reads too, so the split lattice is
# bit-identical to the undivided one on every population with dynamics behind
# it (flow_mask: not the slots streamed in from beyond the domain or from cells
# without dynamics, whose contents depend on the padded extent).
LOCAL_CASES = [
    ("cavity_local_velocity", lambda: W.local_velocity_cavity((17, 15, 19), tau=0.6, lid_u=0.08), "f64"),
    ("cavity_interp_velocity",
     lambda: W.local_velocity_cavity((16, 14, 18), tau=0.6, lid_u=0.08, boundary="interp_velocity"), "f64"),
    ("channel_interp_periodic_z",
     lambda: W.velocity_channel((26, 12, 14), tau=0.6, inlet="interp_velocity", outlet="interp_pressure",
                                periodic_z=True), "f64"),
    ("channel_local_pressure", lambda: W.velocity_channel((24, 13, 11), tau=0.6), "f32"),
    ("step_local_velocity", lambda: W.step_cavity((14, 12, 16), block=(6, 5, 7), tau=0.6, lid_u=0.05), "f64"),
    ("cavity27_local_velocity", lambda: W.local_velocity_cavity((10, 12, 11), tau=0.7, lid_u=0.1, q=27), "f32"),
    # other bulk dynamics; the boundary cells keep their BGK mixin
    ("channel_interp_smagorinsky",
     lambda: dataclasses.replace(W.velocity_channel((26, 12, 14), tau=0.52, inlet="interp_velocity",
                                                    outlet="interp_pressure"), kind="smagorinsky", smago=0.14),
     "f32"),
    ("cavity_local_forced_trt",
     lambda: dataclasses.replace(W.local_velocity_cavity((16, 14, 15), tau=0.6, lid_u=0.08), kind="forced_trt",
                                 magic=0.25, force=(1e-5, 0.0, 0.0)), "f64"),
]
LOCAL_SPLITS = [(2, 1, 1), (1, 3, 2), (2, 2, 2)]


@pytest.mark.parametrize("blocks", LOCAL_SPLITS, ids=["x".join(map(str, b)) for b in LOCAL_SPLITS])
@pytest.mark.parametrize("label,make,storage", LOCAL_CASES, ids=[c[0] for c in LOCAL_CASES])
def test_split_lattice_with_local_boundaries_equals_single_block(gpu_available, label, make, storage, blocks):
    case = make()
    f0 = O.initial_populations(case, storage)
    steps = 23
    print(f"{label} {storage} split {blocks}", flush=True)
    with BlockLattice.from_case(case, storage) as lat:
        lat.upload_populations(f0, "f64")
        lat.initialize()
        lat.collide_and_stream(steps)
        ref = lat.download_populations("f64")
    with SuperLattice(case, storage, blocks) as sup:
        sup.upload_populations(f0, "f64")
        sup.initialize()
        sup.collide_and_stream(steps)
        got = sup.download_populations("f64")
    mask = G.flow_mask(case)
    np.testing.assert_array_equal(got[mask], ref[mask])


@pytest.mark.parametrize("boundary", ["velocity", "interp_velocity"])
def test_split_local_boundaries_do_not_depend_on_stream_order(gpu_available, boundary):
    """Blocks whose post-processors read each other give the same populations
    whichever of them lbmh_stream (and lbmh_initialize) reaches first: the
    joined blocks run their post-processors together, stage by stage, once
    the last of them has streamed.  Blocks of 2-3 cells put every corner and
    edge processor next to an interface, reading edges and planes two blocks
    over; the result must equal the undivided lattice's as well."""
    from openlb_amd import abi
    case = W.local_velocity_cavity((8, 9, 10), tau=0.6, lid_u=0.08, boundary=boundary)
    f0 = O.initial_populations(case, "f64")
    lib = abi.load()
    out = []
    for order in (1, -1):
        with SuperLattice(case, "f64", (4, 4, 5)) as sup:
            sup.upload_populations(f0, "f64")
            blocks = list(sup.blocks.values())[::order]
            for lat in blocks:
                lat.initialize()
            for _ in range(17):
                for lat in blocks:
                    abi.check(lib.lbmh_collide(lat.handle, None))
                for h in sup.halos:
                    abi.check(lib.lbmh_halo_copy(h, None))
                for lat in blocks:
                    abi.check(lib.lbmh_stream(lat.handle, None))
            out.append(sup.download_populations("f64"))
    with BlockLattice.from_case(case, "f64") as lat:
        lat.upload_populations(f0, "f64")
        lat.initialize()
        lat.collide_and_stream(17)
        ref = lat.download_populations("f64")
    mask = G.flow_mask(case)
    np.testing.assert_array_equal(out[0][mask], out[1][mask])
    np.testing.assert_array_equal(out[0][mask], ref[mask])


def _local_fixtures():
    return [n for n in G.multiblock_names() if G.load_multiblock(n)[0].local]


@pytest.mark.parametrize("name", _local_fixtures())
def test_split_initialize_matches_openlb(gpu_available, name):
    """SuperLattice::initialize() on OpenLB's cuboids: every block's
    post-processors on the equilibrium state, the ones at block interfaces
    reading their neighbours in the other blocks, reach OpenLB's state after
    initialize() (its f_0)."""
    case, meta, arrays = G.load_multiblock(name)
    precision = meta["precision"]
    with SuperLattice(case, precision, tuple(meta["blocks"])) as sup:
        sup.upload_populations(O.initial_populations(case, precision), "f64")
        sup.initialize()
        got = sup.download_populations("f64")
    err = G.rel_error(got, arrays["f_0"].astype(np.float64), case.q, G.flow_mask(case))
    print(f"{name} initialize: rel {err:.3e}", flush=True)
    assert err <= TOL[precision]


@pytest.mark.parametrize("name", G.multiblock_names())
def test_split_lattice_matches_openlb_multicuboid(gpu_available, name):
    """SuperLattice on OpenLB's own division reaches OpenLB's populations."""
    case, meta, arrays = G.load_multiblock(name)
    precision = meta["precision"]
    f0 = arrays["f_0"].astype(np.float64) if meta["start"] == "f_0" else O.initial_populations(case, precision)
    flow = G.flow_mask(case)
    with SuperLattice(case, precision, tuple(meta["blocks"])) as sup:
        assert len(sup.blocks) == int(np.prod(meta["blocks"]))
        sup.upload_populations(f0, "f64")
        done = 0
        for s in [s for s in meta["steps"] if s > 0]:
            sup.collide_and_stream(s - done)
            done = s
            got = sup.download_populations("f64")
            ref = arrays[f"f_{s}"].astype(np.float64)
            err = G.rel_error(got, ref, case.q, case.ids == 1)
            err_flow = G.rel_error(got, ref, case.q, flow)
            print(f"{name} blocks {meta['blocks']} step {s}: rel {err:.3e} (cells with dynamics {err_flow:.3e})",
                  flush=True)
            assert err <= TOL[precision]
            assert err_flow <= TOL[precision]


def _with_checkpoint():
    return [n for n in G.multiblock_names() if G.load_multiblock(n)[1]["checkpoint_step"] is not None]


def _decoded(arrays):
    from openlb_amd import checkpoint as CK
    return CK.decode_dat(arrays["dat"].tobytes().decode("ascii"))


@pytest.mark.parametrize("name", _with_checkpoint())
def test_split_lattice_resumes_from_openlb_checkpoint(gpu_available, name):
    """OpenLB's checkpoint of the split lattice (every block, in cuboid order)
    restores OpenLB's k-step populations exactly and continues to OpenLB's
    populations at the next snapshot; loading and saving again gives the file's
    bytes."""
    case, meta, arrays = G.load_multiblock(name)
    k, real = meta["checkpoint_step"], meta["precision"]
    raw = _decoded(arrays)
    flow = G.flow_mask(case)
    with SuperLattice(case, real, tuple(meta["blocks"])) as sup:
        sup.restore(raw, real)
        assert sup.checkpoint(real) == raw
        np.testing.assert_array_equal(sup.download_populations(real), arrays[f"f_{k}"])
        m = min(s for s in meta["steps"] if s > k)
        sup.collide_and_stream(m - k)
        got = sup.download_populations("f64")
    err = G.rel_error(got, arrays[f"f_{m}"].astype(np.float64), case.q, flow)
    print(f"{name}: resumed at {k}, step {m}: rel {err:.3e}", flush=True)
    assert err <= TOL[real]


@pytest.mark.parametrize("name", _with_checkpoint())
def test_split_checkpoint_after_steps_has_openlb_layout(gpu_available, name):
    """Each block's part of a checkpoint written after further steps has
    OpenLB's shifts for that step count, the block's core populations and
    every field OpenLB's set-up writes, overlap included (the oracle's
    restatement, pinned to OpenLB's file by test_multiblock_cpu.py)."""
    from oracle import checkpoint as K
    from openlb_amd.super_lattice import _sub_case, split_axis
    case, meta, arrays = G.load_multiblock(name)
    k, real = meta["checkpoint_step"], meta["precision"]
    with SuperLattice(case, real, tuple(meta["blocks"])) as sup:
        sup.restore(_decoded(arrays), real)
        sup.collide_and_stream(7)
        raw = sup.checkpoint(real)
        cores = {key: lat.download_populations(real) for key, lat in sup.blocks.items()}
        ranges = sup.ranges
    self_periodic = tuple(int(case.periodic[d] and len(ranges[d]) == 1) for d in range(3))
    off = 0
    for key, core in cores.items():
        lo = tuple(ranges[d][key[d]][0] for d in range(3))
        hi = tuple(ranges[d][key[d]][1] for d in range(3))
        sub = _sub_case(case, lo, hi, self_periodic)
        n = K.serializable_size(sub, real, (case, lo))
        shifts, pops, fields = K.unpack(sub, raw[off:off + n], real, (case, lo))
        off += n
        np.testing.assert_array_equal(shifts, K.rotate_shifts(sub, k + 7))
        np.testing.assert_array_equal(K.core_of(sub, pops), core)
        want = K.unpack(sub, K.pack(sub, K.padded_of(sub, core), shifts, real, (case, lo)), real, (case, lo))[2]
        assert fields.keys() == want.keys()
        for f in fields:
            np.testing.assert_array_equal(fields[f], want[f], err_msg=f"{key} {f}")
    assert off == len(raw)
    assert split_axis(case.shape[0], meta["blocks"][0]) == ranges[0]


@pytest.mark.parametrize("name", _with_checkpoint())
def test_split_vtk_pieces_match_openlb(gpu_available, name, tmp_path):
    """SuperVTMwriter3D on the split lattice writes OpenLB's pvd and vtm and one
    piece per cuboid with OpenLB's header and arrays (within one float rounding,
    as the single-block export)."""
    from openlb_amd import vtk as V
    case, meta, arrays = G.load_multiblock(name)
    k, real = meta["checkpoint_step"], meta["precision"]
    conv = V.UnitConverter(0.01, 1e-4, 0.1, 1.0, 1e-5, 1000.0, 100.0, real)
    with SuperLattice(case, real, tuple(meta["blocks"])) as sup:
        sup.restore(_decoded(arrays), real)
        w = V.SuperVTMwriter3D("fields", out_dir=str(tmp_path / "vtkData"))
        for f in (V.SuperLatticeDensity3D(sup), V.SuperLatticeVelocity3D(sup),
                  V.SuperLatticePhysVelocity3D(sup, conv), V.SuperLatticePhysPressure3D(sup, conv)):
            w.add_functor(f)
        w.create_master_file()
        w.write(k)
    out = tmp_path / "vtkData"
    assert (out / "fields.pvd").read_text() == arrays["pvd"].tobytes().decode()
    assert (out / "data" / f"fields_iT{k:07d}.vtm").read_text() == arrays["vtm"].tobytes().decode()
    for ic in range(int(np.prod(meta["blocks"]))):
        path = out / "data" / f"fields_iT{k:07d}iC{ic:05d}.vti"
        ref_text = arrays[f"vti_{ic}"].tobytes().decode()
        head = [ln for ln in path.read_text().splitlines() if ln.startswith("<")]
        assert head == [ln for ln in ref_text.splitlines() if ln.startswith("<")]
        ref_path = tmp_path / f"ref_{ic}.vti"
        ref_path.write_text(ref_text)
        got, ref = V.read_vti_arrays(str(path)), V.read_vti_arrays(str(ref_path))
        for key in ref:
            g, r = got[key].astype(np.float64), ref[key].astype(np.float64)
            tol = np.maximum(np.abs(r) * 2.0 ** -23, 1e-7 * np.abs(r).max())
            assert not (np.abs(g - r) > tol).any(), (ic, key, float(np.abs(g - r).max()))


def test_split_restore_is_all_or_nothing(gpu_available):
    """A checkpoint of the wrong size is refused; one whose last block is
    refused leaves every block as it was, the blocks restored before it
    included."""
    import struct

    from openlb_amd import abi
    case = W.lid_driven_cavity((10, 9, 8), tau=0.6, lid_u=0.05)
    with SuperLattice(case, "f64", (2, 1, 1)) as sup:
        sup.ini_equilibrium(case.rho_u)
        sup.collide_and_stream(3)
        good = sup.checkpoint("f64")
        sup.collide_and_stream(2)
        after = sup.download_populations("f64")
        with pytest.raises(ValueError):
            sup.restore(good[:-8], "f64")
        sizes = [lat.checkpoint_size("f64") for lat in sup.blocks.values()]
        bad = bytearray(good)
        last = sum(sizes[:-1])
        bad[last:last + 8] = struct.pack("<q", 1 << 40)     # the last block's first shift, out of range
        with pytest.raises(abi.LbmhError):
            sup.restore(bytes(bad), "f64")
        np.testing.assert_array_equal(sup.download_populations("f64"), after)
        sup.restore(good, "f64")
        sup.collide_and_stream(2)
        np.testing.assert_array_equal(sup.download_populations("f64"), after)


@pytest.mark.parametrize("label,make,storage", [LOCAL_CASES[1], LOCAL_CASES[3]], ids=[LOCAL_CASES[1][0], LOCAL_CASES[3][0]])
def test_split_local_checkpoint_resumes_bit_identically(gpu_available, label, make, storage):
    """A split lattice with local boundaries saved after some steps and
    restored into a fresh one continues bit-identically (the blocks' bytes in
    cuboid order, their post-processors joined again through the new halos)."""
    case = make()
    f0 = O.initial_populations(case, storage)
    real = "f64" if storage == "f64" else "f32"
    with SuperLattice(case, storage, (2, 2, 1)) as sup:
        sup.upload_populations(f0, "f64")
        sup.initialize()
        sup.collide_and_stream(9)
        data = sup.checkpoint(real)
        sup.collide_and_stream(11)
        ref = sup.download_populations("f64")
    with SuperLattice(case, storage, (2, 2, 1)) as sup:
        sup.restore(data, real)
        sup.collide_and_stream(11)
        got = sup.download_populations("f64")
    mask = G.flow_mask(case)
    np.testing.assert_array_equal(got[mask], ref[mask])
