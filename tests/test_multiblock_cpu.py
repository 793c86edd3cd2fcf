"""Pin OpenLB's multi-cuboid time step to the oracle (CPU only).

The fixtures under tests/golden/multiblock/ were written by OpenLB run on
several cuboids of one process (CuboidDecomposition::splitRegular, overlap 1,
src/geometry/cuboidDecomposition.hh:755-763): every step collides each block,
refreshes each block's overlap from its neighbours and, across periodic faces,
from the periodic images the cuboid decomposition resolves
(src/core/superLattice.hh:601-663, src/communication/superCommunicator.hh:
189-218, src/geometry/cuboidDecomposition.hh:155-177), then streams.  That
must give the populations of the undivided lattice, which the oracle steps as
one block: the division changes no arithmetic.  The bounds are those of the
single-block pin (tests/test_oracle_golden.py): 1e-12 of max |f_i| for FP64,
2e-5 for an FP32 reference run.

The division itself is checked against Cuboid::divide's rule (the sizes of the
split, src/geometry/cuboid.hh:184-211), which SuperLattice in
openlb_amd/super_lattice.py restates."""
import numpy as np
import pytest

import golden_cases as G
from oracle import oracle as O
from openlb_amd.super_lattice import split_axis

TOL = {"f64": 1e-12, "f32": 2e-5}


@pytest.mark.parametrize("name", G.multiblock_names())
def test_oracle_matches_multicuboid_reference(name):
    case, meta, arrays = G.load_multiblock(name)
    precision = meta["precision"]
    lat = O.OracleLattice(case, precision)
    if meta["start"] == "f_0":
        lat.upload(arrays["f_0"].astype(np.float64))
    else:
        lat.upload(O.initial_populations(case, precision))
    done = 0
    fluid = case.ids == 1
    for s in [s for s in meta["steps"] if s > 0]:
        lat.step(s - done, threads=4)
        done = s
        got = lat.download()
        ref = arrays[f"f_{s}"].astype(np.float64)
        err = G.rel_error(got, ref, case.q, fluid)
        print(f"{name} step {s}: rel {err:.3e}")
        assert err <= TOL[precision], f"{name} step {s}: rel error {err:.3e}"
        # cells with dynamics, except populations that stream in from beyond a
        # non-periodic face (what the overlap there holds is each side's own)
        flow = G.flow_mask(case)
        err_all = G.rel_error(got, ref, case.q, flow)
        assert err_all <= TOL[precision], f"{name} step {s} (cells with dynamics): {err_all:.3e}"


@pytest.mark.parametrize("name", G.multiblock_names())
def test_fixture_division_is_splitregular(name):
    """blocks = max(1, n // width) per axis (splitRegular), and split_axis cuts
    each axis into pieces whose sizes differ by at most one, larger first."""
    case, meta, _ = G.load_multiblock(name)
    width = meta["split_width"]
    assert meta["blocks"] == [max(1, n // width) for n in case.shape]
    for n, b in zip(case.shape, meta["blocks"]):
        pieces = split_axis(n, b)
        sizes = [hi - lo for lo, hi in pieces]
        assert sum(sizes) == n and pieces[0][0] == 0 and pieces[-1][1] == n
        assert max(sizes) - min(sizes) <= 1 and sizes == sorted(sizes, reverse=True)


def _split(case, meta):
    """[(block key, origin, block case)] in OpenLB's cuboid order."""
    import itertools

    from openlb_amd.super_lattice import _sub_case
    ranges = [split_axis(n, b) for n, b in zip(case.shape, meta["blocks"])]
    self_periodic = tuple(int(case.periodic[d] and len(ranges[d]) == 1) for d in range(3))
    out = []
    for key in itertools.product(*(range(len(r)) for r in ranges)):
        lo = tuple(ranges[d][key[d]][0] for d in range(3))
        hi = tuple(ranges[d][key[d]][1] for d in range(3))
        out.append((key, lo, _sub_case(case, lo, hi, self_periodic)))
    return ranges, out


def _with_checkpoint():
    return [n for n in G.multiblock_names() if G.load_multiblock(n)[1]["checkpoint_step"] is not None]


@pytest.mark.parametrize("name", _with_checkpoint())
def test_oracle_restates_openlb_multiblock_checkpoint(name):
    """OpenLB's checkpoint of a split lattice is every block's bytes in cuboid
    order; each block's fields (overlap included: RHO = 1 where the overlap
    copies a wall of a neighbouring cuboid) and its shifts are the oracle's,
    and its core populations are OpenLB's k-step populations."""
    from oracle import checkpoint as K
    case, meta, arrays = G.load_multiblock(name)
    k, real = meta["checkpoint_step"], meta["precision"]
    raw = K.decode_dat(arrays["dat"].tobytes().decode("ascii"))
    _, blocks = _split(case, meta)
    off = 0
    for key, lo, sub in blocks:
        n = K.serializable_size(sub, real, (case, lo))
        chunk = raw[off:off + n]
        off += n
        shifts, pops, _ = K.unpack(sub, chunk, real, (case, lo))
        np.testing.assert_array_equal(shifts, K.rotate_shifts(sub, k))
        assert K.pack(sub, pops, shifts, real, (case, lo)) == chunk, key
        sl = (slice(None),) + tuple(slice(lo[d], lo[d] + sub.shape[d]) for d in range(3))
        np.testing.assert_array_equal(K.core_of(sub, pops), arrays[f"f_{k}"][sl].astype(pops.dtype))
    assert off == len(raw)


@pytest.mark.parametrize("name", _with_checkpoint())
def test_split_pieces_reproduce_openlb_vti(name):
    """One VTK piece per cuboid: the block's own functor values (oracle),
    with the overlap points that copy a bulk cell of another cuboid taking
    that cell's values (piece_sources), give OpenLB's arrays bit for bit."""
    import os
    import tempfile

    from openlb_amd import vtk as V
    from openlb_amd.super_lattice import piece_sources
    case, meta, arrays = G.load_multiblock(name)
    k, real = meta["checkpoint_step"], meta["precision"]
    ranges, blocks = _split(case, meta)
    conv = V.UnitConverter(0.01, 1e-4, 0.1, 1.0, 1e-5, 1000.0, 100.0, real).factors()
    lattices = {}
    for key, lo, sub in blocks:
        lat = O.OracleLattice(sub, real)
        sl = (slice(None),) + tuple(slice(lo[d], lo[d] + sub.shape[d]) for d in range(3))
        lat.upload(np.ascontiguousarray(arrays[f"f_{k}"][sl]).astype(np.float64))
        lattices[key] = lat
    for ic, (key, lo, sub) in enumerate(blocks):
        with tempfile.NamedTemporaryFile("wb", suffix=".vti", delete=False) as fh:
            fh.write(arrays[f"vti_{ic}"].tobytes())
            path = fh.name
        try:
            ref = V.read_vti_arrays(path)
        finally:
            os.unlink(path)
        sources = piece_sources(case, ranges, key)
        for field, fname in enumerate(["density", "velocity", "physVelocity", "physPressure"]):
            own = lattices[key].field_export(field, 1, conv).astype(np.float32)
            dim = 3 if field in (1, 2) else 1
            flat = own.reshape(-1, dim)
            for src_key, (dst, src) in sources.items():
                flat[dst] = lattices[src_key].field_export(field, 0, conv).astype(np.float32).reshape(-1, dim)[src]
            np.testing.assert_array_equal(own, ref[fname], err_msg=f"{key} {fname}")


def _mb_library_names():
    import os
    d = os.path.join(G.MULTIBLOCK, "library")
    return sorted(p[:-4] for p in os.listdir(d) if p.endswith(".npz")) if os.path.isdir(d) else []


def test_library_split_checkpoints_cover_every_fixture():
    assert _mb_library_names() == _with_checkpoint()


@pytest.mark.parametrize("name", _mb_library_names())
def test_openlb_resumes_from_library_split_checkpoint(name):
    """OpenLB, on the same cuboids, loads the split lattice's checkpoint the GPU
    library wrote mid-run (gen_library_checkpoints.py) and continues to the
    fixture's last snapshot: it must reach its own uninterrupted populations
    within the single-block bounds."""
    import json
    import os
    import subprocess
    import sys
    import tempfile

    from oracle import checkpoint as K
    from openlb_amd import checkpoint as CK
    case, meta, arrays = G.load_multiblock(name)
    real = meta["precision"]
    z = np.load(os.path.join(G.MULTIBLOCK, "library", name + ".npz"), allow_pickle=False)
    info = json.loads(str(z["meta"]))
    start, m = info["step"], max(meta["steps"])
    assert info["precision"] == real and info["blocks"] == meta["blocks"] and meta["checkpoint_step"] < start < m
    raw = CK.decode_dat(z["dat"].tobytes().decode("ascii"))
    # every block of the library's file has OpenLB's shifts and fields for that step
    _, blocks = _split(case, meta)
    off = 0
    for key, lo, sub in blocks:
        n = K.serializable_size(sub, real, (case, lo))
        shifts, _, fields = K.unpack(sub, raw[off:off + n], real, (case, lo))
        off += n
        np.testing.assert_array_equal(shifts, K.rotate_shifts(sub, start))
        for f in K.fields(sub, (case, lo)):
            np.testing.assert_array_equal(fields[f[0]], K.field_values(sub, f, (case, lo)).astype(fields[f[0]].dtype))
    assert off == len(raw)
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    exe = os.path.join(root, "oracle", "_ref", f"olb_ref_{real}")
    if not os.path.exists(exe):
        pytest.skip("OpenLB reference binary not built (needs the reference sources)")
    sys.path.insert(0, os.path.join(root, "tests", "golden"))
    import gen_golden
    with tempfile.TemporaryDirectory() as d:
        gen_golden.write_case_dir(d, case, [m - start], split=meta["split_width"])
        CK.write_file(os.path.join(d, CK.parallel_file_name("library")), raw)
        subprocess.run([exe, d], check=True, cwd=d, stdout=subprocess.DEVNULL,
                       env=dict(os.environ, OLB_REF_LOAD=os.path.join(d, "library")))
        got = np.fromfile(os.path.join(d, f"pops_{m - start}.bin")).reshape((case.q,) + tuple(case.shape))
    flow = G.flow_mask(case)
    err = G.rel_error(got, arrays[f"f_{m}"].astype(np.float64), case.q, flow)
    print(f"OpenLB resumed from the library's step-{start} split checkpoint of {name}: rel {err:.2e}", flush=True)
    assert err <= TOL[real], err
