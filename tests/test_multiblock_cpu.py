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
        flow = (case.ids != 0)[None] & ~G.streamed_from_outside(case)
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
