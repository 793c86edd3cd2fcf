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


@pytest.mark.parametrize("name", G.multiblock_names())
def test_split_lattice_matches_openlb_multicuboid(gpu_available, name):
    """SuperLattice on OpenLB's own division reaches OpenLB's populations."""
    case, meta, arrays = G.load_multiblock(name)
    precision = meta["precision"]
    f0 = arrays["f_0"].astype(np.float64) if meta["start"] == "f_0" else O.initial_populations(case, precision)
    flow = (case.ids != 0)[None] & ~G.streamed_from_outside(case)
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
