"""Pin the oracle against OpenLB's own CPU path. The vectors in tests/golden/
were written by the compiled reference (tests/golden/gen_golden.py).

Each case starts from OpenLB's populations after ``initialize()`` (snapshot 0),
so the comparison isolates the time step. The time step covers collision, the
overlap-1 halo, PS streaming and the boundary cells. The bound is relative to
max |f_i| of the unshifted populations: 1e-12 for FP64 (the target for the FP64
build) and 2e-5 for an FP32 reference run, where op-order rounding in single
precision accumulates.
"""
import numpy as np
import pytest

import golden_cases as G
from oracle import oracle as O

TOL = {"f64": 1e-12, "f32": 2e-5}


@pytest.mark.parametrize("name", G.names())
def test_oracle_matches_reference(name):
    case, meta, arrays = G.load(name)
    precision = meta["precision"]
    lat = O.OracleLattice(case, precision)
    lat.upload(arrays["f_0"].astype(np.float64))
    done = 0
    fluid = case.ids == 1
    for s in meta["steps"][1:]:
        lat.step(s - done, threads=4)
        done = s
        got = lat.download()
        ref = arrays[f"f_{s}"].astype(np.float64)
        err = G.rel_error(got, ref, case.q, fluid)
        assert err <= TOL[precision], f"{name} step {s}: rel error {err:.3e}"
        # wall cells: OpenLB and the oracle fill the overlap identically here
        err_all = G.rel_error(got, ref, case.q, case.ids != 0)
        assert err_all <= TOL[precision], f"{name} step {s} (all cells): {err_all:.3e}"


# OpenLB accumulates statistics in the lattice's own precision (T) cell by cell:
# in FP32 the running sum over ~3000 cells carries a few 1e-6 of rounding
STATS_TOL = {"f64": 1e-12, "f32": 3e-5}


@pytest.mark.parametrize("name", G.names())
def test_oracle_statistics_match_reference(name):
    """Lattice statistics (cells, average rho, average energy, max |u|) of the
    fixture's last snapshot against what OpenLB reported for it."""
    case, meta, arrays = G.load(name)
    ref = G.statistics(name)
    lat = O.OracleLattice(case, meta["precision"])
    lat.upload(arrays[f"f_{ref['snapshot']}"].astype(np.float64))
    got = lat.statistics()
    err = G.statistics_error(got, ref)
    assert err <= STATS_TOL[meta["precision"]], f"{name}: {err:.3e}"
    if not case.local:
        # the numpy statement of the same definitions, from the per-cell moments
        np.testing.assert_allclose(O.statistics(case, lat.moments()), got,
                                   rtol=1e-12 if meta["precision"] == "f64" else 1e-6)


def test_oracle_statistics_definitions():
    """Hand-built state: two bulk cells, one BounceBack, one NoDynamics and one
    moving wall, so every term of the statistics is known in closed form."""
    from openlb_amd import workloads as W
    ids = np.array([1, 1, 2, 0, 3], dtype=np.uint8).reshape(5, 1, 1)
    case = W.Case("t", 19, "bgk", (5, 1, 1), (1, 1, 1), ids, np.zeros((4, 5, 1, 1)), 1.0, 0.25, 0.0,
                  (0.0, 0.0, 0.0), {3: (1.2, (0.1, 0.0, 0.0))})
    rho_u = np.zeros((4, 5, 1, 1))
    rho_u[0, :, 0, 0] = [0.9, 1.1, 7.0, 9.0, 5.0]
    rho_u[1:, 0, 0, 0] = [0.3, 0.0, 0.4]     # |u|^2 = 0.25
    rho_u[1:, 1, 0, 0] = [0.0, 0.1, 0.0]     # |u|^2 = 0.01
    rho_u[1:, 2, 0, 0] = [5.0, 5.0, 5.0]     # BounceBack: ignored
    got = O.statistics(case, rho_u)
    np.testing.assert_allclose(got, [3, (0.9 + 1.1 + 1.2) / 3, 0.5 * (0.25 + 0.01 + 0.01) / 3, 0.5],
                               rtol=1e-15)


@pytest.mark.parametrize("name", G.names())
def test_oracle_equilibrium_matches_reference_init(name):
    """iniEquilibrium: the oracle's second-order equilibrium against OpenLB's
    populations after initialize(). This includes the F/2 inverse shift on
    forced bulk cells."""
    case, meta, arrays = G.load(name)
    precision = meta["precision"]
    f0 = O.initial_populations(case, precision)
    if case.local:
        # initialize() runs the edge and corner post-processors once
        lat = O.OracleLattice(case, precision)
        lat.upload(f0)
        lat.initialize()
        f0 = lat.download()
    ref = arrays["f_0"].astype(np.float64)
    active = case.ids != 0
    err = G.rel_error(f0, ref, case.q, active)
    # 1-2 ulp: the reference evaluates D3Q19/D3Q27 equilibria with a generated
    # expression order (src/dynamics/lbm.cse.h)
    # the edge/corner reconstruction of initialize() adds a few more ulp
    tol = (4e-16 if precision == "f64" else 2e-7) * (4 if case.local else 1)
    assert err <= tol, f"{name}: {err:.3e}"


@pytest.mark.slow
def test_oracle_tgv64_1000_steps_matches_reference_sample():
    """Config 1 at full size: 64^3 Taylor-Green, FP64, 1000 steps. It is compared with
    a seeded sample of 2048 cells of OpenLB's result and with its per-population sums."""
    meta, arrays = G.load("tgv64_bgk_f64_sample")
    from openlb_amd import workloads as W
    case = W.taylor_green(tuple(meta["shape"]), u0=0.05, tau=0.6)
    assert case.omega == meta["omega"]
    lat = O.OracleLattice(case, "f64")
    lat.upload(O.initial_populations(case, "f64"))
    lat.step(1000)
    got = lat.download().reshape(19, -1)
    idx = arrays["idx"]
    err = G.rel_error(got[:, idx], arrays["f_1000"], 19)
    assert err <= 1e-12, f"sampled rel error {err:.3e}"
    colsum = got.sum(axis=1)
    scale = arrays["colabs_1000"].max()
    assert np.abs(colsum - arrays["colsum_1000"]).max() <= 1e-12 * scale


NORMAL_GEOMETRIES = ["cavity", "channel", "step", "spheres", "spheres_closed"]


@pytest.mark.parametrize("geometry", NORMAL_GEOMETRIES)
def test_oracle_classification_matches_reference(geometry):
    """Boundary type and discrete normal of every wall cell against OpenLB's own
    computeBoundaryTypeAndNormal on the same materials: box faces, edges and
    corners, a re-entrant step (internal edges and corners), sphere packings
    with a periodic axis (irregular patterns, unmatched cells)."""
    z = np.load(f"{G.GOLDEN}/normals.npz", allow_pickle=False)
    ids, ref, per = z[f"{geometry}_ids"], z[f"{geometry}_normals"], z[f"{geometry}_periodic"]
    got = O.classify(ids, per)
    np.testing.assert_array_equal(got, ref)
    # every type occurs somewhere across the geometries; the step has them all
    if geometry == "step":
        assert set(np.unique(ref[0])) >= {-1, 0, 1, 2, 3, 4}
