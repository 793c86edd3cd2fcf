"""GPU parity of the HIP time step, called through the C ABI (include/lbmh.h).

The checkers are the golden vectors of the compiled OpenLB reference
(tests/golden/) and the oracle (oracle/liboracle.so, the CPU restatement).
Both are pinned by test_oracle_golden.py.

Tolerances are relative to max |f_i| of the unshifted populations. They are:
  FP64 storage and arithmetic            1e-12 (the FP64 target: 1000 steps vs OpenLB)
  FP32 storage and arithmetic            2e-5 after <= 200 steps vs the FP64 reference
  FP16 storage with FP32 arithmetic      velocity error <= 1e-3 of max |u| vs FP64
Only cells with bulk dynamics are compared. OpenLB's own CPU and GPU platforms
fill wall-cell populations that stream in from outside the domain differently.
"""
import dataclasses

import numpy as np
import pytest

import golden_cases as G
from oracle import oracle as O
from openlb_amd import abi
from openlb_amd import workloads as W
from openlb_amd.ensemble import Ensemble
from openlb_amd.lattice import BlockLattice

pytestmark = pytest.mark.gpu

TOL = {"f64": 1e-12, "f32": 2e-5}


def run_gpu(case, storage, f0, steps):
    with BlockLattice.from_case(case, storage) as lat:
        lat.upload_populations(f0, "f64")
        out = {}
        done = 0
        for s in steps:
            lat.collide_and_stream(s - done)
            done = s
            out[s] = lat.download_populations("f64")
        return out


@pytest.mark.parametrize("name", G.names())
def test_gpu_matches_reference(gpu_available, name):
    case, meta, arrays = G.load(name)
    precision = meta["precision"]
    steps = meta["steps"][1:]
    got = run_gpu(case, precision, arrays["f_0"].astype(np.float64), steps)
    fluid = case.ids == 1
    for s in steps:
        err = G.rel_error(got[s], arrays[f"f_{s}"], case.q, fluid)
        print(f"{name} {precision} step {s}: rel {err:.3e}", flush=True)
        assert err <= TOL[precision], f"{name} step {s}: {err:.3e}"


@pytest.mark.parametrize("name", [n for n in G.names() if n.endswith("_f64")])
def test_gpu_fp32_storage_vs_fp64_reference(gpu_available, name):
    case, meta, arrays = G.load(name)
    steps = [meta["steps"][-1]]
    got = run_gpu(case, "f32", arrays["f_0"].astype(np.float64), steps)
    err = G.rel_error(got[steps[0]], arrays[f"f_{steps[0]}"], case.q, case.ids == 1)
    print(f"{name} f32-storage: rel {err:.3e}", flush=True)
    assert err <= TOL["f32"]


@pytest.mark.parametrize("name", [n for n in G.names() if n.endswith("_f64")])
def test_gpu_fp16_storage_velocity_error(gpu_available, name):
    """FP16 shifted storage, FP32 collision: the velocity against OpenLB's FP64 run."""
    case, meta, arrays = G.load(name)
    s = meta["steps"][-1]
    with BlockLattice.from_case(case, "f16") as lat:
        lat.upload_populations(arrays["f_0"].astype(np.float64), "f64")
        lat.collide_and_stream(s)
        mom = lat.moments()
    ref = O.OracleLattice(case, "f64")
    ref.upload(arrays[f"f_{s}"].astype(np.float64))
    mref = ref.moments()
    fluid = case.ids == 1
    # characteristic velocity: the larger of the initial and the final flow
    # (decaying periodic modes end far below their start, driven flows above)
    u_char = max(np.abs(mref[1:, fluid]).max(), np.abs(case.rho_u[1:, fluid]).max(),
                 max((abs(v) for _, u in case.walls.values() for v in u), default=0.0))
    uerr = np.abs(mom[1:, fluid] - mref[1:, fluid]).max()
    print(f"{name} f16: |du|max/u_char = {uerr / u_char:.3e}", flush=True)
    assert uerr <= 1e-3 * u_char


def test_gpu_tgv64_fp64_1000_steps(gpu_available):
    """Config 1: 64^3 periodic Taylor-Green, FP64, tau 0.6, 1000 steps. It is checked
    against OpenLB (a seeded sample of 2048 cells plus the per-population sums) at 1e-12,
    and against the oracle on every cell."""
    meta, arrays = G.load("tgv64_bgk_f64_sample")
    case = W.taylor_green((64, 64, 64), u0=0.05, tau=0.6)
    f0 = O.initial_populations(case, "f64")
    got = run_gpu(case, "f64", f0, [1000])[1000].reshape(19, -1)
    idx = arrays["idx"]
    err = G.rel_error(got[:, idx], arrays["f_1000"], 19)
    print(f"tgv64 vs OpenLB sample: rel {err:.3e}", flush=True)
    assert err <= 1e-12
    scale = arrays["colabs_1000"].max()
    assert np.abs(got.sum(axis=1) - arrays["colsum_1000"]).max() <= 1e-12 * scale
    ora = O.OracleLattice(case, "f64")
    ora.upload(f0)
    ora.step(1000)
    err_o = G.rel_error(got, ora.download(), 19)
    print(f"tgv64 vs oracle (all cells): rel {err_o:.3e}", flush=True)
    assert err_o <= 1e-12


def test_gpu_turbulent_channel_fp64_matches_oracle_at_size(gpu_available):
    """Config 3's flow (D3Q27 Smagorinsky-BGK, tau 0.505, C 0.14, periodic in x
    and z between BounceBack walls, laminar profile with noise) at 96x48x48
    for 400 steps: the FP64 device path against the oracle on the bulk cells,
    at the 1e-12 bound of the fixture-size pin.  FP32 storage stays within its
    2e-5 population bound of the oracle's FP64 run, and the FP16 velocity
    error against that run is recorded."""
    case = W.turbulent_channel((96, 48, 48), noise=0.2, seed=11)
    f0 = O.initial_populations(case, "f64")
    steps = 400
    fluid = case.ids == 1
    got = run_gpu(case, "f64", f0, [steps])[steps]
    ora = O.OracleLattice(case, "f64")
    ora.upload(f0)
    ora.step(steps, threads=16)
    want = ora.download()
    err = G.rel_error(got, want, 27, fluid)
    print(f"turbulent channel 96x48x48 FP64, {steps} steps vs oracle: rel {err:.3e}", flush=True)
    assert err <= 1e-12
    ref = O.OracleLattice(case, "f64")
    ref.upload(want)
    u_ref = ref.moments()[1:, fluid]
    umax = float(np.abs(u_ref).max())
    for storage in ("f32", "f16"):
        f = run_gpu(case, storage, O.initial_populations(case, storage), [steps])[steps]
        m = O.OracleLattice(case, "f64")
        m.upload(f)
        du = float(np.abs(m.moments()[1:, fluid] - u_ref).max()) / umax
        perr = G.rel_error(f, want, 27, fluid)
        print(f"  {storage}: max |du| / u_max {du:.2e}, populations rel {perr:.2e}", flush=True)
        if storage == "f32":
            assert perr <= 2e-5


RAGGED = [
    ("cell1_periodic", lambda: W.random_periodic((1, 1, 1), seed=1, tau=0.7)),
    ("line_periodic", lambda: W.random_periodic((1, 3, 67), seed=2, tau=0.8)),
    ("odd_periodic27", lambda: W.random_periodic((5, 3, 7), seed=3, q=27, tau=0.65)),
    ("odd_trt", lambda: W.random_periodic((7, 9, 65), seed=4, kind="trt", tau=0.9, magic=0.1)),
    ("cavity_ragged", lambda: W.lid_driven_cavity((9, 7, 70), tau=0.6, lid_u=0.1)),
    ("channel_forced", lambda: W.porous_channel((33, 17, 19), seed=5, tau=0.55, force=1e-5,
                                               rmin=1, rmax=3)),
    ("channel27_smag", lambda: W.turbulent_channel((20, 12, 10), tau=0.505, noise=0.2, seed=6)),
    ("y_periodic_walls", lambda: _walled(W.random_periodic((6, 9, 5), seed=7, tau=0.7), (0, 1, 0))),
    ("xz_periodic27_thin", lambda: _walled(W.random_periodic((3, 6, 2), seed=8, q=27, tau=0.8), (1, 0, 1))),
]


def _walled(case, periodic):
    """The box with the given periodic axes and BounceBack walls on the faces of
    the others: halo pushes of cells that lie on a periodic and a walled face."""
    ids = case.ids.copy()
    for d in range(3):
        if not periodic[d]:
            lo = [slice(None)] * 3
            hi = [slice(None)] * 3
            lo[d], hi[d] = 0, -1
            ids[tuple(lo)] = 2
            ids[tuple(hi)] = 2
    rho_u = case.rho_u.copy()
    rho_u[1:, ids == 2] = 0.0
    return dataclasses.replace(case, periodic=tuple(periodic), ids=ids, rho_u=rho_u)


def _with_holes(case):
    """Make some interior cells NoDynamics (id 0), and add two more moving walls."""
    ids = case.ids.copy()
    rng = np.random.Generator(np.random.PCG64(99))
    ids[rng.random(ids.shape) < 0.02] = 0
    case.ids = ids
    return case


@pytest.mark.parametrize("storage", ["f64", "f32"])
@pytest.mark.parametrize("label,make", RAGGED, ids=[r[0] for r in RAGGED])
def test_gpu_matches_oracle_ragged(gpu_available, label, make, storage):
    case = make()
    if label == "line_periodic":
        case = _with_holes(case)
    f0 = O.initial_populations(case, storage)
    steps = 40
    got = run_gpu(case, storage, f0, [steps])[steps]
    ora = O.OracleLattice(case, storage)
    ora.upload(f0)
    ora.step(steps, threads=8)
    ref = ora.download()
    err = G.rel_error(got, ref, case.q, case.ids == 1)
    print(f"{label} {storage}: rel {err:.3e}", flush=True)
    assert err <= (1e-13 if storage == "f64" else 1e-5)


@pytest.mark.parametrize("storage", ["f64", "f32", "f16"])
def test_upload_download_roundtrip_is_exact(gpu_available, storage):
    case = W.porous_channel((13, 11, 9), seed=3, rmin=1, rmax=2)
    rng = np.random.Generator(np.random.PCG64(5))
    f = O.round_to_storage(rng.normal(scale=1e-2, size=(19,) + case.shape), storage)
    with BlockLattice.from_case(case, storage) as lat:
        lat.upload_populations(f, "f64")
        back = lat.download_populations("f64")
        np.testing.assert_array_equal(back, f)
        lat.collide_and_stream(3)
        np.testing.assert_array_equal(lat.shifts(), O_shift_closed_form(case, 3))


def O_shift_closed_form(case, steps):
    """shift_i(t) = (-t * c_i . projection) mod n_padded (blockLattice.hh:317-339)."""
    sx, sy, sz = (n + 2 for n in case.shape)
    c = np.array(W_C19 if case.q == 19 else W_C27)
    d = c[:, 0] * sy * sz + c[:, 1] * sz + c[:, 2]
    return np.mod(-steps * d, sx * sy * sz)


W_C19 = [(0, 0, 0), (-1, 0, 0), (0, -1, 0), (0, 0, -1), (-1, -1, 0), (-1, 1, 0), (-1, 0, -1),
         (-1, 0, 1), (0, -1, -1), (0, -1, 1), (1, 0, 0), (0, 1, 0), (0, 0, 1), (1, 1, 0),
         (1, -1, 0), (1, 0, 1), (1, 0, -1), (0, 1, 1), (0, 1, -1)]
W_C27 = [(0, 0, 0), (-1, 0, 0), (0, -1, 0), (0, 0, -1), (-1, -1, 0), (-1, 1, 0), (-1, 0, -1),
         (-1, 0, 1), (0, -1, -1), (0, -1, 1), (-1, -1, -1), (-1, -1, 1), (-1, 1, -1), (-1, 1, 1),
         (1, 0, 0), (0, 1, 0), (0, 0, 1), (1, 1, 0), (1, -1, 0), (1, 0, 1), (1, 0, -1), (0, 1, 1),
         (0, 1, -1), (1, 1, 1), (1, 1, -1), (1, -1, 1), (1, -1, -1)]


@pytest.mark.parametrize("storage", ["f64", "f32", "f16"])
def test_ini_equilibrium_and_moments(gpu_available, storage):
    case = W.porous_channel((16, 12, 10), seed=8, force=3e-4, rmin=1, rmax=2)
    rng = np.random.Generator(np.random.PCG64(8))
    case.rho_u[0] = 1.0 + 0.01 * rng.standard_normal(case.shape)
    case.rho_u[1:] = 0.03 * rng.standard_normal((3,) + case.shape)
    with BlockLattice.from_case(case, storage) as lat:
        lat.ini_equilibrium(case.rho_u)
        f = lat.download_populations("f64")
        mom = lat.moments()
        red = lat.reduce_moments()
    ref = O.initial_populations(case, storage)
    tol = {"f64": 1e-15, "f32": 3e-7, "f16": 2e-3}[storage]
    assert G.rel_error(f, ref, 19, case.ids != 0) <= tol
    fluid = case.ids == 1
    # the moments reproduce the requested rho/u on bulk cells (forced: u incl. F/2)
    mtol = {"f64": 1e-13, "f32": 1e-5, "f16": 5e-3}[storage]
    assert np.abs(mom[0][fluid] - case.rho_u[0][fluid]).max() <= mtol
    assert np.abs(mom[1:, fluid] - case.rho_u[1:, fluid]).max() <= mtol * 10
    ora = O.OracleLattice(case, storage)
    ora.upload(f)
    om = ora.moments()
    rho_sum = om[0][fluid].sum()
    j = (om[1:, fluid] - np.array(case.force)[:, None] * 0.5) * om[0][fluid]
    # FP64 sums on the device vs numpy sums of the oracle's per-cell moments; for
    # FP32/FP16 storage the per-cell moments themselves are single precision
    rtol = 1e-12 if storage == "f64" else 1e-7
    assert abs(red[0] - rho_sum) <= rtol * rho_sum
    np.testing.assert_allclose(red[1:], j.sum(axis=1), rtol=1e-6,
                               atol=1e-12 if storage == "f64" else 1e-5)


@pytest.mark.parametrize("storage", ["f64", "f32"])
def test_force_field_equilibrium_and_moments(gpu_available, storage):
    """A FORCE field that differs from cell to cell: equilibrium initialisation
    (u - F/2 per cell), the moments (+F/2 per cell) and their reduction against
    the oracle's; clearing the field returns to the uniform force."""
    case = W.kolmogorov_flow((12, 10, 8), amplitude=3e-4, tau=0.7, walls=True)
    rng = np.random.Generator(np.random.PCG64(4))
    case.rho_u[0] = 1.0 + 0.01 * rng.standard_normal(case.shape)
    case.rho_u[1:] = 0.03 * rng.standard_normal((3,) + case.shape)
    fluid = case.ids == 1
    with BlockLattice.from_case(case, storage) as lat:
        lat.ini_equilibrium(case.rho_u)
        f = lat.download_populations("f64")
        mom = lat.moments()
        red = lat.reduce_moments()
        lat.set_force_field(None)
        mom0 = lat.moments()
    tol = {"f64": 1e-15, "f32": 3e-7}[storage]
    assert G.rel_error(f, O.initial_populations(case, storage), 19, case.ids != 0) <= tol
    ora = O.OracleLattice(case, storage)
    ora.upload(f)
    om = ora.moments()
    mtol = {"f64": 1e-14, "f32": 1e-6}[storage]
    np.testing.assert_allclose(mom[:, fluid], om[:, fluid], rtol=0, atol=mtol)
    # the requested velocity comes back on every bulk cell, whatever its force
    np.testing.assert_allclose(mom[1:, fluid], case.rho_u[1:, fluid], rtol=0, atol=mtol * 10)
    j = (om[1:, fluid] - 0.5 * case.force_field[:, fluid]) * om[0][fluid]
    np.testing.assert_allclose(red[1:], j.sum(axis=1), rtol=1e-6, atol=1e-12 if storage == "f64" else 1e-5)
    # without the field the force is the uniform one (zero here): u = j / rho
    np.testing.assert_allclose(mom0[1:, fluid], mom[1:, fluid] - 0.5 * case.force_field[:, fluid],
                               rtol=0, atol=mtol * 10)


def test_force_field_is_refused_where_it_does_not_apply(gpu_available):
    case = W.kolmogorov_flow((6, 5, 4))
    with BlockLattice.from_case(case, "f64") as lat:
        with pytest.raises(ValueError):
            lat.set_force_field(np.zeros((3, 6, 5, 3)))
        bad = case.force_field.copy()
        bad[1, 2, 3, 1] = np.nan
        with pytest.raises(abi.LbmhError):
            lat.set_force_field(bad)
    with BlockLattice.from_case(dataclasses.replace(case, kind="bgk", force_field=None), "f64") as lat:
        with pytest.raises(abi.LbmhError):
            lat.set_force_field(case.force_field)
    ens = Ensemble(19, "f64", kind="forced_bgk", max_members=1, population_bytes=1 << 20, cell_capacity=1 << 12)
    try:
        with pytest.raises(ValueError):
            ens.add(case)
    finally:
        ens.close()


def test_full_size_porous_channel_conserves_mass(gpu_available):
    """Config 2 at full size (256^3, FP32, Guo force 1e-6, tau 0.55). The test runs 200 steps.
    Bounce-back and BGK conserve mass, so the FP64 sum of rho over the bulk cells must
    stay put to FP32 rounding. The body force must drive +x momentum, and no population
    may become non-finite."""
    case = W.porous_channel((256, 256, 256))
    with BlockLattice.from_case(case, "f32") as lat:
        lat.ini_equilibrium(case.rho_u)
        m0 = lat.reduce_moments()
        lat.collide_and_stream(200)
        m1 = lat.reduce_moments()
    print("mass", m0[0], m1[0], "jx", m1[1], flush=True)
    assert np.isfinite(m1).all()
    assert abs(m1[0] - m0[0]) <= 1e-6 * m0[0]
    assert m1[1] > 0 and abs(m1[2]) < 1e-3 * m1[1] and abs(m1[3]) < 1e-3 * m1[1]


def test_local_boundary_lattices_default_to_full_occupancy(gpu_available):
    """A lattice with local velocity boundaries drops the LDS cap by default
    (its whole-lattice pass also collides the flat faces); the shape changes
    scheduling only, so a capped run gives the same populations."""
    case = W.local_velocity_cavity((9, 8, 7), tau=0.7, lid_u=0.06)
    f0 = O.initial_populations(case, "f32")
    runs = []
    for cap in (None, 40960):
        with BlockLattice.from_case(case, "f32") as lat:
            if cap is None:
                assert lat.launch_config() == {"lds_bytes": 0, "xcd_group": 4}
            else:
                lat.set_launch_config(cap, 4)
            lat.upload_populations(f0, "f64")
            lat.initialize()
            lat.collide_and_stream(13)
            runs.append(lat.download_populations("f32"))
    np.testing.assert_array_equal(runs[0], runs[1])


def test_launch_configs_give_identical_results(gpu_available):
    """Launch shape is scheduling only: every configuration must give bit-identical
    populations. The autotuner must leave the lattice's own state untouched."""
    case = W.porous_channel((40, 24, 30), seed=12, force=1e-5, rmin=1, rmax=3)
    f0 = O.initial_populations(case, "f32")
    results = []
    for lds, group in [(0, 0), (40960, 0), (65536, 8), (24576, 3), (49152, 1)]:
        with BlockLattice.from_case(case, "f32") as lat:
            lat.upload_populations(f0, "f64")
            lat.set_launch_config(lds, group)
            assert lat.launch_config() == {"lds_bytes": lds, "xcd_group": group}
            lat.collide_and_stream(25)
            results.append(lat.download_populations("f32"))
    for r in results[1:]:
        np.testing.assert_array_equal(r, results[0])
    with BlockLattice.from_case(case, "f32") as lat:
        lat.upload_populations(f0, "f64")
        before = lat.download_populations("f32")
        chosen = lat.autotune(trial_steps=2)
        assert lat.launch_config() == chosen
        assert lat.step_count() == 0
        np.testing.assert_array_equal(lat.download_populations("f32"), before)


STATS_TOL = {"f64": 1e-12, "f32": 3e-5}


@pytest.mark.parametrize("name", G.names())
def test_gpu_statistics_match_reference(gpu_available, name):
    """Device FP64 statistics of the fixture's last snapshot against the values
    OpenLB reported for it (cells, average rho, average energy, max |u|)."""
    case, meta, arrays = G.load(name)
    ref = G.statistics(name)
    with BlockLattice.from_case(case, meta["precision"]) as lat:
        lat.upload_populations(arrays[f"f_{ref['snapshot']}"].astype(np.float64), "f64")
        st = lat.statistics()
    got = [st["cells"], st["average_rho"], st["average_energy"], st["max_u"]]
    err = G.statistics_error(got, ref)
    print(f"{name}: statistics rel {err:.3e}", flush=True)
    assert err <= STATS_TOL[meta["precision"]]


@pytest.mark.parametrize("storage", ["f64", "f32", "f16"])
def test_gpu_statistics_after_steps_match_oracle(gpu_available, storage):
    """Statistics after stepping, on a case with holes (NoDynamics cells), a
    BounceBack obstacle and a moving wall, against the oracle's."""
    case = _with_holes(W.lid_driven_cavity((11, 9, 70), tau=0.6, lid_u=0.1))
    f0 = O.initial_populations(case, storage)
    with BlockLattice.from_case(case, storage) as lat:
        lat.upload_populations(f0, "f64")
        lat.collide_and_stream(15)
        st = lat.statistics()
        f = lat.download_populations("f64")
    ora = O.OracleLattice(case, storage)
    ora.upload(f)
    ref = O.statistics(case, ora.moments())
    got = np.array([st["cells"], st["average_rho"], st["average_energy"], st["max_u"]])
    assert got[0] == ref[0]
    tol = {"f64": 1e-13, "f32": 1e-6, "f16": 1e-6}[storage]
    np.testing.assert_allclose(got[1:], ref[1:], rtol=tol)


@pytest.mark.parametrize("name", [n for n in G.names() if G.load(n)[0].local])
def test_gpu_initialize_matches_reference(gpu_available, name):
    """Device iniEquilibrium followed by initialize() (the edge and corner
    post-processors once) against OpenLB's populations after initialize()."""
    case, meta, arrays = G.load(name)
    precision = meta["precision"]
    with BlockLattice.from_case(case, precision) as lat:
        lat.ini_equilibrium(case.rho_u)
        lat.initialize()
        got = lat.download_populations("f64")
    err = G.rel_error(got, arrays["f_0"], case.q, case.ids != 0)
    print(f"{name}: initialize rel {err:.3e}", flush=True)
    assert err <= (4e-15 if precision == "f64" else 1e-6)


@pytest.mark.parametrize("storage", ["f64", "f32"])
def test_gpu_local_boundaries_match_oracle_larger(gpu_available, storage):
    """Larger cases, 300 steps, all cells, against the oracle: a 40x36x44 cavity with
    LocalVelocity walls and lid, a duct with a LocalVelocity inlet and LocalPressure
    outlet, and the same two with InterpolatedVelocity / InterpolatedPressure."""
    for case in (W.local_velocity_cavity((40, 36, 44), tau=0.56, lid_u=0.1),
                 W.velocity_channel((60, 18, 22), u_in=0.04, rho_out=1.0, tau=0.58),
                 W.local_velocity_cavity((30, 34, 28), tau=0.58, lid_u=0.1, boundary="interp_velocity"),
                 W.velocity_channel((48, 20, 18), u_in=0.04, rho_out=1.0, tau=0.6, inlet="interp_velocity",
                                    outlet="interp_pressure")):
        ora = O.OracleLattice(case, storage)
        ora.upload(O.initial_populations(case, storage))
        ora.initialize()
        f0 = ora.download()
        got = run_gpu(case, storage, f0, [300])[300]
        ora.step(300, threads=8)
        err = G.rel_error(got, ora.download(), case.q, case.ids != 0)
        print(f"{case.name} {storage}: rel {err:.3e}", flush=True)
        assert err <= (1e-12 if storage == "f64" else 2e-5)


@pytest.mark.parametrize("storage", ["f64", "f32", "f16"])
def test_split_collide_stream_equals_step_with_local_boundaries(gpu_available, storage):
    """lbmh_step and the split lbmh_collide / lbmh_stream calls (the phases a
    multi-block step interleaves with its overlap copies) give the same
    populations, bit for bit, on a cavity with local velocity faces, edges
    and corners."""
    from openlb_amd import abi
    case = W.local_velocity_cavity((19, 14, 23), tau=0.6, lid_u=0.08)
    f0 = O.initial_populations(case, storage)
    runs = []
    for whole_step in (True, False):
        with BlockLattice.from_case(case, storage) as lat:
            lat.upload_populations(f0, "f64")
            lat.initialize()
            if whole_step:
                lat.collide_and_stream(17)
            else:
                L = abi.load()
                for _ in range(17):
                    abi.check(L.lbmh_collide(lat.handle, None))
                    abi.check(L.lbmh_stream(lat.handle, None))
            runs.append(lat.download_populations(storage))
    np.testing.assert_array_equal(runs[0], runs[1])
