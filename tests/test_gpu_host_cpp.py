"""The C++ host layer (openlb_amd/host/olb_hip.h) drives the C ABI the way
OpenLB's benchmark harness drives SuperLattice. This runs its port of
examples/laminar/cavity3dBenchmark and checks its populations against the
oracle on the same cavity."""
import os
import subprocess

import numpy as np
import pytest

import golden_cases as G
from oracle import oracle as O
from openlb_amd import workloads as W

pytestmark = pytest.mark.gpu

EXE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                   "openlb_amd", "_lib", "cavity3d_benchmark")


@pytest.mark.parametrize("variant", ["bounce_back", "local_velocity"])
@pytest.mark.parametrize("storage", ["f64", "f32"])
def test_cpp_cavity_benchmark_matches_oracle(gpu_available, storage, variant, tmp_path):
    size, steps = 24, 50
    dump = tmp_path / "pop.bin"
    cmd = [EXE, "--size", str(size), "--steps", str(steps), "--dump", str(dump), "--stats"]
    if storage == "f64":
        cmd.append("--fp64")
    if variant == "bounce_back":
        cmd.append("--bounce-back")
    out = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    lines = out.stdout.strip().splitlines()
    fields = [s.strip() for s in [ln for ln in lines if "," in ln][-1].split(",")]
    assert fields[:5] == [str(size), str(steps), "1", "1", "1"]
    assert float(fields[5]) > 0
    case = W.cavity_benchmark(size) if variant == "bounce_back" else W.cavity_benchmark_local(size)
    got = np.fromfile(dump, dtype=np.float64).reshape((19,) + case.shape)
    ora = O.OracleLattice(case, storage)
    ora.upload(O.initial_populations(case, storage))
    ora.initialize()
    ora.step(10 + steps, threads=8)
    mask = case.ids == 1 if variant == "bounce_back" else case.ids != 0
    err = G.rel_error(got, ora.download(), 19, mask)
    print(f"C++ host layer cavity {storage}: rel {err:.3e}", flush=True)
    assert err <= (1e-13 if storage == "f64" else 1e-5)
    # getStatistics() of the C++ layer: statistics of the dumped state
    st = [float(v) for v in [ln for ln in lines if ln.startswith("statistics ")][-1].split()[1:]]
    ora.upload(got)
    ref = ora.statistics()
    assert st[0] == ref[0]
    np.testing.assert_allclose(st[1:], ref[1:], rtol=1e-12 if storage == "f64" else 1e-6)


def test_cpp_layer_reports_errors_as_exceptions(gpu_available):
    """A misuse surfaces as an exception message and a non-zero exit, not a crash."""
    out = subprocess.run([EXE, "--size", "4", "--steps", "1", "--tau", "0.4"], capture_output=True,
                         text=True, timeout=60)
    assert out.returncode == 1
    assert "omega must lie in (0, 2)" in out.stderr


@pytest.mark.parametrize("variant", ["bounce_back", "local_velocity"])
@pytest.mark.parametrize("storage", ["f64", "f32"])
def test_cpp_save_and_load_resume_the_run(gpu_available, storage, variant, tmp_path):
    """lattice.save(name) / lattice.load(name) of the C++ layer: OpenLB's file
    <name>_rank0_size1.dat, whose contents have OpenLB's layout, and a run that
    resumes from it ends where the uninterrupted run ends, bit for bit."""
    from oracle import checkpoint as K
    from openlb_amd import checkpoint as CK

    size = 16
    base = [EXE, "--size", str(size)] + (["--fp64"] if storage == "f64" else []) + \
        (["--bounce-back"] if variant == "bounce_back" else [])

    def run(*extra):
        out = subprocess.run(base + list(extra), capture_output=True, text=True, timeout=120)
        assert out.returncode == 0, out.stdout + out.stderr

    name = str(tmp_path / "cavity")
    run("--steps", "7", "--save", name, "--dump", str(tmp_path / "a.bin"))
    run("--steps", "5", "--load", name, "--dump", str(tmp_path / "b.bin"))
    run("--steps", "22", "--dump", str(tmp_path / "c.bin"))
    b = np.fromfile(tmp_path / "b.bin")
    c = np.fromfile(tmp_path / "c.bin")
    np.testing.assert_array_equal(b, c)
    case = W.cavity_benchmark(size) if variant == "bounce_back" else W.cavity_benchmark_local(size)
    real = storage
    raw = CK.read_file(name + "_rank0_size1.dat")
    shifts, pops, _ = K.unpack(case, raw, real)
    np.testing.assert_array_equal(shifts, K.rotate_shifts(case, 17))
    a = np.fromfile(tmp_path / "a.bin").reshape((19,) + case.shape)
    np.testing.assert_array_equal(K.core_of(case, pops).astype(np.float64), a)


@pytest.mark.parametrize("storage", ["f64", "f32"])
def test_cpp_vtk_writer_matches_python_writer(gpu_available, storage, tmp_path):
    """SuperVTMwriter3D of the C++ layer writes the same three files as the
    package's writer, which test_vtk_cpu.py pins to OpenLB's own output byte
    for byte, for the same state and converter."""
    from openlb_amd import vtk as V
    from openlb_amd.lattice import BlockLattice

    size, steps = 16, 20
    cdir = tmp_path / "cpp"
    (cdir / "data").mkdir(parents=True)
    cmd = [EXE, "--size", str(size), "--steps", str(steps), "--vtk", str(cdir), "--dump", str(tmp_path / "p.bin")]
    if storage == "f64":
        cmd.append("--fp64")
    out = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    vals = [float(v) for v in [ln for ln in out.stdout.splitlines() if ln.startswith("vtk ")][-1].split()[1:]]

    class Conv:
        def factors(self):
            return np.array(vals[:3])

    case = W.cavity_benchmark_local(size)
    pops = np.fromfile(tmp_path / "p.bin").reshape((19,) + case.shape)
    with BlockLattice.from_case(case, storage) as lat:
        lat.upload_populations(pops, "f64")
        w = V.SuperVTMwriter3D("cavity3d", out_dir=str(tmp_path / "py"))
        w.add_functor(V.SuperLatticePhysVelocity3D(lat, Conv()))
        w.add_functor(V.SuperLatticePhysPressure3D(lat, Conv()))
        w.create_master_file()
        w.write(10 + steps, delta=vals[3])
    it = 10 + steps
    for rel in ("cavity3d.pvd", f"data/cavity3d_iT{it:07d}.vtm", f"data/cavity3d_iT{it:07d}iC00000.vti"):
        assert (cdir / rel).read_text() == (tmp_path / "py" / rel).read_text(), rel


UQ_EXE = os.path.join(os.path.dirname(EXE), "uq_cavity3d")


@pytest.mark.parametrize("method", ["gpc", "monte_carlo"])
def test_cpp_uq_ensemble_study(gpu_available, method, tmp_path):
    """The C++ Ensemble runs a UQ study of LocalVelocity cavities (one member
    per sampling point, set up with the SuperLattice calls): the sampling
    points are those of OpenLB's UQ module (as restated in openlb_amd/uq.py,
    pinned to the reference), every member's physVelocity equals the same
    cavity run alone, and the mean / std fields are uq.mean / uq.std of the
    members."""
    from openlb_amd import uq as U
    from openlb_amd import vtk as V
    from openlb_amd.lattice import BlockLattice

    size, steps = 10, 30
    n = size + 1
    cmd = [UQ_EXE, "--size", str(size), "--steps", str(steps), "--dump", str(tmp_path / "uq.bin")]
    if method == "monte_carlo":
        cmd += ["--mc", "6", "--seed", "123456"]
    out = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    lines = out.stdout.splitlines()
    samples = np.array([float(v) for v in [ln for ln in lines if ln.startswith("samples")][0].split()[1:]])
    tau = float([ln for ln in lines if ln.startswith("tau")][0].split()[1])
    u = U.UncertaintyQuantification(method)
    if method == "gpc":
        u.initialize_gpc(2, 5, U.uniform(0.9, 1.1))
        np.testing.assert_allclose(samples, u.sampling_points()[:, 0], rtol=0, atol=1e-15)
    else:
        u.initialize_monte_carlo(6, U.uniform(0.9, 1.1), 123456)
        np.testing.assert_array_equal(samples, u.sampling_points()[:, 0])
    P = (n + 2) ** 3 * 3
    raw = np.fromfile(tmp_path / "uq.bin", dtype=np.uint8)
    members = np.frombuffer(raw[: len(samples) * P * 4].tobytes(), dtype=np.float32).reshape(len(samples), P)
    rest = np.frombuffer(raw[len(samples) * P * 4:].tobytes(), dtype=np.float64)
    mean, std = rest[:P], rest[P:]
    f32 = np.float32
    conv = V.UnitConverter(f32(1) / f32(size), f32(0.01) / f32(size), 1.0, 1.0, 1e-3, 1.0, 0.0, "f32")
    for k, s in enumerate(samples):
        case = W.local_velocity_cavity((n, n, n), tau=tau, lid_u=float(f32(0.01) * f32(s)))
        case.omega = float(f32(1.0 / tau))
        with BlockLattice.from_case(case, "f32") as lat:
            lat.ini_equilibrium(case.rho_u)
            lat.initialize()
            lat.collide_and_stream(steps)
            np.testing.assert_array_equal(members[k], V.SuperLatticePhysVelocity3D(lat, conv).evaluate(1))
    scale = np.abs(members).max()
    np.testing.assert_allclose(mean, u.mean(members), rtol=0, atol=1e-12 * scale)
    np.testing.assert_allclose(std, u.std(members), rtol=0, atol=1e-12 * scale)
