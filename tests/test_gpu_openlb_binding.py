"""OpenLB's own cavity3dBenchmark program (examples/laminar/cavity3dBenchmark/
cavity3d.cpp: D3Q19 BGK, LocalVelocity boundaries on the lid and the walls,
Re 1000) built in the build container on OpenLB's CPU_SISD platform and on the
GPU_HIP platform (integration/openlb), whose blocks run their time step through
liblbmh; as written (value type float) and with `using T = double;`.  Each run
writes OpenLB's VTK files; the physical velocity and pressure of the GPU_HIP
runs are compared with the CPU runs'."""
import os
import subprocess

import numpy as np
import pytest

from openlb_amd import vtk

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BUILD = os.path.join(ROOT, "integration", "openlb", "_build")
BINARIES = {k: os.path.join(BUILD, k) for k in ("cavity3d_gpu_hip", "cavity3d_cpu", "cavity3d_gpu_hip_f64",
                                                 "cavity3d_cpu_f64")}

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not all(os.path.exists(b) for b in BINARIES.values()),
                                 reason="cavity3d binaries not built (make -C integration/openlb)")]

SIZE, STEPS = 24, 60
FIELDS = ("physVelocity", "physPressure")


@pytest.fixture(scope="module")
def runs(tmp_path_factory):
    out = {}
    for name, binary in BINARIES.items():
        cwd = tmp_path_factory.mktemp(name)
        print(f"running {name} --size {SIZE} --steps {STEPS}", flush=True)
        p = subprocess.run([binary, "--size", str(SIZE), "--steps", str(STEPS)], cwd=cwd, capture_output=True,
                           text=True, timeout=240)
        assert p.returncode == 0, p.stdout[-2000:] + p.stderr[-2000:]
        line = [ln for ln in p.stdout.splitlines() if ln.count(",") == 5][-1]
        size, steps = (int(float(v)) for v in line.split(",")[:2])
        assert (size, steps) == (SIZE, STEPS)
        out[name] = {f: vtk.read_vti_arrays(os.path.join(cwd, "tmp", "vtkData", "data",
                                                          f"{f}_iT0000000iC00000.vti"))[f] for f in FIELDS}
        out[name]["mlups"] = float(line.split(",")[5])
    return out


def max_err(a, b):
    return float(np.abs(a.astype(np.float64) - b.astype(np.float64)).max())


def test_double_precision_run_matches_openlb(gpu_available, runs):
    """T = double on both platforms: the GPU_HIP run equals OpenLB's CPU run to
    FP64 rounding (the lattice's FP64 path is pinned to OpenLB at 1e-12)."""
    for f in FIELDS:
        a, b = runs["cavity3d_gpu_hip_f64"][f], runs["cavity3d_cpu_f64"][f]
        scale = float(np.abs(b).max())
        err = max_err(a, b)
        print(f"f64 {f}: max |gpu_hip - cpu| = {err:.3e} of max {scale:.3e}", flush=True)
        assert err <= 1e-6 * scale  # the writer stores float: 1e-6 of the field's range is a few float ulps


def test_float_run_as_close_to_double_as_openlbs_own(gpu_available, runs):
    """T = float, the program as written: at tau = 0.5007 rounding is amplified,
    so two float implementations differ by about as much as each differs from the
    double run.  The GPU_HIP float run must stay as close to OpenLB's double run
    as OpenLB's own float run is (within a factor 2)."""
    for f in FIELDS:
        ref = runs["cavity3d_cpu_f64"][f]
        gpu = max_err(runs["cavity3d_gpu_hip"][f], ref)
        cpu = max_err(runs["cavity3d_cpu"][f], ref)
        print(f"f32 {f}: |gpu_hip - cpu f64| = {gpu:.3e}, |cpu f32 - cpu f64| = {cpu:.3e}", flush=True)
        assert gpu <= 2.0 * cpu + 1e-7


DUCT = {k: os.path.join(BUILD, k) for k in ("duct3d_gpu_hip", "duct3d_cpu")}
FORCED = {k: os.path.join(BUILD, k) for k in ("duct3d_forced_gpu_hip", "duct3d_forced_cpu")}


@pytest.mark.skipif(not all(os.path.exists(b) for b in DUCT.values()),
                    reason="duct3d binaries not built (make -C integration/openlb)")
@pytest.mark.parametrize("family", ["local", "interpolated"])
def test_duct_boundaries_match_openlb(gpu_available, family, tmp_path):
    """integration/openlb/examples/duct3d.cpp, an OpenLB program of this
    repository (T = double): a duct with BounceBack walls, a velocity inlet and
    a pressure outlet, LocalVelocity + LocalPressure or InterpolatedVelocity +
    InterpolatedPressure (the plane finite-difference post-processors), built
    on CPU_SISD and on GPU_HIP.  The GPU_HIP run's physical velocity and
    pressure equal the CPU run's to the float precision of the VTK output."""
    out = {}
    for name, binary in DUCT.items():
        cwd = tmp_path / name
        cwd.mkdir()
        cmd = [binary, "--steps", "300"] + (["--interpolated"] if family == "interpolated" else [])
        print("running", " ".join(cmd), flush=True)
        p = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True, timeout=240)
        assert p.returncode == 0, p.stdout[-2000:] + p.stderr[-2000:]
        out[name] = {f: vtk.read_vti_arrays(str(cwd / "tmp" / "vtkData" / "data" / f"{f}_iT0000000iC00000.vti"))[f]
                     for f in FIELDS}
        line = [ln for ln in p.stdout.splitlines() if ln.count(",") == 5][-1]
        out[name]["platform"] = int(line.split(",")[5])
    # the GPU_HIP build runs its block on another platform than CPU_SISD
    assert out["duct3d_gpu_hip"]["platform"] != out["duct3d_cpu"]["platform"]
    for f in FIELDS:
        a, b = out["duct3d_gpu_hip"][f], out["duct3d_cpu"][f]
        scale = float(np.abs(b).max())
        err = max_err(a, b)
        print(f"duct {family} {f}: max |gpu_hip - cpu| = {err:.3e} of max {scale:.3e}", flush=True)
        assert scale > 0 and err <= 1e-6 * scale


def _pieces(cwd, field: str) -> list:
    """The per-cuboid pieces SuperVTMwriter3D wrote for a field, in cuboid order."""
    import glob
    files = sorted(glob.glob(os.path.join(str(cwd), "tmp", "vtkData", "data", f"{field}_iT0000000iC*.vti")))
    return [vtk.read_vti_arrays(f)[field] for f in files]


@pytest.mark.skipif(not all(os.path.exists(b) for b in DUCT.values()),
                    reason="duct3d binaries not built (make -C integration/openlb)")
@pytest.mark.parametrize("cuboids", [1, 4])
def test_duct_smagorinsky_bulk_matches_openlb(gpu_available, cuboids, tmp_path):
    """A SmagorinskyBGK bulk next to InterpolatedVelocity / InterpolatedPressure
    faces with their default BGK mixin, the set-up of aorta3d.cpp: the GPU_HIP
    run equals the CPU run piece by piece, on one cuboid and on four."""
    out = {}
    for name in ("duct3d_gpu_hip", "duct3d_cpu"):
        cwd = tmp_path / name
        cwd.mkdir()
        cmd = [DUCT[name], "--steps", "300", "--interpolated", "--smagorinsky", "--cuboids", str(cuboids)]
        print("running", " ".join(cmd), flush=True)
        p = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True, timeout=240)
        assert p.returncode == 0, p.stdout[-2000:] + p.stderr[-2000:]
        out[name] = {f: _pieces(cwd, f) for f in FIELDS}
    for f in FIELDS:
        gpu, cpu = out["duct3d_gpu_hip"][f], out["duct3d_cpu"][f]
        assert len(gpu) == len(cpu) == cuboids
        scale = max(float(np.abs(b).max()) for b in cpu)
        err = max(max_err(a, b) for a, b in zip(gpu, cpu))
        print(f"duct smagorinsky {cuboids} cuboids {f}: max |gpu_hip - cpu| = {err:.3e} of max {scale:.3e}",
              flush=True)
        assert scale > 0 and err <= 1e-6 * scale


@pytest.mark.skipif(not all(os.path.exists(b) for b in DUCT.values()),
                    reason="duct3d binaries not built (make -C integration/openlb)")
@pytest.mark.parametrize("family", ["local", "interpolated"])
def test_duct_on_several_cuboids_matches_openlb(gpu_available, family, tmp_path):
    """The duct on 6 cuboids (CuboidDecomposition, OpenLB's default overlap of
    3): the inlet and outlet faces and their post-processors cross cuboid
    interfaces.  Each GPU_HIP block classifies its boundary cells with the
    materials of its overlap and reads the neighbours of its post-processors
    from the neighbouring blocks (lbmh_set_overlap_ids); every piece of the
    GPU_HIP run's output equals the CPU run's piece to the VTK float precision,
    and the CPU run on 6 cuboids equals OpenLB's own run on one."""
    out = {}
    runs = [("duct3d_gpu_hip", 6), ("duct3d_cpu", 6), ("duct3d_cpu", 1)]
    for name, n in runs:
        cwd = tmp_path / f"{name}_{n}"
        cwd.mkdir()
        cmd = [DUCT[name], "--steps", "300", "--cuboids", str(n)]
        cmd += ["--interpolated"] if family == "interpolated" else []
        print("running", " ".join(cmd), flush=True)
        p = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True, timeout=240)
        assert p.returncode == 0, p.stdout[-2000:] + p.stderr[-2000:]
        assert f"cuboids {n}" in p.stdout
        out[(name, n)] = {f: _pieces(cwd, f) for f in FIELDS}
    for f in FIELDS:
        gpu, cpu = out[("duct3d_gpu_hip", 6)][f], out[("duct3d_cpu", 6)][f]
        assert len(gpu) == len(cpu) == 6
        scale = max(float(np.abs(b).max()) for b in cpu)
        err = max(max_err(a, b) for a, b in zip(gpu, cpu))
        print(f"duct {family} 6 cuboids {f}: max |gpu_hip - cpu| = {err:.3e} of max {scale:.3e}", flush=True)
        assert scale > 0 and err <= 1e-6 * scale
        # OpenLB itself: the 6 pieces hold the one-cuboid run's values (a
        # piece also holds its overlap-1 ring; compare the value sets' extremes)
        one = out[("duct3d_cpu", 1)][f][0]
        assert abs(max(float(b.max()) for b in cpu) - float(one.max())) <= 1e-6 * scale
        assert abs(min(float(b.min()) for b in cpu) - float(one.min())) <= 1e-6 * scale


@pytest.mark.parametrize("cuboids", [2, 4])
def test_cavity_on_several_cuboids_matches_openlb(gpu_available, cuboids, tmp_path):
    """cavity3dBenchmark (T = double, LocalVelocity on every wall) with
    --cuboids-per-process: the edges and corners of the cavity and their
    post-processors sit on cuboid interfaces.  Every piece of the GPU_HIP run
    equals the CPU run's."""
    out = {}
    for name in ("cavity3d_gpu_hip_f64", "cavity3d_cpu_f64"):
        cwd = tmp_path / name
        cwd.mkdir()
        cmd = [BINARIES[name], "--size", str(SIZE), "--steps", str(STEPS), "--cuboids-per-process", str(cuboids)]
        print("running", " ".join(cmd), flush=True)
        p = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True, timeout=240)
        assert p.returncode == 0, p.stdout[-2000:] + p.stderr[-2000:]
        out[name] = {f: _pieces(cwd, f) for f in FIELDS}
    for f in FIELDS:
        gpu, cpu = out["cavity3d_gpu_hip_f64"][f], out["cavity3d_cpu_f64"][f]
        assert len(gpu) == len(cpu) >= 2
        scale = max(float(np.abs(b).max()) for b in cpu)
        err = max(max_err(a, b) for a, b in zip(gpu, cpu))
        print(f"cavity {cuboids} cuboids {f}: max |gpu_hip - cpu| = {err:.3e} of max {scale:.3e}", flush=True)
        assert scale > 0 and err <= 1e-6 * scale


@pytest.mark.skipif(not all(os.path.exists(b) for b in DUCT.values()),
                    reason="duct3d binaries not built (make -C integration/openlb)")
@pytest.mark.parametrize("first,second", [("duct3d_gpu_hip", "duct3d_cpu"), ("duct3d_cpu", "duct3d_gpu_hip")])
def test_checkpoint_moves_between_platforms(gpu_available, first, second, tmp_path):
    """OpenLB's own serializer on the duct (interpolated faces, 2 cuboids): one
    platform runs 150 steps and saves (SuperLattice::save after the switch to
    Evaluation), the other loads that file and runs 150 more.  The result
    equals a CPU_SISD run of 300 steps to the VTK float precision: a GPU_HIP
    block's checkpoint is OpenLB's CPU bytes, and a loaded one reaches the
    device."""
    common = ["--interpolated", "--cuboids", "2"]
    ck = str(tmp_path / "duct")
    runs = [(first, ["--steps", "150", "--save", ck]), (second, ["--steps", "150", "--load", ck]),
            ("duct3d_cpu", ["--steps", "300"])]
    out = []
    for k, (name, extra) in enumerate(runs):
        cwd = tmp_path / f"run{k}"
        cwd.mkdir()
        cmd = [DUCT[name]] + common + extra
        print("running", " ".join(cmd), flush=True)
        p = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True, timeout=240)
        assert p.returncode == 0, p.stdout[-2000:] + p.stderr[-2000:]
        out.append({f: _pieces(cwd, f) for f in FIELDS})
    assert os.path.exists(ck + "_rank0_size1.dat")
    for f in FIELDS:
        resumed, ref = out[1][f], out[2][f]
        scale = max(float(np.abs(b).max()) for b in ref)
        err = max(max_err(a, b) for a, b in zip(resumed, ref))
        print(f"{first} -> {second} {f}: max |resumed - cpu 300| = {err:.3e} of max {scale:.3e}", flush=True)
        assert scale > 0 and err <= 1e-6 * scale


@pytest.mark.skipif(not all(os.path.exists(b) for b in FORCED.values()),
                    reason="duct3d_forced binaries not built (make -C integration/openlb)")
@pytest.mark.parametrize("family", ["local", "interpolated"])
def test_duct_forced_bulk_matches_openlb(gpu_available, family, tmp_path):
    """The duct built with a ForcedBGK bulk and a body force (D3Q19<FORCE>),
    its faces keeping their BGK mixin, on 2 cuboids: the edge and plane
    processors read the forced bulk neighbours' velocity with F/2 added.  The
    GPU_HIP run equals the CPU run piece by piece."""
    out = {}
    for name in ("duct3d_forced_gpu_hip", "duct3d_forced_cpu"):
        cwd = tmp_path / name
        cwd.mkdir()
        cmd = [FORCED[name], "--steps", "300", "--cuboids", "2"] + (["--interpolated"] if family == "interpolated" else [])
        print("running", " ".join(cmd), flush=True)
        p = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True, timeout=240)
        assert p.returncode == 0, p.stdout[-2000:] + p.stderr[-2000:]
        out[name] = {f: _pieces(cwd, f) for f in FIELDS}
    for f in FIELDS:
        gpu, cpu = out["duct3d_forced_gpu_hip"][f], out["duct3d_forced_cpu"][f]
        assert len(gpu) == len(cpu) == 2
        scale = max(float(np.abs(b).max()) for b in cpu)
        err = max(max_err(a, b) for a, b in zip(gpu, cpu))
        print(f"duct forced {family} {f}: max |gpu_hip - cpu| = {err:.3e} of max {scale:.3e}", flush=True)
        assert scale > 0 and err <= 1e-6 * scale
