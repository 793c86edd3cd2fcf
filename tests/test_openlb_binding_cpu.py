"""The GPU_HIP platform as OpenLB code: OpenLB's own cavity3dBenchmark program
(examples/laminar/cavity3dBenchmark/cavity3d.cpp, unchanged) compiles with
PLATFORMS = CPU_SISD GPU_HIP and links against liblbmh.so.  Builds in the
container that holds the reference tree (integration/openlb/apply.py patches a
scratch copy of it); elsewhere there is nothing to build from."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REFERENCE = "/root/reference"
INTEGRATION = os.path.join(ROOT, "integration", "openlb")
LIB = os.path.join(ROOT, "openlb_amd", "_lib", "liblbmh.so")

needs_reference = pytest.mark.skipif(not os.path.isdir(os.path.join(REFERENCE, "src")),
                                     reason="the OpenLB source tree is only present in the build container")


@pytest.mark.slow
@needs_reference
def test_cavity3d_builds_on_gpu_hip(tmp_path):
    if not os.path.exists(LIB):
        pytest.skip("liblbmh.so is not built (python -m openlb_amd.build)")
    scratch = tmp_path / "olb"
    out = subprocess.run(["make", "-s", "-j8", "-C", INTEGRATION, "_build/cavity3d_gpu_hip", f"OLB={scratch}",
                          f"REF={REFERENCE}"], capture_output=True, text=True, timeout=1200)
    assert out.returncode == 0, out.stdout[-3000:] + out.stderr[-3000:]
    binary = os.path.join(INTEGRATION, "_build", "cavity3d_gpu_hip")
    assert os.path.exists(binary)
    # the patched tree declares the platform next to CUDA's
    platform_h = (scratch / "src" / "core" / "platform" / "platform.h").read_text()
    assert "GPU_HIP" in platform_h
    # no GPU here: the program sets up geometry, dynamics and boundaries on the
    # host, maps every dynamics type to the device's cell roles, and stops at
    # the first device allocation (on a GPU machine it runs to the end)
    run = subprocess.run([binary, "--size", "12", "--steps", "2", "--no-results"], cwd=tmp_path,
                         capture_output=True, text=True, timeout=300)
    if run.returncode != 0:
        assert "lbmh_lattice_create" in run.stderr, run.stderr[-2000:]
        assert "not available on this platform" not in run.stderr


@needs_reference
def test_apply_refuses_a_changed_reference(tmp_path):
    """Every edit is anchored; a reference whose anchor lines moved fails loudly."""
    import importlib.util
    spec = importlib.util.spec_from_file_location("olb_apply", os.path.join(INTEGRATION, "apply.py"))
    apply = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(apply)
    text = "enum struct Platform : std::uint8_t {\n  CPU_SISD,\n};\n"
    with pytest.raises(apply.AnchorError):
        apply.insert_after(text, "  GPU_CUDA, /// GPU using CUDA\n", "  GPU_HIP,\n", "platform.h")
    cases = ("#ifdef PLATFORM_GPU_CUDA\n    case Platform::GPU_CUDA:\n      return f(1);\n#endif\n" * 2)
    with pytest.raises(apply.AnchorError):
        apply.duplicate_cuda_cases(cases, "dispatch.h", expected=3)
    doubled = apply.duplicate_cuda_cases(cases, "dispatch.h", expected=2)
    assert doubled.count("case Platform::GPU_HIP:") == 2 and doubled.count("case Platform::GPU_CUDA:") == 2
