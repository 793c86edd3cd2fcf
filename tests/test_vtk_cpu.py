"""VTK output in OpenLB's format, pinned on the CPU against OpenLB's own files.

tests/golden/checkpoints/<name>.npz holds, besides OpenLB's checkpoint after
k steps, the files OpenLB's SuperVTMwriter3D wrote for that same state
(fields.pvd, fields_iT*.vtm, fields_iT*iC00000.vti with the functors density,
velocity, physVelocity and physPressure; gen_golden.py --checkpoints-only).

* The oracle's restatement of the functors (olo_field_export: the writer's
  point order, the overlap, each cell's momenta, the unit conversion) gives
  OpenLB's float arrays bit for bit from the k-step populations.
* The package's writer (openlb_amd/vtk.py), fed those arrays, writes OpenLB's
  three files byte for byte (XML, base64, zlib stream).
"""
import os
import tempfile

import numpy as np
import pytest

import golden_cases as G
from oracle import oracle as O
from openlb_amd import vtk as V

FUNCTORS = ["density", "velocity", "physVelocity", "physPressure"]


def openlb_arrays(name):
    files = G.load_vtk(name)
    with tempfile.NamedTemporaryFile("w", suffix=".vti", delete=False) as fh:
        fh.write(files["vti"])
        path = fh.name
    try:
        return V.read_vti_arrays(path)
    finally:
        os.unlink(path)


def converter(real):
    # the converter the reference driver writes its physical functors with
    return V.UnitConverter(0.01, 1e-4, 0.1, 1.0, 1e-5, 1000.0, 100.0, real)


@pytest.mark.parametrize("name", G.checkpoint_names())
def test_oracle_functors_reproduce_openlb_arrays(name):
    case, meta, arrays, _ = G.load_checkpoint(name)
    k, real = meta["checkpoint_step"], meta["precision"]
    ref = openlb_arrays(name)
    assert list(ref) == FUNCTORS
    lat = O.OracleLattice(case, real)
    lat.upload(arrays[f"f_{k}"])
    conv = converter(real).factors()
    for field, key in enumerate(FUNCTORS):
        np.testing.assert_array_equal(lat.field_export(field, 1, conv), ref[key], err_msg=key)


def test_point_order_and_overlap_against_populations():
    """The velocity array in VTK point order (x fastest) over the overlap,
    checked on the core against u computed from the populations directly."""
    name = "tgv19_bgk_f64"
    case, meta, arrays, _ = G.load_checkpoint(name)
    k = meta["checkpoint_step"]
    f = arrays[f"f_{k}"] + G.weights(case.q)[:, None, None, None]
    c = np.array(__import__("oracle.checkpoint", fromlist=["VELOCITIES"]).VELOCITIES[case.q], dtype=np.float64)
    rho = f.sum(axis=0)
    u = np.einsum("qd,qxyz->dxyz", c, f) / rho
    P = [n + 2 for n in case.shape]
    ref = openlb_arrays(name)["velocity"].reshape(P[2], P[1], P[0], 3)
    core = ref[1:-1, 1:-1, 1:-1].transpose(3, 2, 1, 0)      # -> [d][x][y][z]
    np.testing.assert_allclose(core, u, rtol=0, atol=1e-7 * np.abs(u).max())
    # fully periodic: the overlap layer repeats the opposite face
    np.testing.assert_array_equal(ref[0, 1:-1, 1:-1], ref[-2, 1:-1, 1:-1])
    np.testing.assert_array_equal(ref[1:-1, 1:-1, 0], ref[1:-1, 1:-1, -2])


class _Fixed(V._Functor):
    """A functor with given values (the writer test needs no device)."""

    def __init__(self, name, dim, values, shape):
        self.name, self.dim, self._values, self._shape = name, dim, values, shape
        self.member = None

    def shape(self):
        return self._shape

    def evaluate(self, overlap=1):
        return self._values


@pytest.mark.parametrize("name", G.checkpoint_names())
def test_writer_reproduces_openlb_files(name, tmp_path):
    case, meta, _, _ = G.load_checkpoint(name)
    k = meta["checkpoint_step"]
    ref = openlb_arrays(name)
    files = G.load_vtk(name)
    w = V.SuperVTMwriter3D("fields", out_dir=str(tmp_path / "vtkData"))
    for key in FUNCTORS:
        w.add_functor(_Fixed(key, 3 if "elocity" in key else 1, ref[key], tuple(case.shape)))
    w.create_master_file()
    w.write(k)
    out = tmp_path / "vtkData"
    assert (out / "fields.pvd").read_text() == files["pvd"]
    assert (out / "data" / f"fields_iT{k:07d}.vtm").read_text() == files["vtm"]
    assert (out / "data" / f"fields_iT{k:07d}iC00000.vti").read_text() == files["vti"]


def test_writer_formats(tmp_path):
    """Uncompressed binary and ascii pieces read back; several writes append to the .pvd."""
    vals = np.arange(4 * 3 * 5, dtype=np.float32) / 7
    for binary, compress in ((True, False), (False, False)):
        w = V.SuperVTMwriter3D("f", binary=binary, compress=compress, out_dir=str(tmp_path / f"o{int(binary)}"))
        w.add_functor(_Fixed("density", 1, vals, (2, 1, 3)))
        w.create_master_file()
        w.write(0)
        w.write(10)
        got = V.read_vti_arrays(str(tmp_path / f"o{int(binary)}" / "data" / "f_iT0000010iC00000.vti"))["density"]
        np.testing.assert_allclose(got, vals, rtol=1e-5 if not binary else 0)
        pvd = (tmp_path / f"o{int(binary)}" / "f.pvd").read_text()
        assert pvd.count("<DataSet ") == 2 and pvd.endswith("</Collection>\n</VTKFile>\n")
    assert V.create_file_name("cavity3d", 120, 3) == "cavity3d_iT0000120iC00003"
