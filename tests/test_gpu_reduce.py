"""GPU tests of lbmh_reduce_field, the device reduction of a cell field
(OpenLB's SuperLatticeFieldReductionO with SumO / MinO / MaxO, and the
SuperSum3D / SuperMin3D / SuperMax3D / SuperAverage3D functors over a
material indicator).

The checkers are the oracle's per-cell values, which test_vtk_cpu.py pins to
OpenLB's own VTK arrays bit for bit (its functor evaluation, float32) and
test_oracle_golden.py to OpenLB's populations (its FP64 moments):
* over every cell with dynamics, against numpy reductions of the oracle's
  float32 functor values: within one float rounding per value;
* over the bulk cells, against the oracle's FP64 moments (u including F/2):
  FP64 lattices within 1e-12 of the largest magnitude, FP32 within 1e-6;
* the FORCE field, against the case's field: sums to FP64 rounding, minimum
  and maximum exactly.
"""
import numpy as np
import pytest

import golden_cases as G
from oracle import oracle as O
from openlb_amd import abi
from openlb_amd import checkpoint as CK
from openlb_amd import vtk as V
from openlb_amd import workloads as W
from openlb_amd.lattice import BlockLattice

pytestmark = pytest.mark.gpu

FIELDS = ["density", "velocity", "physVelocity", "physPressure"]


def converter(real):
    return V.UnitConverter(0.01, 1e-4, 0.1, 1.0, 1e-5, 1000.0, 100.0, real)


def reduce_np(values, op):
    return {"sum": values.sum(axis=0), "min": values.min(axis=0), "max": values.max(axis=0)}[op]


@pytest.mark.parametrize("name", G.checkpoint_names())
def test_reductions_match_the_oracle_functors(gpu_available, name):
    case, meta, arrays, text = G.load_checkpoint(name)
    k, real = meta["checkpoint_step"], meta["precision"]
    conv = converter(real).factors()
    nx, ny, nz = case.shape
    ora = O.OracleLattice(case, real)
    ora.upload(arrays[f"f_{k}"])
    with BlockLattice.from_case(case, real) as lat:
        lat.restore(CK.decode_dat(text), real)
        for field_id, field in enumerate(FIELDS):
            dim = 3 if "elocity" in field else 1
            # VTK point order, x fastest -> [x][y][z][component]
            vals = ora.field_export(field_id, 0, conv).astype(np.float64).reshape(nz, ny, nx, dim)
            vals = vals.transpose(2, 1, 0, 3)
            for ids in (None, sorted(int(i) for i in np.unique(case.ids) if i not in (0, 1))):
                sel = case.ids != 0 if ids is None else np.isin(case.ids, ids)
                if not sel.any():
                    continue
                pick = vals[sel]
                for op in ("sum", "min", "max"):
                    got, cells = lat.reduce_field(field, op, ids, conv)
                    want = reduce_np(pick, op)
                    assert cells == int(sel.sum())
                    scale = np.abs(pick).sum(axis=0) if op == "sum" else np.abs(want)
                    tol = 2.0 ** -23 * scale + 1e-30
                    assert np.all(np.abs(got - want) <= tol), (field, op, ids, got, want)


@pytest.mark.parametrize("name", [n for n in G.names() if n.endswith("_f64") or n.endswith("_f32")])
def test_bulk_reductions_match_oracle_moments(gpu_available, name):
    case, meta, arrays = G.load(name)
    real = meta["precision"]
    snap = max(int(s) for s in meta["steps"])
    f = arrays[f"f_{snap}"].astype(np.float64)
    ora = O.OracleLattice(case, real)
    ora.upload(f)
    mom = ora.moments()
    bulk = case.ids == 1
    if not bulk.any():
        pytest.skip("no bulk cells")
    tol = {"f64": 1e-12, "f32": 1e-6}[real]
    with BlockLattice.from_case(case, real) as lat:
        lat.upload_populations(f, "f64")
        for field, vals in (("density", mom[0][bulk][:, None]), ("velocity", mom[1:, bulk].T)):
            for op in ("sum", "min", "max"):
                got, cells = lat.reduce_field(field, op, [1])
                want = reduce_np(vals, op)
                assert cells == int(bulk.sum())
                scale = np.abs(vals).sum(axis=0) if op == "sum" else np.abs(vals).max(axis=0)
                assert np.all(np.abs(got - want) <= tol * scale + 1e-300), (name, field, op, got, want)


def test_force_reductions(gpu_available):
    case = W.kolmogorov_flow((12, 10, 8), amplitude=3e-4, tau=0.7, walls=True)
    bulk = case.ids == 1
    F = case.force_field[:, bulk].T
    with BlockLattice.from_case(case, "f64") as lat:
        got, cells = lat.reduce_field("force", "sum")
        assert cells == int((case.ids != 0).sum())
        np.testing.assert_allclose(got, F.sum(axis=0), rtol=0, atol=1e-15 * np.abs(F).sum())
        assert np.array_equal(lat.reduce_field("force", "max", [1])[0], F.max(axis=0))
        assert np.array_equal(lat.reduce_field("force", "min", [1])[0], F.min(axis=0))
        # walls carry no force
        assert np.array_equal(lat.reduce_field("force", "max", [2])[0], np.zeros(3))
    uniform = W.couette_channel((9, 7, 6), tau=0.7, wall_u=0.05, force=1e-5)
    with BlockLattice.from_case(uniform, "f32") as lat:
        got, cells = lat.reduce_field("force", "sum", [1])
        n = int((uniform.ids == 1).sum())
        assert cells == n
        want = np.array([float(np.float32(f)) for f in uniform.force]) * n
        np.testing.assert_allclose(got, want, rtol=1e-12, atol=0)


def test_empty_selection_and_errors(gpu_available):
    case = W.taylor_green((6, 5, 4))
    with BlockLattice.from_case(case, "f32") as lat:
        s, n = lat.reduce_field("density", "sum", [])
        assert n == 0 and s[0] == 0.0
        assert lat.reduce_field("velocity", "min", [])[0][0] == np.finfo(np.float64).max
        assert lat.reduce_field("velocity", "max", [])[0][0] == -np.finfo(np.float64).max
        L = abi.load()
        out = np.zeros(3)
        ptr = out.ctypes.data_as(abi.ctypes.c_void_p)
        with pytest.raises(abi.LbmhError, match="unknown field"):
            abi.check(L.lbmh_reduce_field(lat.handle, 9, 0, None, None, ptr, None))
        with pytest.raises(abi.LbmhError, match="unknown reduction"):
            abi.check(L.lbmh_reduce_field(lat.handle, 0, 7, None, None, ptr, None))
        with pytest.raises(abi.LbmhError, match="conversion"):
            abi.check(L.lbmh_reduce_field(lat.handle, 3, 0, None, None, ptr, None))
