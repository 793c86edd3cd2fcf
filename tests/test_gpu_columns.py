"""Population columns in OpenLB's host storage order (lbmh_column_upload /
lbmh_column_download): the copies a GPU_HIP block makes on a processing-context
switch.  The host layout is OpenLB's CyclicColumn over a block with overlap p
(src/core/platform/cpu/sisd/column.h:160-170, blockStructure.h:65-78),
restated here with numpy: padded cell (x, y, z), -p <= x < nx + p, has id
((x+p) (ny+2p) + (y+p)) (nz+2p) + (z+p) and sits at element (id + shift) mod count."""
import numpy as np
import pytest

from openlb_amd import workloads as W
from openlb_amd.lattice import BlockLattice
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def host_ids(shape, p, lo=0):
    """ids of the padded cells (x, y, z) in [lo, n + ...) per axis, host padding p"""
    nx, ny, nz = shape
    P = (nx + 2 * p, ny + 2 * p, nz + 2 * p)
    x, y, z = np.meshgrid(np.arange(-1, nx + 1), np.arange(-1, ny + 1), np.arange(-1, nz + 1), indexing="ij")
    return ((x + p) * P[1] + (y + p)) * P[2] + (z + p), int(np.prod(P))


@pytest.mark.parametrize("storage", ["f64", "f32", "f16"])
@pytest.mark.parametrize("padding", [1, 3])
def test_column_upload_places_openlb_cells(gpu_available, storage, padding):
    case = W.random_periodic((7, 5, 6), seed=11)
    rng = np.random.default_rng(3)
    real = np.float64 if storage == "f64" else np.float32
    ids, count = host_ids(case.shape, padding)
    core = ids[1:-1, 1:-1, 1:-1]
    with BlockLattice.from_case(case, storage) as lat:
        want = np.empty((case.q,) + case.shape)
        for i in range(case.q):
            column = (rng.standard_normal(count) * 1e-2).astype(real)
            shift = int(rng.integers(-count + 1, count))
            lat.upload_column(i, column, shift, padding)
            want[i] = column[(core + shift) % count]
        got = lat.download_populations("f64")
    tol = 2e-5 if storage == "f16" else 0.0  # FP16 keeps f * 2^12 to 11 bits: |f| <= 0.05 rounds within 1.6e-5
    np.testing.assert_allclose(got, want, rtol=0, atol=tol)


@pytest.mark.parametrize("storage", ["f64", "f32"])
@pytest.mark.parametrize("padding", [1, 3])
def test_column_round_trip_after_steps(gpu_available, storage, padding):
    """Upload from OpenLB's layout, step on the device, download into a host
    column with its own shift: the core cells and the first overlap layer equal
    the device state, every other host entry is left untouched."""
    case = W.random_periodic((6, 7, 5), seed=5, q=27, kind="trt")
    real = np.float64 if storage == "f64" else np.float32
    rng = np.random.default_rng(8)
    ids, count = host_ids(case.shape, padding)
    f0 = O.initial_populations(case, storage)
    with BlockLattice.from_case(case, storage) as lat:
        lat.upload_populations(f0, "f64")
        lat.collide_and_stream(7)
        ref = lat.download_populations("f64")
        host_shift = int(rng.integers(-count + 1, count))
        columns = []
        for i in range(case.q):
            column = np.full(count, -7.0, dtype=real)
            lat.download_column(i, column, host_shift, padding)
            columns.append(column)
        core = ids[1:-1, 1:-1, 1:-1]
        for i in range(case.q):
            np.testing.assert_array_equal(columns[i][(core + host_shift) % count], ref[i].astype(real))
            touched = np.zeros(count, dtype=bool)
            touched[(ids.ravel() + host_shift) % count] = True
            assert (columns[i][~touched] == -7.0).all()
        # and back: a fresh lattice loaded from these host columns continues identically
        with BlockLattice.from_case(case, storage) as other:
            for i in range(case.q):
                other.upload_column(i, columns[i], host_shift, padding)
            other.collide_and_stream(3)
            lat.collide_and_stream(3)
            np.testing.assert_array_equal(other.download_populations("f64"), lat.download_populations("f64"))


def test_column_argument_errors(gpu_available):
    from openlb_amd import abi
    case = W.taylor_green((4, 4, 4))
    with BlockLattice.from_case(case, "f32") as lat:
        count = 6 ** 3
        col = np.zeros(count, dtype=np.float32)
        with pytest.raises(abi.LbmhError, match="padding"):
            lat.upload_column(0, col, 0, 0)
        with pytest.raises(abi.LbmhError, match="shift"):
            lat.upload_column(0, col, count, 1)
        with pytest.raises(abi.LbmhError, match="population index"):
            lat.upload_column(19, col, 0, 1)
