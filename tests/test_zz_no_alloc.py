"""GPU test of lbmh_reduce_moments_noalloc, the moment reduction that allocates
nothing (its partial sums live in the library's fixed per-device workspace).

It must give lbmh_reduce_moments' four FP64 sums bit for bit, through the C
ABI, on two calls in a row of which the second reduces a smaller lattice (the
workspace keeps the first call's partials, which must not leak into the
second), and again on the first lattice afterwards."""
import ctypes

import numpy as np
import pytest

from oracle import oracle as O
from openlb_amd import abi
from openlb_amd import workloads as W
from openlb_amd.lattice import BlockLattice

pytestmark = pytest.mark.gpu


class DeviceDoubles:
    """A few doubles of device memory owned by the test (hipMalloc through ctypes)."""

    def __init__(self, n):
        self.hip = ctypes.CDLL("libamdhip64.so")
        self.hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.hip.hipMemset.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t]
        self.hip.hipFree.argtypes = [ctypes.c_void_p]
        self.n = n
        self.ptr = ctypes.c_void_p()
        assert self.hip.hipMalloc(ctypes.byref(self.ptr), 8 * n) == 0
        assert self.hip.hipMemset(self.ptr, 0xFF, 8 * n) == 0   # NaN pattern: a missed write shows

    def read(self):
        out = np.empty(self.n, dtype=np.float64)
        assert self.hip.hipDeviceSynchronize() == 0
        assert self.hip.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), self.ptr, 8 * self.n, 2) == 0
        return out

    def free(self):
        self.hip.hipFree(self.ptr)


def _reduce_both(lat):
    lib = abi.load()
    ref = lat.reduce_moments()
    dev = DeviceDoubles(4)
    try:
        abi.check(lib.lbmh_reduce_moments_noalloc(lat.handle, dev.ptr, None))
        got = dev.read()
    finally:
        dev.free()
    return np.asarray(ref, dtype=np.float64), got


def test_noalloc_reduction_equals_reduce_moments(gpu_available):
    big = W.porous_channel((70, 64, 60), seed=3, tau=0.6, force=1e-5, rmin=1, rmax=3)   # 1024 workgroups
    small = W.taylor_green((12, 10, 8), u0=0.05, tau=0.6)                             # 4 workgroups
    print("no-alloc reduction: porous channel 70x64x60 f64, then Taylor-Green 12x10x8 f32", flush=True)
    with BlockLattice.from_case(big, "f64") as a, BlockLattice.from_case(small, "f32") as b:
        a.upload_populations(O.initial_populations(big, "f64"), "f64")
        a.collide_and_stream(7)
        b.upload_populations(O.initial_populations(small, "f32"), "f64")
        b.collide_and_stream(5)
        for lat, label in ((a, "large"), (b, "smaller"), (a, "large again")):
            ref, got = _reduce_both(lat)
            print(f"{label}: {got.tolist()}", flush=True)
            assert np.isfinite(got).all()
            np.testing.assert_array_equal(got, ref, err_msg=label)
        # a lattice with only NoDynamics cells sums to zero
        with BlockLattice(19, (5, 4, 3), storage="f64") as empty:
            empty.define_dynamics(np.zeros((5, 4, 3), dtype=np.uint8))
            ref, got = _reduce_both(empty)
            np.testing.assert_array_equal(got, np.zeros(4))
            np.testing.assert_array_equal(ref, np.zeros(4))


def test_noalloc_reduction_rejects_null_output(gpu_available):
    lib = abi.load()
    with BlockLattice(19, (4, 4, 4), storage="f32") as lat:
        with pytest.raises(abi.LbmhError, match="null"):
            abi.check(lib.lbmh_reduce_moments_noalloc(lat.handle, None, None))
