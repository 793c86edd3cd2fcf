"""ctypes binding of the C ABI in include/lbmh.h (library openlb_amd/_lib/liblbmh.so).

Loading fails with an exception if the HIP library has not been built. The
Python layer never computes a time step itself.
"""
from __future__ import annotations

import ctypes
import os

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "_lib", "liblbmh.so")

F64, F32, F16 = 0, 1, 2
STORAGE = {"f64": F64, "f32": F32, "f16": F16}
KINDS = {"bgk": 0, "forced_bgk": 1, "trt": 2, "forced_trt": 3, "smagorinsky": 4}

OK, E_INVALID, E_DEVICE, E_MEMORY, E_STATE = 0, -1, -2, -3, -4

# every function include/lbmh.h declares
EXPORTS = (
    "lbmh_lattice_create", "lbmh_lattice_destroy", "lbmh_set_dynamics_ids",
    "lbmh_set_bulk_params", "lbmh_set_force_field", "lbmh_set_wall", "lbmh_populations_upload",
    "lbmh_populations_download", "lbmh_column_upload", "lbmh_column_download", "lbmh_ini_equilibrium", "lbmh_step",
    "lbmh_get_step_count", "lbmh_get_shifts", "lbmh_checkpoint_size", "lbmh_checkpoint_save",
    "lbmh_checkpoint_load", "lbmh_field_export", "lbmh_moments", "lbmh_reduce_field", "lbmh_reduce_moments",
    "lbmh_reduce_moments_noalloc", "lbmh_statistics",
    "lbmh_set_local_boundary", "lbmh_classify_boundary", "lbmh_initialize",
    "lbmh_classify_boundary_overlap", "lbmh_set_overlap_ids", "lbmh_halo_copy_many",
    "lbmh_collide", "lbmh_stream", "lbmh_halo_create", "lbmh_halo_destroy", "lbmh_halo_copy",
    "lbmh_synchronize", "lbmh_device_synchronize", "lbmh_lattice_footprint", "lbmh_abi_version", "lbmh_device_count",
    "lbmh_device_arch", "lbmh_last_error", "lbmh_set_launch_config", "lbmh_get_launch_config",
    "lbmh_autotune", "lbmh_pattern_ceiling",
    "lbmh_ensemble_create", "lbmh_ensemble_destroy", "lbmh_ensemble_member_bytes", "lbmh_ensemble_add",
    "lbmh_ensemble_remove", "lbmh_ensemble_set_wall", "lbmh_ensemble_set_local_boundary",
    "lbmh_ensemble_initialize", "lbmh_ensemble_set_steps", "lbmh_ensemble_status",
    "lbmh_ensemble_ini_equilibrium", "lbmh_ensemble_upload", "lbmh_ensemble_download",
    "lbmh_ensemble_moments", "lbmh_ensemble_field_export", "lbmh_ensemble_checkpoint_size", "lbmh_ensemble_checkpoint_save",
    "lbmh_ensemble_checkpoint_load", "lbmh_ensemble_step", "lbmh_ensemble_step_graph",
    "lbmh_ensemble_set_launch_config", "lbmh_ensemble_autotune", "lbmh_ensemble_usage", "lbmh_ensemble_status_all",
)


class LatticeDesc(ctypes.Structure):
    _fields_ = [
        ("q", ctypes.c_int),
        ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
        ("periodic", ctypes.c_int * 3),
        ("storage", ctypes.c_int),
        ("bulk_kind", ctypes.c_int),
        ("device", ctypes.c_int),
    ]


class BulkParams(ctypes.Structure):
    _fields_ = [
        ("omega", ctypes.c_double),
        ("magic", ctypes.c_double),
        ("smago", ctypes.c_double),
        ("force", ctypes.c_double * 3),
    ]


class LaunchConfig(ctypes.Structure):
    _fields_ = [("lds_bytes", ctypes.c_int), ("xcd_group", ctypes.c_int)]


class EnsembleDesc(ctypes.Structure):
    _fields_ = [
        ("q", ctypes.c_int), ("storage", ctypes.c_int), ("bulk_kind", ctypes.c_int),
        ("device", ctypes.c_int), ("max_members", ctypes.c_int),
        ("population_bytes", ctypes.c_size_t), ("cell_capacity", ctypes.c_size_t),
        ("boundary_capacity", ctypes.c_size_t),
    ]


class LbmhError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"lbmh error {code}: {message}")
        self.code = code


_lib = None


def load() -> ctypes.CDLL:
    """Load liblbmh.so. It raises if the library is missing, and there is no fallback."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"HIP backend library not built: {LIB_PATH} is missing "
            "(run `python -m openlb_amd.build` or __graft_entry__.build())")
    L = ctypes.CDLL(LIB_PATH)
    vp, i32, i64 = ctypes.c_void_p, ctypes.c_int, ctypes.c_int64
    L.lbmh_lattice_create.argtypes = [ctypes.POINTER(LatticeDesc), ctypes.POINTER(vp)]
    L.lbmh_lattice_destroy.argtypes = [vp]
    L.lbmh_set_dynamics_ids.argtypes = [vp, vp]
    L.lbmh_set_bulk_params.argtypes = [vp, ctypes.POINTER(BulkParams)]
    L.lbmh_set_force_field.argtypes = [vp, vp]
    L.lbmh_reduce_field.argtypes = [vp, i32, i32, vp, vp, vp, vp]
    L.lbmh_set_wall.argtypes = [vp, i32, ctypes.c_double, ctypes.POINTER(ctypes.c_double)]
    L.lbmh_populations_upload.argtypes = [vp, vp, i32]
    L.lbmh_populations_download.argtypes = [vp, vp, i32]
    L.lbmh_column_upload.argtypes = [vp, i32, vp, i64, i32, i32]
    L.lbmh_column_download.argtypes = [vp, i32, vp, i64, i32, i32]
    L.lbmh_ini_equilibrium.argtypes = [vp, vp]
    L.lbmh_step.argtypes = [vp, i64, vp]
    L.lbmh_get_step_count.argtypes = [vp, ctypes.POINTER(i64)]
    L.lbmh_get_shifts.argtypes = [vp, vp]
    L.lbmh_field_export.argtypes = [vp, i32, i32, vp, vp, ctypes.c_size_t]
    L.lbmh_ensemble_field_export.argtypes = [vp, i32, i32, i32, vp, vp, ctypes.c_size_t]
    L.lbmh_checkpoint_size.argtypes = [vp, i32, ctypes.POINTER(ctypes.c_size_t)]
    L.lbmh_checkpoint_save.argtypes = [vp, i32, vp, ctypes.c_size_t]
    L.lbmh_checkpoint_load.argtypes = [vp, i32, vp, ctypes.c_size_t]
    L.lbmh_ensemble_checkpoint_size.argtypes = [vp, i32, i32, ctypes.POINTER(ctypes.c_size_t)]
    L.lbmh_ensemble_checkpoint_save.argtypes = [vp, i32, i32, vp, ctypes.c_size_t]
    L.lbmh_ensemble_checkpoint_load.argtypes = [vp, i32, i32, vp, ctypes.c_size_t]
    L.lbmh_moments.argtypes = [vp, vp]
    L.lbmh_reduce_moments.argtypes = [vp, vp]
    L.lbmh_reduce_moments_noalloc.argtypes = [vp, vp, vp]
    L.lbmh_statistics.argtypes = [vp, vp]
    L.lbmh_set_local_boundary.argtypes = [vp, i32, i32, ctypes.c_double, vp]
    L.lbmh_classify_boundary.argtypes = [vp, vp, vp, vp]
    L.lbmh_classify_boundary_overlap.argtypes = [vp, vp, vp, vp, vp]
    L.lbmh_set_overlap_ids.argtypes = [vp, vp]
    L.lbmh_halo_copy_many.argtypes = [vp, ctypes.c_size_t, vp]
    L.lbmh_initialize.argtypes = [vp]
    L.lbmh_collide.argtypes = [vp, vp]
    L.lbmh_stream.argtypes = [vp, vp]
    L.lbmh_halo_create.argtypes = [vp, vp, vp, vp, ctypes.c_size_t, ctypes.POINTER(vp)]
    L.lbmh_halo_destroy.argtypes = [vp]
    L.lbmh_halo_copy.argtypes = [vp, vp]
    L.lbmh_synchronize.argtypes = [vp]
    L.lbmh_device_synchronize.argtypes = [i32]
    L.lbmh_lattice_footprint.argtypes = [vp, ctypes.POINTER(ctypes.c_size_t)]
    L.lbmh_device_count.argtypes = [ctypes.POINTER(i32)]
    L.lbmh_device_arch.argtypes = [i32, ctypes.c_char_p, ctypes.c_size_t]
    L.lbmh_set_launch_config.argtypes = [vp, ctypes.POINTER(LaunchConfig)]
    L.lbmh_get_launch_config.argtypes = [vp, ctypes.POINTER(LaunchConfig)]
    L.lbmh_autotune.argtypes = [vp, i32, ctypes.POINTER(LaunchConfig)]
    L.lbmh_pattern_ceiling.argtypes = [vp, i64, vp]
    i3 = ctypes.c_int * 3
    szp = ctypes.POINTER(ctypes.c_size_t)
    L.lbmh_ensemble_create.argtypes = [ctypes.POINTER(EnsembleDesc), ctypes.POINTER(vp)]
    L.lbmh_ensemble_destroy.argtypes = [vp]
    L.lbmh_ensemble_member_bytes.argtypes = [vp, i3, szp, szp]
    L.lbmh_ensemble_add.argtypes = [vp, i3, i3, vp, ctypes.POINTER(BulkParams), i64,
                                    ctypes.POINTER(i32)]
    L.lbmh_ensemble_remove.argtypes = [vp, i32]
    L.lbmh_ensemble_set_wall.argtypes = [vp, i32, i32, ctypes.c_double, ctypes.POINTER(ctypes.c_double)]
    L.lbmh_ensemble_set_local_boundary.argtypes = [vp, i32, i32, i32, ctypes.c_double, vp]
    L.lbmh_ensemble_initialize.argtypes = [vp, i32]
    L.lbmh_ensemble_set_steps.argtypes = [vp, i32, i64]
    L.lbmh_ensemble_status.argtypes = [vp, i32, ctypes.POINTER(i64), ctypes.POINTER(i64)]
    L.lbmh_ensemble_ini_equilibrium.argtypes = [vp, i32, vp]
    L.lbmh_ensemble_upload.argtypes = [vp, i32, vp, i32]
    L.lbmh_ensemble_download.argtypes = [vp, i32, vp, i32]
    L.lbmh_ensemble_moments.argtypes = [vp, i32, vp]
    L.lbmh_ensemble_step.argtypes = [vp, i64, vp]
    L.lbmh_ensemble_step_graph.argtypes = [vp, i64, vp]
    L.lbmh_ensemble_set_launch_config.argtypes = [vp, ctypes.POINTER(LaunchConfig)]
    L.lbmh_ensemble_autotune.argtypes = [vp, i32, ctypes.POINTER(LaunchConfig)]
    L.lbmh_ensemble_usage.argtypes = [vp, ctypes.POINTER(i32), szp, szp, szp]
    L.lbmh_ensemble_status_all.argtypes = [vp, vp, vp]
    L.lbmh_last_error.restype = ctypes.c_char_p
    for name in EXPORTS:
        if name != "lbmh_last_error":
            getattr(L, name).restype = i32
    _lib = L
    return L


def check(code: int):
    if code != OK:
        raise LbmhError(code, load().lbmh_last_error().decode())


def device_count() -> int:
    n = ctypes.c_int(0)
    check(load().lbmh_device_count(ctypes.byref(n)))
    return n.value


def device_arch(device: int = 0) -> str:
    buf = ctypes.create_string_buffer(64)
    check(load().lbmh_device_arch(device, buf, 64))
    return buf.value.decode()
