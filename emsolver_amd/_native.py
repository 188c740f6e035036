"""Loader for the in-tree HIP library ``emsolver_amd/lib/libemsolver_tlm.so``.

There is no fallback: if the library has not been built (``python -c
"import __graft_entry__ as g; g.build()"``) the import of the product path
fails with an explicit error instead of silently computing on the CPU.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libemsolver_tlm.so")

_lib = None

P = C.c_void_p
i64p = C.POINTER(C.c_int64)
i32p = C.POINTER(C.c_int32)
fp = C.POINTER(C.c_float)
dp = C.POINTER(C.c_double)
u8p = C.POINTER(C.c_uint8)


class GridDesc(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32), ("dl", C.c_float),
                ("rho", C.c_float), ("pulse_type", C.c_int32), ("device", C.c_int32),
                ("fp32_halo", C.c_int32)]


class BatchTuning(C.Structure):
    _fields_ = [("planes", C.c_int32), ("block_waves", C.c_int32), ("grid_blocks", C.c_int32),
                ("graph_ticks", C.c_int32), ("planes16", C.c_int32)]


class VoxelGrid(C.Structure):
    _fields_ = [("origin", C.c_float * 3), ("dl", C.c_float)]


class FarFieldSummary(C.Structure):
    _fields_ = [("u_max", C.c_float), ("directivity", C.c_float),
                ("peak_theta_deg", C.c_float), ("peak_phi_deg", C.c_float)]


class Tuning(C.Structure):
    _fields_ = [("kernel", C.c_int32), ("planes", C.c_int32), ("block_waves", C.c_int32),
                ("nontemporal", C.c_int32), ("grid_blocks", C.c_int32), ("fold", C.c_int32),
                ("occupancy", C.c_int32)]


# Every symbol declared in include/GpuTlmHip.h with its ctypes signature.
SIGNATURES = {
    "tlm_device_count": (C.c_int, []),
    "tlm_device_name": (C.c_char_p, [C.c_int]),
    "tlm_create": (C.c_int, [C.POINTER(GridDesc), C.POINTER(P)]),
    "tlm_destroy": (None, [P]),
    "tlm_last_error": (C.c_char_p, [P]),
    "tlm_upload": (C.c_int, [P, u8p, fp, fp, C.c_int, fp, fp]),
    "tlm_set_sources": (C.c_int, [P, i64p, i32p, fp, C.c_int]),
    "tlm_set_schedule": (C.c_int, [P, fp, fp, fp, C.c_int, C.c_int]),
    "tlm_set_probes": (C.c_int, [P, i64p, C.c_int, C.c_float]),
    "tlm_set_probe_dft": (C.c_int, [P, dp, dp, C.c_int]),
    "tlm_set_ports": (C.c_int, [P, i64p, i32p, i32p, C.c_int]),
    "tlm_set_faces": (C.c_int, [P, i64p, i32p, C.c_int]),
    "tlm_set_huygens": (C.c_int, [P, i64p, i32p, C.c_int]),
    "tlm_set_energy_every": (C.c_int, [P, C.c_int]),
    "tlm_run": (C.c_int, [P, C.c_int, C.c_int]),
    "tlm_sync": (C.c_int, [P]),
    "tlm_read_state": (C.c_int, [P, fp, fp]),
    "tlm_read_probes": (C.c_int, [P, C.c_int, C.c_int, fp]),
    "tlm_read_probe_dft": (C.c_int, [P, dp, dp]),
    "tlm_read_ports": (C.c_int, [P, C.c_int, C.c_int, fp]),
    "tlm_read_js": (C.c_int, [P, fp, fp]),
    "tlm_read_huygens": (C.c_int, [P, fp]),
    "tlm_read_energy": (C.c_int, [P, i32p, dp, C.c_int]),
    "tlm_time_steps": (C.c_int, [P, C.c_int, C.c_int, C.c_int, fp, fp, C.POINTER(C.c_int)]),
    "tlm_launch_shape": (C.c_int, [P, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                   C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "tlm_set_tuning": (C.c_int, [P, C.POINTER(Tuning)]),
    "tlm_get_tuning": (C.c_int, [P, C.POINTER(Tuning)]),
    "tlm_copy_bandwidth": (C.c_int, [C.c_int, C.c_size_t, C.c_int, fp, fp]),
    "tlm_scatter_unitarity_error": (C.c_float, [C.c_float]),
    "tlm_build_pec_faces": (C.c_int64, [C.c_int, C.c_int, C.c_int, u8p, i64p, i32p, i32p,
                                        C.c_int64]),
    "tlm_build_huygens": (C.c_int64, [C.c_int, C.c_int, C.c_int, u8p, C.c_int, i64p, i32p,
                                      i32p, C.c_int64]),
    "tlm_dft_tables": (None, [dp, C.c_int, C.c_double, C.c_int, dp, dp]),
    "tlm_schedule_tables": (None, [C.c_int, C.c_float, C.c_float, C.c_int, C.c_int, C.c_float,
                                   C.c_float, fp, fp, fp]),
    "tlm_set_plane": (C.c_int, [P, C.c_int, C.c_int]),
    "tlm_set_frames": (C.c_int, [P, C.c_int, C.c_int]),
    "tlm_read_plane": (C.c_int, [P, fp, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "tlm_read_frames": (C.c_int, [P, i32p, fp, fp, C.c_int]),
    "tlm_read_scales": (C.c_int, [P, fp, fp, i32p]),
    "tlm_batch_create": (C.c_int, [C.c_int, C.POINTER(P)]),
    "tlm_batch_destroy": (None, [P]),
    "tlm_batch_last_error": (C.c_char_p, [P]),
    "tlm_batch_add": (C.c_int, [P, P, C.c_int]),
    "tlm_batch_remove": (C.c_int, [P, P]),
    "tlm_batch_run": (C.c_int, [P, C.c_int]),
    "tlm_batch_sync": (C.c_int, [P]),
    "tlm_batch_members": (C.c_int, [P, C.POINTER(P), i32p, C.c_int]),
    "tlm_batch_set_tuning": (C.c_int, [P, C.POINTER(BatchTuning)]),
    "tlm_batch_get_tuning": (C.c_int, [P, C.POINTER(BatchTuning)]),
    "tlm_batch_time": (C.c_int, [P, C.c_int, fp, i64p]),
    "tlm_batch_autotune": (C.c_int, [P, C.c_int, C.POINTER(BatchTuning)]),
    "tlm_device_memory": (C.c_int, [C.c_int, C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]),
    "tlm_estimate_bytes": (C.c_int64, [C.POINTER(GridDesc), C.c_int]),
    "tlm_device_encode": (C.c_int, [C.c_int, C.c_int, fp, C.c_float, C.POINTER(C.c_uint16),
                                    C.c_int64]),
    "tlm_mat_fill": (C.c_int, [P, C.c_uint8]),
    "tlm_mat_cells_box": (C.c_int, [P, i32p, i32p, C.c_uint8]),
    "tlm_voxelize_mesh": (C.c_int, [P, C.POINTER(VoxelGrid), fp, C.c_int32, i32p, C.c_int32, fp,
                                    C.c_uint8, C.c_int]),
    "tlm_voxelize_wire": (C.c_int, [P, C.POINTER(VoxelGrid), fp, C.c_int32, C.c_float, fp,
                                    C.c_uint8]),
    "tlm_carve_feed": (C.c_int, [P, C.POINTER(VoxelGrid), fp, fp, fp, i64p, C.c_int32, i32p]),
    "tlm_mat_read": (C.c_int, [P, u8p]),
    "tlm_mat_write": (C.c_int, [P, u8p]),
    "tlm_upload_mat": (C.c_int, [P, fp, fp, C.c_int]),
    "tlm_pec_faces_mat": (C.c_int64, [P]),
    "tlm_huygens_mat": (C.c_int64, [P, C.c_int]),
    "tlm_read_faces": (C.c_int64, [P, i64p, i32p, C.c_int64]),
    "tlm_read_huygens_faces": (C.c_int64, [P, i64p, i32p, C.c_int64]),
    "tlm_far_field": (C.c_int, [C.c_int, C.c_int, fp, i32p, i32p, fp, C.c_int, C.c_float,
                                C.c_int, C.c_int, fp, C.POINTER(FarFieldSummary)]),
    "tlm_far_field_ctx": (C.c_int, [P, fp, C.c_int, C.c_float, C.c_int, C.c_int, fp,
                                    C.POINTER(FarFieldSummary)]),
    "tlm_far_field_ctx_noalloc": (C.c_int, [P, fp, C.c_int, C.c_float, C.c_int, C.c_int, fp,
                                    C.POINTER(FarFieldSummary)]),
}


class NativeLibraryMissing(ImportError):
    pass


def load():
    """Load and type the C ABI. Raises NativeLibraryMissing if not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise NativeLibraryMissing(
            f"{LIB_PATH} is missing: build the HIP library first "
            "(python -c 'import __graft_entry__ as g; g.build()')")
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def ptr(a, ct):
    return None if a is None else a.ctypes.data_as(C.POINTER(ct))
