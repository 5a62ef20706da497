"""ctypes binding of librsh_scatter.so (C ABI declared in include/rsh_scatter.h).

The library is loaded from this package directory only. If it is missing the
import fails with an error that names the build command: there is no Python
or CPU fallback for any entry point.
"""
from __future__ import annotations

import ctypes as C
import os

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG_DIR, "librsh_scatter.so")

if not os.path.exists(LIB_PATH):
    raise ImportError(
        f"{LIB_PATH} is not built; run `python -m relativistic_solver_hllc_amd.build` "
        "(hipcc, gfx950) before using relativistic_solver_hllc_amd")

lib = C.CDLL(LIB_PATH)

i32, i64, u32, u64, f64, f32 = C.c_int32, C.c_int64, C.c_uint32, C.c_uint64, C.c_double, C.c_float
vp, cp = C.c_void_p, C.c_char_p
P_i32, P_f64, P_f32, P_vp = C.POINTER(i32), C.POINTER(f64), C.POINTER(f32), C.POINTER(vp)
P_i64 = C.POINTER(i64)


class Tile(C.Structure):
    _fields_ = [("waves", i32), ("chunk_steps", i32)]


class ProblemDesc(C.Structure):
    _fields_ = [
        ("device", i32),
        ("n_cells", i64),
        ("faces", i32),
        ("n_dirs", i32),
        ("dirs_xyz", vp),
        ("weights", vp),
        ("full_area", f64),
        ("k_begin", i32),
        ("k_end", i32),
        ("phase", i32),
        ("g", f64),
        ("normalize", i32),
    ]


class RequestDesc(C.Structure):
    _fields_ = [
        ("n_cells", i64),
        ("n_dirs", i32),
        ("phase", i32),
        ("g", f64),
        ("albedo", f64),
        ("tolerance", f64),
        ("max_iterations", i32),
        ("form", i32),
        ("seed", u64),
    ]


class RequestInfo(C.Structure):
    _fields_ = [
        ("status", i32),
        ("iterations", i32),
        ("residual", f64),
        ("t_submit", f64),
        ("t_admit", f64),
        ("t_done", f64),
        ("footprint_bytes", i64),
    ]


# name -> (restype, argtypes)
PROTOTYPES = {
    "rsh_abi_version": (C.c_int, []),
    "rsh_last_error": (cp, []),
    "rsh_fibonacci_directions": (C.c_int, [i32, vp, vp, P_f64]),
    "rsh_read_directions": (C.c_int, [cp, i32, P_i32, vp, vp, P_f64]),
    "rsh_write_directions": (C.c_int, [cp, i32, vp, vp, f64]),
    "rsh_partition": (C.c_int, [i32, i32, vp, vp]),
    "rsh_device_count": (C.c_int, [P_i32]),
    "rsh_set_device": (C.c_int, [i32]),
    "rsh_dirset_create": (C.c_int, [i32, i32, vp, vp, f64, i32, f64, i32, P_vp]),
    "rsh_dirset_destroy": (C.c_int, [vp]),
    "rsh_dirset_phase_matrix": (C.c_int, [vp, i32, i32, vp]),
    "rsh_scatter_f64_dev": (C.c_int, [vp, vp, i32, i64, i32, i32, vp, i64, vp, i32, C.POINTER(Tile)]),
    "rsh_split_sizes": (C.c_int, [i32, i32, i64, i32, C.POINTER(i64), C.POINTER(i64)]),
    "rsh_pack_split_dev": (C.c_int, [i32, vp, i32, i64, i32, vp, vp, vp]),
    "rsh_scatter_split_dev": (C.c_int, [vp, i32, vp, vp, i32, i64, i32, i32, vp, i64, i32, vp, i32]),
    "rsh_fill_lognormal_dev": (C.c_int, [vp, i64, f64, f64, u64, vp]),
    "rsh_fill_uniform_dev": (C.c_int, [vp, i64, f64, f64, u64, vp]),
    "rsh_autotune_f64": (C.c_int, [i32, i32, i32, i64, C.POINTER(Tile), P_f32]),
    "rsh_tuned_tile_f64": (C.c_int, [i32, i32, C.POINTER(Tile)]),
    "rsh_batch_footprint": (C.c_int, [i32, i64, i32, C.POINTER(i64)]),
    "rsh_batch_create": (C.c_int, [i32, i32, i32, P_vp]),
    "rsh_batch_destroy": (C.c_int, [vp]),
    "rsh_batch_add": (C.c_int, [vp, vp, i64, vp, vp, f64, i32, P_i32]),
    "rsh_batch_add_seeded": (C.c_int, [vp, vp, i64, u64, f64, i32, P_i32]),
    "rsh_batch_reserve": (C.c_int, [vp, i64]),
    "rsh_batch_set_split_kernel": (C.c_int, [vp, i32]),
    "rsh_batch_autotune": (C.c_int, [i32, i32, i32, i64, i32, C.POINTER(Tile), P_f32]),
    "rsh_batch_tuned": (C.c_int, [i32, i32, i32, C.POINTER(Tile)]),
    "rsh_batch_set_f64_tile": (C.c_int, [vp, C.POINTER(Tile)]),
    "rsh_batch_slot_config": (C.c_int, [vp, i32, C.POINTER(Tile)]),
    "rsh_batch_remove": (C.c_int, [vp, i32]),
    "rsh_batch_run": (C.c_int, [vp, i32]),
    "rsh_batch_sync": (C.c_int, [vp]),
    "rsh_batch_residuals": (C.c_int, [vp, vp, vp]),
    "rsh_batch_get_state": (C.c_int, [vp, i32, vp]),
    "rsh_batch_stream": (C.c_int, [vp, P_vp]),
    "rsh_selftest_mfma_f64": (C.c_int, [i32, P_i32]),
    "rsh_dev_alloc": (C.c_int, [i32, i64, P_vp]),
    "rsh_dev_free": (C.c_int, [vp]),
    "rsh_host_alloc_pinned": (C.c_int, [i64, P_vp]),
    "rsh_host_free_pinned": (C.c_int, [vp]),
    "rsh_dev_memcpy": (C.c_int, [vp, vp, i64, i32, vp]),
    "rsh_dev_memcpy2d": (C.c_int, [vp, i64, vp, i64, i64, i64, i32]),
    "rsh_dev_fill_f64": (C.c_int, [vp, i64, f64, vp]),
    "rsh_dev_scale_f64": (C.c_int, [vp, i64, f64, vp]),
    "rsh_dev_copy_f64": (C.c_int, [vp, vp, i64, vp]),
    "rsh_mfma_peak": (C.c_int, [i32, i32, P_f64]),
    "rsh_scatter_f64_dev_noalloc": (C.c_int, [vp, vp, i32, i64, i32, i32, vp, i64, vp]),
    "rsh_noalloc_scratch_bytes": (i64, []),
    "rsh_dev_rel_error": (C.c_int, [vp, i32, i64, vp, i64, i64, i64, f64, P_f64, P_f64, P_i64, P_i64, vp]),
    "rsh_dev_synchronize": (C.c_int, [i32]),
    "rsh_dev_mem_info": (C.c_int, [i32, C.POINTER(i64), C.POINTER(i64)]),
    "rsh_stream_create": (C.c_int, [i32, P_vp]),
    "rsh_stream_destroy": (C.c_int, [vp]),
    "rsh_stream_synchronize": (C.c_int, [vp]),
    "rsh_event_create": (C.c_int, [P_vp]),
    "rsh_event_destroy": (C.c_int, [vp]),
    "rsh_event_record": (C.c_int, [vp, vp]),
    "rsh_event_elapsed_ms": (C.c_int, [vp, vp, P_f32]),
    "rsh_server_create": (C.c_int, [i32, i64, i32, i32, i32, P_vp]),
    "rsh_server_destroy": (C.c_int, [vp]),
    "rsh_server_submit": (C.c_int, [vp, C.POINTER(RequestDesc), P_i32]),
    "rsh_server_cancel": (C.c_int, [vp, i32]),
    "rsh_server_step": (C.c_int, [vp, P_i32, P_i32]),
    "rsh_server_poll": (C.c_int, [vp, i32, C.POINTER(RequestInfo)]),
    "rsh_server_result": (C.c_int, [vp, i32, vp]),
    "rsh_server_stats": (C.c_int, [vp, C.POINTER(i64), P_i32, P_i32, C.POINTER(i64)]),
    "rsh_server_clock": (f64, []),
    "rsh_server_run_trace": (C.c_int, [vp, i32, vp, vp, vp, P_f64]),
    "rsh_create": (C.c_int, [C.POINTER(ProblemDesc), P_vp]),
    "rsh_destroy": (C.c_int, [vp]),
    "rsh_upload_illum": (C.c_int, [vp, vp, i64, i64, i64, i32]),
    "rsh_copy_illum_sync": (C.c_int, [vp, vp, i64]),
    "rsh_scatter": (C.c_int, [vp, i32, i32, i32, vp]),
    "rsh_scatter_from_host": (C.c_int, [vp, vp, i32, i32, vp, i32]),
    "rsh_sync": (C.c_int, [vp, i32]),
    "rsh_problem_buffers": (C.c_int, [vp, P_vp, P_vp]),
    "rsh_problem_stream": (C.c_int, [vp, i32, P_vp]),
    "rsh_event_times": (C.c_int, [vp, P_f32]),
    "rsh_read_simple_bin": (C.c_int, [cp, i32, vp, i64, C.POINTER(i64)]),
    "rsh_pack_split_avg_dev": (C.c_int, [i32, vp, i64, i32, vp, vp, vp]),
    "rsh_write_simple_bin": (C.c_int, [cp, i32, vp, i64]),
    "rsh_moments_f64_dev": (C.c_int, [vp, vp, i32, i64, vp, vp, vp, vp, vp, vp, vp, vp, i32, vp]),
    "rsh_problem_set_geometry": (C.c_int, [vp, vp, vp, vp]),
    "rsh_moments": (C.c_int, [vp, i32, vp, vp, vp, vp, vp]),
    "rsh_problem_moment_buffers": (C.c_int, [vp, P_vp, P_vp, P_vp, P_vp, P_vp]),
    "rsh_write_solution": (C.c_int, [vp, cp]),
    "rsh_mesh_cube_tetra": (C.c_int, [i32, vp, vp]),
    "rsh_tetra_geometry": (C.c_int, [i64, vp, vp, vp, vp, vp]),
    "rsh_sweep_prepare_direction": (C.c_int, [i64, vp, vp, vp, vp, vp, vp, vp, i64, C.POINTER(i64)]),
    "rsh_mesh_vtk_arrays": (C.c_int, [i64, vp, vp, vp, vp, vp, vp]),
    "rsh_make_nodes": (C.c_int, [i64, vp, vp]),
    "rsh_make_direction": (C.c_int, [i64, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, i64, C.POINTER(i64)]),
    "rsh_sweep_create": (C.c_int, [i32, i64, i32, vp, vp, P_vp]),
    "rsh_sweep_destroy": (C.c_int, [vp]),
    "rsh_sweep_set_direction": (C.c_int, [vp, i32, vp, vp, vp, vp, i64]),
    "rsh_sweep_levels": (C.c_int, [vp, i32, P_i32]),
    "rsh_sweep_run": (C.c_int, [vp, i32, vp, i64, vp, vp, vp, vp]),
}

for _name, (_res, _args) in PROTOTYPES.items():
    _fn = getattr(lib, _name)  # AttributeError here = symbol missing from the build
    _fn.restype = _res
    _fn.argtypes = _args

STATUS = {0: "RSH_OK", 1: "RSH_E_ARG", 2: "RSH_E_HIP", 3: "RSH_E_NOMEM", 4: "RSH_E_STATE", 5: "RSH_E_IO",
          6: "RSH_E_UNSUPPORTED"}


class RshError(RuntimeError):
    def __init__(self, status: int, where: str):
        msg = lib.rsh_last_error().decode(errors="replace")
        super().__init__(f"{where}: {STATUS.get(status, status)}: {msg}")
        self.status = status


def check(status: int, where: str) -> None:
    if status != 0:
        raise RshError(status, where)


def call(name: str, *args) -> None:
    check(getattr(lib, name)(*args), name)
