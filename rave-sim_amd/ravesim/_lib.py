"""ctypes binding of librsim.so (declarations: include/ravesim.h).

The library is built in-tree next to this file (`make -C rave-sim_amd`).
There is no fallback: if the shared library is missing, importing this module
raises, so nothing can silently run a CPU path in its place.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "librsim.so"
CLI_PATH = HERE / "fastwave"

RS_OK = 0
RS_ERR_INVALID = 1
RS_ERR_CONFIG = 2
RS_ERR_IO = 3
RS_ERR_DEVICE = 4
RS_ERR_NOMEM = 5
RS_ERR_UNSUPPORTED = 6
RS_C8 = 8
RS_C16 = 16
RS_MODE_FUSED = 0
RS_MODE_PLAIN = 1
RS_NOALLOC_MAX_PIXELS = 65536

# every symbol declared in include/ravesim.h
EXPORTS = (
    "rs_version", "rs_last_error", "rs_ctx_create", "rs_ctx_destroy", "rs_ctx_synchronize",
    "rs_run_simdir", "rs_validate_simdir", "rs_field_create", "rs_field_destroy",
    "rs_field_upload", "rs_field_download", "rs_field_copy", "rs_point_source", "rs_propagate",
    "rs_apply_grating", "rs_apply_dba_grating", "rs_apply_sample", "rs_scale", "rs_square_and_downsample", "rs_square_and_downsample_noalloc",
    "rs_detector_pixels",
    "rs_detector_bounds", "rs_batch_run", "rs_batch_run_mixed", "rs_set_timing", "rs_get_pass_stats", "rs_copy_bandwidth", "rs_autotune", "rs_set_split",
    "rs_tune_get", "rs_tune_save", "rs_tune_load",
    "rs_server_create", "rs_server_submit", "rs_server_poll", "rs_server_wait", "rs_server_stats",
    "rs_server_destroy",
)


class RsError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[rs error {code}] {msg}")
        self.code = code


class rs_run_options(C.Structure):
    _fields_ = [("dtype", C.c_int), ("mode", C.c_int), ("write_output", C.c_int),
                ("reserved", C.c_int), ("detected", C.POINTER(C.c_double)),
                ("detected_capacity", C.c_int64), ("shape_out", C.POINTER(C.c_int64)),
                ("history_dz", C.c_double), ("keypoints_dir", C.c_char_p)]


class rs_grid(C.Structure):
    _fields_ = [("N", C.c_int64), ("dx", C.c_double), ("chunk_size", C.c_int64)]


class rs_grating(C.Structure):
    _fields_ = [("pitch", C.c_double), ("dc0", C.c_double), ("dc1", C.c_double),
                ("z_start", C.c_double), ("thickness", C.c_double), ("nr_steps", C.c_int32),
                ("nr_x_positions", C.c_int32), ("x_positions", C.POINTER(C.c_double)),
                ("substrate_thickness", C.c_double),
                ("db_a_re", C.c_double), ("db_a_im", C.c_double),
                ("db_b_re", C.c_double), ("db_b_im", C.c_double),
                ("db_s_re", C.c_double), ("db_s_im", C.c_double)]


class rs_dba_grating(C.Structure):
    _fields_ = [("pitch0", C.c_double), ("pitch1", C.c_double), ("dc0", C.c_double * 2),
                ("dc1", C.c_double * 2), ("z_start", C.c_double), ("thickness", C.c_double),
                ("nr_steps", C.c_int32), ("nr_x_positions", C.c_int32),
                ("x_positions", C.POINTER(C.c_double)), ("substrate_thickness", C.c_double),
                ("db_a_re", C.c_double), ("db_a_im", C.c_double),
                ("db_b_re", C.c_double), ("db_b_im", C.c_double),
                ("db_s_re", C.c_double), ("db_s_im", C.c_double)]


class rs_sample(C.Structure):
    _fields_ = [("z_start", C.c_double), ("pixel_size_x", C.c_double), ("pixel_size_z", C.c_double),
                ("rows", C.c_int64), ("cols", C.c_int64), ("grid", C.POINTER(C.c_int8)),
                ("nr_materials", C.c_int32), ("nr_x_positions", C.c_int32),
                ("deltabeta", C.POINTER(C.c_double)), ("x_positions", C.POINTER(C.c_double))]


class rs_detector(C.Structure):
    _fields_ = [("size", C.c_double), ("pixel_x", C.c_double), ("pixel_y", C.c_double)]


class rs_batch_member(C.Structure):
    _fields_ = [("wl", C.c_double), ("weight", C.c_double), ("source_kind", C.c_int32),
                ("out_index", C.c_int32), ("source_x", C.c_double), ("source_z", C.c_double),
                ("gratings", C.POINTER(rs_grating)), ("cutoff_angles", C.POINTER(C.c_double))]


class rs_batch_desc(C.Structure):
    _fields_ = [("grid", rs_grid), ("det", rs_detector), ("z_detector", C.c_double),
                ("dtype", C.c_int32), ("mode", C.c_int32), ("nr_gratings", C.c_int32),
                ("nr_members", C.c_int32), ("members", C.POINTER(rs_batch_member)),
                ("nr_outputs", C.c_int32), ("max_resident", C.c_int32),
                ("use_graph", C.c_int32), ("lanes", C.c_int32), ("spans", C.c_int32)]


class rs_batch_sim(C.Structure):
    _fields_ = [("grid", rs_grid), ("det", rs_detector), ("z_detector", C.c_double),
                ("nr_gratings", C.c_int32), ("nr_members", C.c_int32),
                ("members", C.POINTER(rs_batch_member)), ("out_offset", C.c_int64)]


class rs_mixed_desc(C.Structure):
    _fields_ = [("dtype", C.c_int32), ("mode", C.c_int32), ("nr_sims", C.c_int32),
                ("use_graph", C.c_int32), ("sims", C.POINTER(rs_batch_sim))]


class rs_request(C.Structure):
    _fields_ = [("grid", rs_grid), ("det", rs_detector), ("z_detector", C.c_double),
                ("dtype", C.c_int32), ("nr_gratings", C.c_int32), ("nr_energies", C.c_int32),
                ("source_kind", C.c_int32), ("energies_ev", C.POINTER(C.c_double)),
                ("weights", C.POINTER(C.c_double)), ("gratings", C.POINTER(rs_grating)),
                ("cutoff_angles", C.POINTER(C.c_double)), ("source_x", C.c_double),
                ("source_z", C.c_double)]


class rs_tune_entry(C.Structure):
    _fields_ = [("N", C.c_int64), ("dtype", C.c_int32), ("members", C.c_int32), ("lgR", C.c_int32),
                ("lgA", C.c_int32), ("tileA", C.c_int32), ("tileB", C.c_int32), ("ms_per_slice", C.c_double)]


class rs_pass_stat(C.Structure):
    _fields_ = [("launches", C.c_int64), ("total_ms", C.c_double), ("bytes", C.c_double)]


# include/ravesim.h RS_VERSION: the struct layouts declared above
VERSION = 300


def _load() -> C.CDLL:
    if not LIB_PATH.exists():
        raise ImportError(f"{LIB_PATH} is not built; run `make -C {HERE.parent}` (hipcc, gfx950)")
    lib = C.CDLL(str(LIB_PATH))
    P = C.c_void_p
    sig = {
        "rs_version": (C.c_int, []),
        "rs_last_error": (C.c_char_p, []),
        "rs_ctx_create": (C.c_int, [C.c_int, C.POINTER(P)]),
        "rs_ctx_destroy": (C.c_int, [P]),
        "rs_ctx_synchronize": (C.c_int, [P]),
        "rs_run_simdir": (C.c_int, [P, C.c_char_p, C.c_int, C.POINTER(rs_run_options)]),
        "rs_validate_simdir": (C.c_int, [C.c_char_p, C.c_int, C.POINTER(C.c_int64), C.c_int64]),
        "rs_field_create": (C.c_int, [P, C.POINTER(rs_grid), C.c_int, C.c_int, C.POINTER(P)]),
        "rs_field_destroy": (C.c_int, [P]),
        "rs_field_upload": (C.c_int, [P, P]),
        "rs_field_download": (C.c_int, [P, P]),
        "rs_field_copy": (C.c_int, [P, P]),
        "rs_point_source": (C.c_int, [P, C.c_double, C.c_double, C.c_double]),
        "rs_propagate": (C.c_int, [P, C.c_double, C.c_double, C.c_double]),
        "rs_apply_grating": (C.c_int, [P, C.c_double, C.POINTER(rs_grating), C.c_int, C.c_double]),
        "rs_apply_dba_grating": (C.c_int, [P, C.c_double, C.POINTER(rs_dba_grating), C.c_int,
                                           C.c_double]),
        "rs_apply_sample": (C.c_int, [P, C.c_double, C.POINTER(rs_sample), C.c_int, C.c_double]),
        "rs_scale": (C.c_int, [P, C.c_double, C.c_double]),
        "rs_square_and_downsample": (C.c_int, [P, C.POINTER(rs_detector), C.c_double,
                                               C.POINTER(C.c_double), C.c_int64]),
        "rs_square_and_downsample_noalloc": (C.c_int, [P, C.POINTER(rs_detector), C.c_double,
                                                       C.POINTER(C.c_double), C.c_int64]),
        "rs_detector_pixels": (C.c_int64, [C.POINTER(rs_detector)]),
        "rs_detector_bounds": (C.c_int, [C.c_int64, C.c_double, C.POINTER(rs_detector),
                                         C.POINTER(C.c_int64), C.c_int64]),
        "rs_batch_run": (C.c_int, [P, C.POINTER(rs_batch_desc), C.POINTER(C.c_double), C.c_int64]),
        "rs_batch_run_mixed": (C.c_int, [P, C.POINTER(rs_mixed_desc), C.POINTER(C.c_double), C.c_int64]),
        "rs_set_timing": (C.c_int, [P, C.c_int]),
        "rs_autotune": (C.c_int, [P, C.c_int64, C.c_int, C.c_int, C.POINTER(C.c_int32), C.POINTER(C.c_double)]),
        "rs_set_split": (C.c_int, [P, C.c_int64, C.c_int, C.c_int, C.c_int]),
        "rs_tune_get": (C.c_int, [P, C.c_int64, C.c_int, C.c_int, C.POINTER(rs_tune_entry), C.POINTER(C.c_int32)]),
        "rs_tune_save": (C.c_int, [P, C.c_char_p]),
        "rs_tune_load": (C.c_int, [P, C.c_char_p, C.POINTER(C.c_int32)]),
        "rs_server_create": (C.c_int, [P, C.c_double, C.c_int, C.POINTER(P)]),
        "rs_server_submit": (C.c_int, [P, C.POINTER(rs_request), C.POINTER(C.c_int64)]),
        "rs_server_poll": (C.c_int, [P, C.c_int64, C.POINTER(C.c_int32), C.POINTER(C.c_double), C.c_int64,
                                     C.POINTER(C.c_int64), C.POINTER(C.c_double)]),
        "rs_server_wait": (C.c_int, [P, C.c_int64, C.c_double, C.POINTER(C.c_int32)]),
        "rs_server_stats": (C.c_int, [P, C.POINTER(C.c_double)]),
        "rs_server_destroy": (C.c_int, [P]),
        "rs_get_pass_stats": (C.c_int, [P, C.POINTER(rs_pass_stat), C.c_int]),
        "rs_copy_bandwidth": (C.c_int, [P, C.c_int64, C.c_int, C.POINTER(C.c_double)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if lib.rs_version() != VERSION:
        raise ImportError(f"{LIB_PATH} is version {lib.rs_version()}, this binding's structs are version "
                          f"{VERSION}; rebuild with `make -C {HERE.parent}`")
    return lib


lib = _load()


def check(code: int) -> None:
    if code != RS_OK:
        raise RsError(code, lib.rs_last_error().decode(errors="replace"))
