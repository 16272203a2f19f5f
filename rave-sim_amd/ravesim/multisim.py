"""Drop-in for big-wave's per-source entry point.

`run_single_simulation` has the signature of big-wave/multisim.py:50-56 and
the same effect: it reads sim_dir/config.yaml, computed.yaml and
%08d/subconfig.yaml and writes %08d/detected.npy (float64, (n_steps, n_px)),
plus u_%04d.npy when save_final_u_vectors is set.  The work runs on the GPU
through librsim.so (rs_run_simdir); the field stays in device memory, so
`scratch_dir` is accepted for signature compatibility and left untouched.
"""
from __future__ import annotations

import ctypes as C
import logging
import time
from pathlib import Path
from typing import Optional

import numpy as np

from . import _lib
from ._lib import check, lib
from .core import Context

logger = logging.getLogger("big-wave")

_default_ctx: Optional[Context] = None

# the tuning table shipped with the package: digit-group splits measured on an
# MI355X; rs_tune_load takes only the entries of the context's own device
DEFAULT_TUNE_TABLE = Path(__file__).resolve().parent / "data" / "tuned_mi355x.txt"


def _ctx(device: int = 0) -> Context:
    """The context of calls that pass none, with the shipped tuning table."""
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(device)
        if DEFAULT_TUNE_TABLE.exists():
            _default_ctx.tune_load(DEFAULT_TUNE_TABLE)
    return _default_ctx


def run_single_simulation(sim_dir, source_idx: int, scratch_dir=None, save_keypoints_path=None,
                          history_dz: Optional[float] = None, *, dtype: Optional[str] = None,
                          plain: bool = False, ctx: Optional[Context] = None) -> np.ndarray:
    """Run source `source_idx` of `sim_dir`; returns the detector rows it saved.

    dtype: None (take config.yaml's, default c16) or 'c8' / 'c16' to force it.
    plain: one kernel per operation instead of fused passes (baseline form).
    history_dz: record big-wave's history into history.npy / history_z.npy /
    history_x.npy of the source's directory (big-wave/multisim.py:130-141).
    save_keypoints_path: write keypoint_%02d_%d.npy there (big-wave/wavesim.py:35-58).
    """
    if history_dz is not None and not float(history_dz) > 0:
        raise ValueError("history_dz must be positive")
    start = time.perf_counter()
    sim_dir = Path(sim_dir)
    info = validate_simdir(sim_dir, source_idx)
    n_steps, n_px = int(info[2]), int(info[3])
    out = np.zeros((n_steps, n_px), dtype=np.float64)
    shape = (C.c_int64 * 2)()
    opts = _lib.rs_run_options(
        dtype={None: 0, "c8": _lib.RS_C8, "c16": _lib.RS_C16}[dtype],
        mode=_lib.RS_MODE_PLAIN if plain else _lib.RS_MODE_FUSED, write_output=1, reserved=0,
        detected=out.ctypes.data_as(C.POINTER(C.c_double)), detected_capacity=out.size,
        shape_out=C.cast(shape, C.POINTER(C.c_int64)),
        history_dz=float(history_dz) if history_dz is not None else 0.0,
        keypoints_dir=str(save_keypoints_path).encode() if save_keypoints_path is not None else None)
    check(lib.rs_run_simdir((ctx or _ctx()).handle, str(sim_dir).encode(), int(source_idx), C.byref(opts)))
    logger.info(f"Finished running simulation for source {source_idx} in {time.perf_counter() - start} seconds")
    return out


def validate_simdir(sim_dir, source_idx: int) -> np.ndarray:
    """Host-only parse + validation: [N, dtype bits, n_steps, n_px, n_elements, n_propagates]."""
    info = np.zeros(6, dtype=np.int64)
    check(lib.rs_validate_simdir(str(sim_dir).encode(), int(source_idx),
                                 info.ctypes.data_as(C.POINTER(C.c_int64)), 6))
    return info
