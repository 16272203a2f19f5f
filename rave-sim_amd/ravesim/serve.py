"""Serving form: a scheduler thread in librsim.so admits requests by free
device memory, keeps their fields resident and batches the passes of all
running requests into shared launches (continuous batching).  Each request's
detector rows do not depend on what it was batched with."""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import Optional, Sequence

import numpy as np

from . import _lib
from ._lib import check, lib
from .core import Context, Grating, energy_to_wavelength


@dataclass
class Request:
    """A setup at one or more energies (incoherent weighted sum).

    gratings[e][k] is grating k at energy e (delta/beta differ per energy);
    a grating with several x_positions is phase stepped."""

    N: int
    dx: float
    energies_ev: Sequence[float]
    gratings: Sequence[Sequence[Grating]]
    cutoff_angles: Sequence[float]
    det_size: float
    det_px: float
    z_detector: float
    det_py: float = 1.0
    weights: Optional[Sequence[float]] = None
    dtype: object = np.complex64
    plane_wave: bool = True
    source_x: float = 0.0
    source_z: float = 0.0
    chunk_size: Optional[int] = None


@dataclass
class Result:
    rows: np.ndarray
    wait_s: float
    service_s: float
    latency_s: float


class Server:
    def __init__(self, ctx: Context, reserve_bytes: float = 0.0, max_jobs: int = 0):
        self.ctx = ctx
        h = C.c_void_p()
        check(lib.rs_server_create(ctx.handle, reserve_bytes, max_jobs, C.byref(h)))
        self._h = h

    def submit(self, r: Request) -> int:
        ne = len(r.energies_ev)
        ng = len(r.gratings[0])
        keep = []
        garr = (_lib.rs_grating * max(ne * ng, 1))()
        for e in range(ne):
            if len(r.gratings[e]) != ng:
                raise ValueError("same number of gratings at every energy")
            for k, g in enumerate(r.gratings[e]):
                cg, xs = g.to_c()
                garr[e * ng + k] = cg
                keep.append(xs)
        en = (C.c_double * ne)(*[float(v) for v in r.energies_ev])
        w = (C.c_double * ne)(*[float(v) for v in (r.weights or [1.0] * ne)])
        ang = (C.c_double * (ng + 1))(*[float(a) for a in r.cutoff_angles])
        q = _lib.rs_request(
            grid=_lib.rs_grid(N=int(r.N), dx=float(r.dx), chunk_size=int(r.chunk_size or r.N)),
            det=_lib.rs_detector(size=r.det_size, pixel_x=r.det_px, pixel_y=r.det_py),
            z_detector=r.z_detector,
            dtype=_lib.RS_C8 if np.dtype(r.dtype) == np.complex64 else _lib.RS_C16,
            nr_gratings=ng, nr_energies=ne, source_kind=1 if r.plane_wave else 0,
            energies_ev=C.cast(en, C.POINTER(C.c_double)), weights=C.cast(w, C.POINTER(C.c_double)),
            gratings=C.cast(garr, C.POINTER(_lib.rs_grating)),
            cutoff_angles=C.cast(ang, C.POINTER(C.c_double)),
            source_x=r.source_x, source_z=r.source_z)
        rid = C.c_int64()
        check(lib.rs_server_submit(self._h, C.byref(q), C.byref(rid)))
        return rid.value

    def state(self, rid: int) -> int:
        st = C.c_int32()
        shape = (C.c_int64 * 2)()
        check(lib.rs_server_poll(self._h, rid, C.byref(st), None, 0, shape, None))
        return st.value

    def result(self, rid: int, timeout_s: float = 3600.0) -> Result:
        done = C.c_int32()
        check(lib.rs_server_wait(self._h, rid, timeout_s, C.byref(done)))
        if not done.value:
            raise TimeoutError(f"request {rid} not done after {timeout_s} s")
        st = C.c_int32()
        shape = (C.c_int64 * 2)()
        check(lib.rs_server_poll(self._h, rid, C.byref(st), None, 0, shape, None))
        out = np.zeros((shape[0], shape[1]), dtype=np.float64)
        t = (C.c_double * 3)()
        check(lib.rs_server_poll(self._h, rid, C.byref(st), out.ctypes.data_as(C.POINTER(C.c_double)),
                                 out.size, shape, t))
        return Result(rows=out, wait_s=t[0], service_s=t[1], latency_s=t[2])

    def stats(self) -> dict:
        o = (C.c_double * 6)()
        check(lib.rs_server_stats(self._h, o))
        return {"ticks": o[0], "launches": o[1], "jobs_per_launch": o[2], "jobs_done": o[3],
                "max_resident_jobs": o[4], "busy_s": o[5]}

    def close(self) -> None:
        if getattr(self, "_h", None) is not None:
            check(lib.rs_server_destroy(self._h))
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def phase_sweep(N: int, energies_ev: Sequence[float], n_steps: int = 8, *, dx: float = 2e-9,
                slices: int = 64, dtype=np.complex64, det_size: float = 1e-3,
                weights: Optional[Sequence[float]] = None, pitch: float = 4.2e-6) -> Request:
    """Talbot-Lau phase-stepping request: plane wave at G0 (Au), G1 (Si) at
    0.2 m, G2 (Au) at 0.4 m stepped over n_steps positions k p / n_steps,
    detector at the end of G2.  delta ~ E^-2 and beta ~ E^-4 from the 46 keV
    values of the configuration."""
    h = 60e-6
    grs = []
    for e in energies_ev:
        s = 46000.0 / float(e)
        au = complex(1.1e-6 * s ** 2, 4.0e-8 * s ** 4)
        si = complex(2.6e-7 * s ** 2, 1.0e-10 * s ** 4)
        xs = [k * pitch / n_steps for k in range(n_steps)]
        grs.append([Grating(pitch=pitch, dc=(0.5, 0.5), z_start=0.0, thickness=h, nr_steps=slices, db_a=au),
                    Grating(pitch=pitch, dc=(0.5, 0.5), z_start=0.2, thickness=h, nr_steps=slices, db_a=si),
                    Grating(pitch=pitch, dc=(0.5, 0.5), z_start=0.4, thickness=h, nr_steps=slices, db_a=au,
                            x_positions=xs)])
    return Request(N=N, dx=dx, energies_ev=list(energies_ev), gratings=grs, cutoff_angles=[np.pi / 2] * 4,
                   det_size=min(det_size, N * dx / 2), det_px=1e-6, z_detector=0.4 + h, weights=weights,
                   dtype=dtype)
