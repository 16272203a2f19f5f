"""complex128 restatement of big-wave's in-RAM split-step path (test oracle).

TEST INFRASTRUCTURE ONLY -- see oracle/__init__.py.  Every function cites the
big-wave file:line it follows.  Fields are plain numpy arrays; the chunked
iteration of big-wave's `NumpyVector.modify_chunked` (big-wave/vector.py:265-271)
is kept wherever the chunk boundary changes the arithmetic (grating x
coordinates built with one `np.linspace` per chunk, the detector histogram that
is accumulated per chunk).  FFTs are scipy.fft with its default single worker,
as big-wave calls them (big-wave/vector.py:250-254).

Pins (what checks each function independently of this file):
  fftfreq_chunk          np.fft.fftfreq; big-wave/test.py:90-97 (tests/test_oracle_kat.py)
  apply_frequency_cutoff big-wave/test.py:187-199 reference lambda |fftfreq/dx| < f
  grating_mask           big-wave/test.py:121-185 templates + stepping invariance;
                         fast-wave/test/fwtest.cu:71-87 mask
  convert_energy_wavelength big-wave/test.py:201-207
  square_and_downsample  big-wave/test.py:209-250 KAT [1.25, 12.5, 15.249999999999998]
  propagate, point source, run_simulation, setup rules, delta/beta:
                         end-to-end golden vector big-wave/test.py:853-960
                         (25 pixels, atol 1e-14), plus closed-form checks
                         (free-space propagation of a Gaussian, unitarity,
                         composition of two hops) in tests/test_oracle_math.py
  sample_apply           closed forms in tests/test_oracle_math.py: a one-pixel
                         row is the hat function of linear interpolation, a
                         uniform grid is the uniform grating slab, and moving
                         x_position by s samples rolls the output by +s
  dba_grating_mask/apply reductions to the plain grating (dc1 >= 1) and to a
                         uniform slab (dc1 = 0), the second pitch's condition
                         against exact rational arithmetic per slice, and the
                         AND of both conditions (tests/test_oracle_dba.py)
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from pathlib import Path
from typing import Optional, Sequence

import numpy as np
from scipy import fft as spfft

# Physical constants exactly as big-wave/propagation.py:26-28 writes them.
PLANCK_H = 6.62607004 * 10 ** (-34)
C_0 = 299792458
EV_TO_JOULE = 1.602176634 * 10 ** (-19)
Z_TOLERANCE = 1e-8  # big-wave/wavesim.py:138


# ----------------------------------------------------------------------------
# parameter records
# ----------------------------------------------------------------------------


@dataclass
class SimParams:
    """big-wave/propagation.py:12-23."""

    N: int
    dx: float
    z_detector: float
    detector_size: float
    detector_pixel_size_x: float
    detector_pixel_size_y: float
    wl: float
    chunk_size: int


@dataclass
class Grating:
    """big-wave/optical_element.py:116-130 (deltabetas already resolved)."""

    pitch: float
    dc: tuple
    z_start: float
    thickness: float
    nr_steps: int
    x_positions: np.ndarray
    substrate_thickness: float
    db_a: complex = 0j
    db_b: complex = 0j
    db_substrate: complex = 0j

    def get_thickness(self) -> float:
        return self.thickness + self.substrate_thickness


@dataclass
class DBAGrating:
    """big-wave/optical_element.py:227-243 (deltabetas already resolved)."""

    pitch0: float
    pitch1: float
    dc0: tuple
    dc1: tuple
    z_start: float
    thickness: float
    nr_steps: int
    x_positions: np.ndarray
    substrate_thickness: float
    db_a: complex = 0j
    db_b: complex = 0j
    db_substrate: complex = 0j

    def get_thickness(self) -> float:
        return self.thickness + self.substrate_thickness


@dataclass
class Sample:
    """big-wave/optical_element.py:380-430 (db_list[0] is vacuum)."""

    z_start: float
    pixel_size_x: float
    pixel_size_z: float
    grid: np.ndarray
    db_list: np.ndarray
    x_positions: np.ndarray

    def get_thickness(self) -> float:
        return self.pixel_size_z * self.grid.shape[0]


@dataclass
class PointSource:
    x: float
    z: float


@dataclass
class VectorSource:
    vec: np.ndarray
    z: float


# ----------------------------------------------------------------------------
# primitives (big-wave/propagation.py)
# ----------------------------------------------------------------------------


def convert_energy_wavelength(e_or_wl: float) -> float:
    """lambda = h c / (E e); big-wave/propagation.py:69-71."""
    return PLANCK_H * C_0 / (e_or_wl * EV_TO_JOULE)


def fftfreq_chunk(n: int, start: int, length: int) -> np.ndarray:
    """Slice [start, start+length) of np.fft.fftfreq(n); big-wave/propagation.py:154-175.

    Positive half holds 0..ceil(n/2)-1, negative half -(n//2)..-1; the values
    are the integer bins times 1.0/n, as numpy computes them.
    """
    idx = np.arange(start, start + length, dtype=np.int64)
    npos = (n - 1) // 2 + 1
    signed = np.where(idx < npos, idx, idx - n)
    return signed * (1.0 / n)


def cutoff_removal_range(n: int, freq: float, dx: float) -> Optional[tuple]:
    """Index range zeroed by the frequency cutoff, or None.

    big-wave/propagation.py:83-90: start = ceil(n*freq*dx); nothing is removed
    when start > n/2; otherwise indices [start, n - start + 1) are zeroed.
    """
    frac = freq * dx
    start = math.ceil(n * frac)
    if start > n / 2:
        return None
    return start, n - start + 1


def apply_frequency_cutoff(U: np.ndarray, freq: float, dx: float, chunk_size: int) -> None:
    """Zero the |f| >= f_c band in place; big-wave/propagation.py:74-97."""
    rng = cutoff_removal_range(len(U), freq, dx)
    if rng is None:
        return
    U[rng[0]:rng[1]] = 0.0


class Fields:
    """The pair (u, U) of big-wave's driver; U may hold FFT(u) (skip_fft)."""

    def __init__(self, u: np.ndarray):
        self.u = u
        self.U = np.zeros_like(u)


def _propagator_chunk(n: int, start: int, length: int, dx: float, wl: float, dz: float) -> np.ndarray:
    # big-wave/propagation.py:124-128
    fx = fftfreq_chunk(n, start, length) / dx
    return np.exp(-1j * 2 * np.pi / wl * dz) * np.exp(1j * np.pi * wl * dz * (fx ** 2))


def propagate(f: Fields, dx: float, wl: float, dz: float, chunk_size: int,
              cutoff_frequency: float, skip_fft: bool = False) -> None:
    """Fresnel free-space step; big-wave/propagation.py:100-138.

    U = FFT(u) (unless skip_fft), U *= exp(-2 pi i dz/wl) exp(i pi wl dz f^2)
    chunk by chunk, cutoff, u = IFFT(U).  U keeps the propagated spectrum.
    """
    n = len(f.u)
    if not skip_fft:
        f.U = spfft.fft(f.u)
    U = f.U
    for i in range(0, n, chunk_size):
        seg = U[i:i + chunk_size]
        seg *= _propagator_chunk(n, i, len(seg), dx, wl, dz)
    apply_frequency_cutoff(U, cutoff_frequency, dx, chunk_size)
    f.u = spfft.ifft(U)


def propagate_analytically(N: int, dx: float, wl: float, chunk_size: int, z: float,
                           x_source: float, dtype=np.complex128) -> np.ndarray:
    """Point-source field exp(-2 pi i r/wl)/sqrt(r); big-wave/propagation.py:141-151."""
    u = np.zeros(N, dtype=dtype)
    for idx in range(0, N, chunk_size):
        length = min(chunk_size, N - idx)
        x = (idx + np.arange(length, dtype=np.float64) - N / 2) * dx - x_source
        r = np.sqrt(x ** 2 + z ** 2)
        u[idx:idx + length] = np.exp(r * (-2j * np.pi / wl)) / np.sqrt(r)
    return u


def grid_density_check(dz: float, x_source: float, N: int, dx: float, wl: float) -> None:
    """Nyquist check of the analytic source; big-wave/propagation.py:31-45."""

    def dist_x(idx: int) -> float:
        return np.sqrt(((N / 2 - idx) * dx + np.abs(x_source)) ** 2 + dz ** 2)

    if not np.abs(dist_x(0) - dist_x(1)) <= wl / 2:
        raise ValueError("Nyquist not satisfied, dx too large")


def detector_x_vector(detector_size: float, px: float) -> np.ndarray:
    """Pixel centres (arange(n) - n//2)*px with n = size // px as a float;
    big-wave/util.py:78-82."""
    n = detector_size // px
    return (np.arange(n) - n // 2) * px


def square_and_downsample(u: np.ndarray, p: SimParams, current_z: float) -> np.ndarray:
    """Detector intensity; big-wave/propagation.py:279-320.

    |u|^2 is binned by np.histogram over n_px+1 equal bins of
    [-(n_px/2) px, (n_px/2 + 1) px) per chunk (the extra last bin is dropped),
    then scaled by cos(atan(x/z)) dx py / sqrt(x^2 + z^2).
    """
    outsize = int(p.detector_size // p.detector_pixel_size_x)
    assert outsize <= p.N
    out = np.zeros(outsize, dtype=np.float64)
    rng = (-(outsize / 2) * p.detector_pixel_size_x,
           (outsize / 2 + 1) * p.detector_pixel_size_x)
    for idx in range(0, len(u), p.chunk_size):
        chunk = u[idx:idx + p.chunk_size]
        pos = (idx + np.arange(len(chunk)) - p.N / 2) * p.dx
        w = chunk.real * chunk.real + chunk.imag * chunk.imag
        out[:] += np.histogram(pos, bins=outsize + 1, range=rng, weights=w)[0][:-1]
    x = detector_x_vector(p.detector_size, p.detector_pixel_size_x)
    ang = np.arctan(x / current_z)
    return out * np.cos(ang) * p.dx * p.detector_pixel_size_y / np.sqrt(x ** 2 + current_z ** 2)


# ----------------------------------------------------------------------------
# optical elements (big-wave/optical_element.py)
# ----------------------------------------------------------------------------


def material_factor(deltabeta, thickness: float, wl: float):
    """exp(2 pi i t/wl (delta + i beta)); big-wave/optical_element.py:36-39."""
    return np.exp(2j * np.pi * thickness / wl * deltabeta)


def lookup_deltabeta(name: str, density: float, table: Sequence) -> complex:
    """Closest-density match with tolerance 1e-5; big-wave/optical_element.py:42-63."""
    cands = sorted((abs(d - density), db) for (m, d), db in table if m == name)
    if not cands:
        raise KeyError(f"Material {name} not found in deltabeta table")
    if not cands[0][0] < 1e-5:
        raise KeyError(f"Material {name} found with density error of {cands[0][0]}")
    return cands[0][1]


def grating_mask(dx: float, start: int, length: int, g: Grating, z_fraction: float,
                 step: int) -> np.ndarray:
    """Material-A mask of one chunk; big-wave/optical_element.py:341-357.

    dc is interpolated in FP64, x is np.linspace(x0, x0 + length*dx, length,
    endpoint=False) with x0 = start*dx + x_positions[step], and the mask is
    mod(x/pitch + dc/2, 1) < dc.
    """
    dc = (1 - z_fraction) * g.dc[0] + z_fraction * g.dc[1]
    x_start = start * dx + g.x_positions[step]
    x_stop = x_start + length * dx
    x = np.linspace(x_start, x_stop, length, endpoint=False)
    return np.mod(x / g.pitch + dc / 2, 1) < dc


def grating_apply(f: Fields, g: Grating, p: SimParams, cutoff_freq: float, step: int,
                  history=None) -> None:
    """Slice loop of a grating plus its substrate; big-wave/optical_element.py:134-206.
    With a history, one entry after every slice at z_start + (i + 1) z_step."""
    z_step = g.thickness / g.nr_steps
    fa = material_factor(g.db_a, z_step, p.wl)
    fb = material_factor(g.db_b, z_step, p.wl)
    for i in range(g.nr_steps):
        zf = i / g.nr_steps
        for idx in range(0, p.N, p.chunk_size):
            length = min(p.chunk_size, p.N - idx)
            mask = grating_mask(p.dx, idx - p.N // 2, length, g, zf, step)
            fac = np.where(mask, fa, fb).astype(f.u.dtype)
            f.u[idx:idx + length] *= fac
        propagate(f, p.dx, p.wl, z_step, p.chunk_size, cutoff_freq)
        if history is not None:
            z = g.z_start + (i + 1) * z_step
            history.push(square_and_downsample(f.u, p, z), z)
    if g.substrate_thickness > 0:
        f.u *= np.complex128(material_factor(g.db_substrate, g.substrate_thickness, p.wl))
        propagate(f, p.dx, p.wl, g.substrate_thickness, p.chunk_size, cutoff_freq)


def dba_grating_mask(dx: float, start: int, length: int, g: DBAGrating, z_fraction: float,
                     step: int) -> np.ndarray:
    """Material-A mask of one chunk of a DBA grating; big-wave/optical_element.py:359-375.

    Both duty cycles are interpolated in FP64, x is the chunk's np.linspace as
    for the plain grating, and the mask is
    (mod(x/pitch0 + dc0/2, 1) < dc0) * (mod(x/pitch1, 1) < dc1).
    """
    dc0 = (1 - z_fraction) * g.dc0[0] + z_fraction * g.dc0[1]
    dc1 = (1 - z_fraction) * g.dc1[0] + z_fraction * g.dc1[1]
    x_start = start * dx + g.x_positions[step]
    x_stop = x_start + length * dx
    x = np.linspace(x_start, x_stop, length, endpoint=False)
    phase0 = np.mod(x / g.pitch0 + dc0 / 2, 1)
    phase1 = np.mod(x / g.pitch1, 1)
    return (phase0 < dc0) * (phase1 < dc1)


def dba_grating_apply(f: Fields, g: DBAGrating, p: SimParams, cutoff_freq: float,
                      step: int, history=None) -> None:
    """Slice loop of a DBA grating plus its substrate; big-wave/optical_element.py:247-316.
    Factor A where the mask holds, B elsewhere (np.place with mask and 1 - mask)."""
    z_step = g.thickness / g.nr_steps
    fa = material_factor(g.db_a, z_step, p.wl)
    fb = material_factor(g.db_b, z_step, p.wl)
    for i in range(g.nr_steps):
        zf = i / g.nr_steps
        for idx in range(0, p.N, p.chunk_size):
            length = min(p.chunk_size, p.N - idx)
            mask = dba_grating_mask(p.dx, idx - p.N // 2, length, g, zf, step)
            fac = np.where(mask, fa, fb).astype(f.u.dtype)
            f.u[idx:idx + length] *= fac
        propagate(f, p.dx, p.wl, z_step, p.chunk_size, cutoff_freq)
        if history is not None:
            z = g.z_start + (i + 1) * z_step
            history.push(square_and_downsample(f.u, p, z), z)
    if g.substrate_thickness > 0:
        f.u *= np.complex128(material_factor(g.db_substrate, g.substrate_thickness, p.wl))
        propagate(f, p.dx, p.wl, g.substrate_thickness, p.chunk_size, cutoff_freq)


def sample_apply(f: Fields, s: Sample, p: SimParams, cutoff_freq: float, step: int,
                 history=None) -> None:
    """Row-by-row sample; big-wave/optical_element.py:432-476.

    Each z-row's delta+i*beta is linearly interpolated (np.interp over the
    complex values, zero outside; numpy's complex interp forms the slope as
    difference * (1 / dx)) from pixel positions arange(len)*px + x_pos - len*px/2 onto the field x
    coordinates, then exp(2 pi i dz/wl db) is applied and the row propagated.
    """
    rowlen = s.grid.shape[1]
    row_x = np.arange(rowlen) * s.pixel_size_x + s.x_positions[step] - rowlen * s.pixel_size_x * 0.5
    for r in range(s.grid.shape[0]):
        if history is not None:
            z = s.z_start + r * s.pixel_size_z
            history.push(square_and_downsample(f.u, p, z), z)
        row_db = s.db_list[s.grid[r, :]]
        for idx in range(0, p.N, p.chunk_size):
            length = min(p.chunk_size, p.N - idx)
            start = idx - p.N / 2
            xc = np.arange(start, start + length) * p.dx
            db = np.interp(xc, row_x, row_db, left=0.0, right=0.0)
            f.u[idx:idx + length] *= material_factor(db, s.pixel_size_z, p.wl)
        propagate(f, p.dx, p.wl, s.pixel_size_z, p.chunk_size, cutoff_freq)


def element_apply(f: Fields, el, p: SimParams, cutoff_freq: float, step: int,
                  history=None) -> None:
    if isinstance(el, Grating):
        grating_apply(f, el, p, cutoff_freq, step, history)
    elif isinstance(el, DBAGrating):
        dba_grating_apply(f, el, p, cutoff_freq, step, history)
    elif isinstance(el, Sample):
        sample_apply(f, el, p, cutoff_freq, step, history)
    else:
        raise TypeError(type(el))


# ----------------------------------------------------------------------------
# sources (big-wave/source.py)
# ----------------------------------------------------------------------------


# ----------------------------------------------------------------------------
# history (big-wave/history.py)
# ----------------------------------------------------------------------------


class History:
    """Intensity rows with their z values; big-wave/history.py:10-40."""

    def __init__(self, dz: float):
        self.dz = dz
        self.entries: list = []
        self.zs: list = []

    def push(self, downscaled: np.ndarray, z: float) -> None:
        self.entries.append(downscaled)
        self.zs.append(z)

    def get_history(self) -> np.ndarray:
        """(n_px, entries), np.empty((0, 0)) when empty."""
        if not self.entries:
            return np.empty((0, 0))
        return np.stack(self.entries, axis=-1)


def generate_analytic_history(history: History, zs: np.ndarray, x_source: float, p: SimParams,
                              cutoff_freq: float) -> None:
    """Point-source intensity rows in closed form; big-wave/history.py:43-66."""
    gradient = cutoff_freq * p.wl
    n = p.detector_size // p.detector_pixel_size_x
    x = (np.arange(n) - n // 2) * p.detector_pixel_size_x - x_source
    for z in zs:
        angle = np.arctan(x / z)
        r2 = x ** 2 + z ** 2
        intensity = p.detector_pixel_size_x * p.detector_pixel_size_y * np.cos(angle) / r2
        intensity[np.abs(x) > gradient * z] = 0
        history.push(intensity, z)


def propagate_with_history(f: Fields, p: SimParams, dz: float, cutoff_freq: float,
                           current_z: float, skip_fft: bool = False,
                           history: Optional[History] = None) -> None:
    """big-wave/history.py:69-135: entries at zs = arange(0, dz, history.dz)
    (the last moved to dz when closer than half a step, else dz appended), each
    propagated from the segment's start spectrum; u and U end as after one
    propagation over dz."""
    if history is None:
        propagate(f, p.dx, p.wl, dz, p.chunk_size, cutoff_freq, skip_fft)
        return
    if dz == 0:
        return
    zs = np.arange(0, dz, history.dz)
    if dz - zs[-1] < history.dz / 2.0 and len(zs) > 1:
        zs[-1] = dz
    else:
        zs = np.append(zs, [dz])
    if not skip_fft:
        f.U = spfft.fft(f.u)
    U_start = f.U.copy()
    for i in range(1, len(zs)):
        propagate(f, p.dx, p.wl, zs[i] - zs[0], p.chunk_size, cutoff_freq, True)
        z = zs[i] + current_z
        history.push(square_and_downsample(f.u, p, z), z)
        if i < len(zs) - 1:
            f.U = U_start.copy()


def source_propagate_to(src, z_out: float, p: SimParams, cutoff_freq: float, dtype,
                        history: Optional[History] = None) -> Fields:
    """big-wave/source.py:46-102.  Point source: analytic field then FFT, cutoff,
    IFFT (U does not hold FFT(u) afterwards), analytic history rows at
    arange(0, z_out, dz)[1:].  Vector source: copy when dz == 0, otherwise FFT
    and propagate with skip_fft."""
    if isinstance(src, PointSource):
        dz = z_out - src.z
        grid_density_check(dz, src.x, p.N, p.dx, p.wl)
        u = propagate_analytically(p.N, p.dx, p.wl, p.chunk_size, dz, src.x, dtype)
        f = Fields(u)
        f.U = spfft.fft(f.u)
        apply_frequency_cutoff(f.U, cutoff_freq, p.dx, p.chunk_size)
        f.u = spfft.ifft(f.U)
        if history is not None:
            zs = np.arange(0, z_out, history.dz)[1:]
            generate_analytic_history(history, zs, src.x, p, cutoff_freq)
        return f
    dz = z_out - src.z
    f = Fields(src.vec.astype(dtype, copy=True))
    if dz == 0:
        return f
    f.U = spfft.fft(src.vec.astype(dtype))
    propagate_with_history(f, p, dz, cutoff_freq, src.z, True, history)
    return f


# ----------------------------------------------------------------------------
# per-source driver (big-wave/wavesim.py)
# ----------------------------------------------------------------------------


def run_simulation(p: SimParams, source, elements: list, cutoff_angles: Sequence[float],
                   dtype=np.complex128, final_u: Optional[list] = None,
                   history: Optional[History] = None, keypoints: Optional[dict] = None) -> list:
    """One source through all elements incl. phase steps; big-wave/wavesim.py:84-298.

    Returns the list of detector rows (one per phase step).  Snapshot before
    the first element with >1 x_positions; the final hop reuses U (skip_fft)
    on the first pass only; elements with one x_position use index 0 in later
    phase steps.  If `final_u` is a list, the final u of every phase step is
    appended to it.  A history records the first pass only; `keypoints`
    receives u before (element, 0) and after (element, 1) every element of
    the first pass (wavesim.py:35-58, 182-197).
    """
    fcs = [np.sin(a) / p.wl for a in cutoff_angles]
    fc = fcs[0]
    current_z = elements[0].z_start if elements else p.z_detector
    f = source_propagate_to(source, current_z, p, fc, dtype, history)

    snapshot = None
    for i, el in enumerate(elements):
        if current_z < el.z_start:
            propagate_with_history(f, p, el.z_start - current_z, fc, current_z, False, history)
        current_z = el.z_start
        if len(el.x_positions) > 1 and snapshot is None:
            snapshot = (current_z, i, f.u.copy(), len(el.x_positions))
        fc = fcs[i + 1]
        if keypoints is not None:
            keypoints[(i, 0)] = f.u.copy()
        element_apply(f, el, p, fc, 0, history)
        if keypoints is not None:
            keypoints[(i, 1)] = f.u.copy()
        current_z = el.z_start + el.get_thickness()

    if current_z < p.z_detector - Z_TOLERANCE:
        propagate_with_history(f, p, p.z_detector - current_z, fc, current_z,
                               len(elements) > 0, history)

    rows = [square_and_downsample(f.u, p, p.z_detector)]
    if final_u is not None:
        final_u.append(f.u.copy())

    if snapshot is not None:
        z0, i0, u0, nsteps = snapshot
        for step in range(1, nsteps):
            current_z = z0
            f = Fields(u0.copy())
            fc = fcs[i0]
            for i in range(i0, len(elements)):
                el = elements[i]
                if current_z < el.z_start:
                    propagate(f, p.dx, p.wl, el.z_start - current_z, p.chunk_size, fc)
                current_z = el.z_start
                fc = fcs[i + 1]
                element_apply(f, el, p, fc, step if len(el.x_positions) > 1 else 0)
                current_z = el.z_start + el.get_thickness()
            if current_z < p.z_detector - Z_TOLERANCE:
                propagate(f, p.dx, p.wl, p.z_detector - current_z, p.chunk_size, fc)
            rows.append(square_and_downsample(f.u, p, p.z_detector))
            if final_u is not None:
                final_u.append(f.u.copy())
    return rows


# ----------------------------------------------------------------------------
# sim-dir setup rules (big-wave/multisim.py, propagation.py:178-276)
# ----------------------------------------------------------------------------


def compute_cutoff_angles(detector_size: float, dx: float, energy_range, z_source: float,
                          z_distances: list, element_heights: list, z_detector: float,
                          max_x: float):
    """Cutoff angles and max |x| per element start; big-wave/propagation.py:178-276."""
    assert len(z_distances) == len(element_heights)
    wl_min = convert_energy_wavelength(energy_range[0])
    wl_max = convert_energy_wavelength(energy_range[1])
    direct = np.arctan((detector_size / 2 + max_x) / z_detector)
    if 2 / (np.sin(direct) / wl_max) < dx:
        raise ValueError("grid sampled too coarsely for the highest energy")
    if wl_min * 2 / dx < 1:
        grad_emin = np.tan(np.arcsin(wl_min * 2 / dx))
    else:
        grad_emin = float(np.inf)
    cur_x, cur_z = max_x, z_source
    grads, xs = [], []
    for nz in list(z_distances) + [z_detector]:
        if cur_z == z_detector:
            raise ValueError("element at the detector z")
        grads.append(min((detector_size / 2 + cur_x) / (z_detector - cur_z), grad_emin))
        cur_x += grads[-1] * (nz - cur_z)
        xs.append(float(cur_x))
        cur_z = nz
    assert cur_x >= detector_size / 2
    if len(z_distances) > 0 and np.abs(z_detector - (z_distances[-1] + element_heights[-1])) < 1e-8:
        grads[-1] = grads[-2]
        xs[-1] = xs[-2] + grads[-1] * element_heights[-1]
    return [float(np.arctan(g)) for g in grads], xs


def draw_point_sources(seed: int, n: int, energy_range, x_range):
    """Energies first, then x positions, from one RandomState; big-wave/multisim.py:358-379."""
    rs = np.random.RandomState(seed)
    energies = rs.rand(n) * (energy_range[1] - energy_range[0]) + energy_range[0]
    xs = rs.rand(n) * (x_range[1] - x_range[0]) + x_range[0]
    return energies, xs


def params_from_config(sp: dict, wl: float) -> SimParams:
    """big-wave/config.py:45-55 with the wavelength filled in (multisim.py:112)."""
    return SimParams(N=int(sp["N"]), dx=float(sp["dx"]), z_detector=float(sp["z_detector"]),
                     detector_size=float(sp["detector_size"]),
                     detector_pixel_size_x=float(sp["detector_pixel_size_x"]),
                     detector_pixel_size_y=float(sp["detector_pixel_size_y"]),
                     wl=wl, chunk_size=int(sp["chunk_size"]))


def elements_from_config(els: list, table: Sequence, base_dir: Optional[Path] = None) -> list:
    """Grating / sample records from config dicts; big-wave/config.py:76-136 plus
    the deltabeta resolution of optical_element.py:212-225, 423-430."""

    def db(mat):
        if mat is None:
            return 0j
        return complex(lookup_deltabeta(mat[0], float(mat[1]), table))

    out = []
    for d in els:
        if d["type"] == "grating":
            out.append(Grating(pitch=float(d["pitch"]), dc=(float(d["dc"][0]), float(d["dc"][1])),
                               z_start=float(d["z_start"]), thickness=float(d["thickness"]),
                               nr_steps=int(d["nr_steps"]),
                               x_positions=np.array(d["x_positions"], dtype=np.float64),
                               substrate_thickness=float(d["substrate_thickness"]),
                               db_a=db(d.get("mat_a")), db_b=db(d.get("mat_b")),
                               db_substrate=db(d.get("mat_substrate"))))
        elif d["type"] == "dba_grating":
            # keys of fast-wave/src/config_parsing.cpp:107-137 (big-wave's
            # config.py:103-116 names the same keys)
            out.append(DBAGrating(pitch0=float(d["pitch0"]), pitch1=float(d["pitch1"]),
                                  dc0=(float(d["dc0"][0]), float(d["dc0"][1])),
                                  dc1=(float(d["dc1"][0]), float(d["dc1"][1])),
                                  z_start=float(d["z_start"]), thickness=float(d["thickness"]),
                                  nr_steps=int(d["nr_steps"]),
                                  x_positions=np.array(d["x_positions"], dtype=np.float64),
                                  substrate_thickness=float(d["substrate_thickness"]),
                                  db_a=db(d.get("mat_a")), db_b=db(d.get("mat_b")),
                                  db_substrate=db(d.get("mat_substrate"))))
        elif d["type"] == "sample":
            grid = np.load(Path(base_dir) / d["grid_path"]) if base_dir is not None else d["grid"]
            assert grid.dtype == np.int8
            dbl = np.zeros(len(d["materials"]) + 1, dtype=np.complex128)
            for k, m in enumerate(d["materials"]):
                dbl[k + 1] = db(m)
            out.append(Sample(z_start=float(d["z_start"]), pixel_size_x=float(d["pixel_size_x"]),
                              pixel_size_z=float(d["pixel_size_z"]), grid=grid, db_list=dbl,
                              x_positions=np.array(d["x_positions"], dtype=np.float64)))
        else:
            raise ValueError(f"unsupported element type {d['type']}")
    return out


def parse_dtype(cfg: dict):
    """c16 unless `dtype: c8`; big-wave/config.py:171-188."""
    val = cfg.get("dtype", "c16")
    return {"c8": np.complex64, "c16": np.complex128}[val]


def run_simdir(sim_dir, source_idx: int, dtype=None, history_dz: Optional[float] = None,
               keypoints: Optional[dict] = None, final_u: Optional[list] = None):
    """Oracle for one source of a sim dir; big-wave/multisim.py:50-150.

    Reads config.yaml / computed.yaml / %08d/subconfig.yaml (safe YAML loader)
    and returns the float64 (n_steps, n_px) detector array that big-wave saves
    as detected.npy.  `dtype` overrides the config dtype (the oracle is normally
    run in complex128).  With `history_dz` it returns (detected, history) where
    history is the History object (multisim.py:130-141 saves
    history.get_history() and its z values).
    """
    import yaml

    sim_dir = Path(sim_dir)
    cfg = yaml.safe_load((sim_dir / "config.yaml").read_text())
    comp = yaml.safe_load((sim_dir / "computed.yaml").read_text())
    sub = yaml.safe_load((sim_dir / f"{source_idx:08d}" / "subconfig.yaml").read_text())
    table = [((e[0][0], float(e[0][1])), complex(e[1][0], e[1][1])) for e in sub["deltabeta_table"]]
    energy = float(sub["energy"])
    p = params_from_config(cfg["sim_params"], convert_energy_wavelength(energy))
    els = elements_from_config(cfg["elements"], table, sim_dir)
    s = sub["source"]
    if dtype is None:
        dtype = parse_dtype(cfg)
    if s["type"] == "point":
        src = PointSource(x=float(s["x"]), z=float(s["z"]))
    else:
        src = VectorSource(vec=np.load(s["input_path"]), z=float(s["z"]))
    hist = History(float(history_dz)) if history_dz is not None else None
    rows = run_simulation(p, src, els, [float(a) for a in comp["cutoff_angles"]], dtype,
                          final_u=final_u, history=hist, keypoints=keypoints)
    det = np.asarray(rows, dtype=np.float64)
    return det if hist is None else (det, hist)
