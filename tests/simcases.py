"""Seeded synthetic inputs for the benchmark configurations, as sim dirs.

Each builder returns (config dict, list of sub-configs, cutoff angles) in
big-wave's sim-dir schema.  Full-size and reduced-size variants share the
physics (energies, gratings, distances); reduced variants shrink N, slices and
the detector so that the complex128 oracle finishes in seconds.

Choices the workload descriptions leave open: chunk_size = N unless given,
dx = 2 nm, plane waves are a vector source of ones at the first grating,
no frequency cutoff (all cutoff angles pi/2) for plane-wave chains, the
detector follows the last element directly when the config says so.
"""
from __future__ import annotations

import math
from pathlib import Path

import numpy as np

# delta/beta of the monochromatic chains (Si at 25 keV, Au and Si at 46 keV)
DB_SI_25KEV = (7.7e-7, 3.5e-10)
DB_AU_46KEV = (1.1e-6, 4.0e-8)
DB_SI_46KEV = (2.6e-7, 1.0e-10)


def grating(pitch, z_start, thickness, nr_steps, mat_a, x_positions=(0.0,), dc=(0.5, 0.5),
            substrate=0.0, mat_b=None, mat_s=None):
    return {"type": "grating", "pitch": pitch, "dc": list(dc), "z_start": z_start,
            "thickness": thickness, "nr_steps": nr_steps, "x_positions": [float(x) for x in x_positions],
            "substrate_thickness": substrate, "mat_a": mat_a, "mat_b": mat_b, "mat_substrate": mat_s}


def sim_params(N, dx, z_det, det_size, px, py=1.0, chunk=None):
    return {"N": int(N), "dx": dx, "z_detector": z_det, "detector_size": det_size,
            "detector_pixel_size_x": px, "detector_pixel_size_y": py, "chunk_size": int(chunk or N)}


def write_ones(path: Path, N: int) -> str:
    np.save(path, np.ones(N, dtype=np.complex128))
    return str(path)


def mono_chain(tmp: Path, N=2 ** 20, slices=32, det_size=1e-3, dtype="c16", chunk=None):
    """Plain monochromatic chain: plane wave at 25 keV, one Si phase grating
    (p 4.2 um, h 30 um, 32 slices), 0.1 m hop, 1 um pixels."""
    dx = 2e-9
    h = 30e-6
    cfg = {"sim_params": sim_params(N, dx, h + 0.1, det_size, 1e-6, 1.0, chunk),
           "use_disk_vector": False, "save_final_u_vectors": False, "dtype": dtype,
           "elements": [grating(4.2e-6, 0.0, h, slices, ["Si", 2.34])]}
    src = write_ones(tmp / "ones.npy", N)
    subs = [{"energy": 25000.0, "source": {"type": "vector", "input_path": src, "z": 0.0},
             "deltabeta_table": [[["Si", 2.34], list(DB_SI_25KEV)]]}]
    return cfg, subs, [math.pi / 2] * 2


def talbot_lau(tmp: Path, N=2 ** 26, slices=64, det_size=None, dtype="c8", chunk=None, g2_steps=1):
    """Talbot-Lau 3-grating chain at 46 keV: G0 Au absorption, G1 Si phase,
    G2 = G0, 60 um high, 0.2 m start to start; plane wave at G0, detector at
    the end of G2.  g2_steps > 1 phase-steps G2 at k p / g2_steps."""
    dx = 2e-9
    h, p = 60e-6, 4.2e-6
    if det_size is None:
        det_size = min(1e-3, N * dx / 2)
    xs = [k * p / g2_steps for k in range(g2_steps)]
    els = [grating(p, 0.0, h, slices, ["Au", 19.32]),
           grating(p, 0.2, h, slices, ["Si", 2.34]),
           grating(p, 0.4, h, slices, ["Au", 19.32], x_positions=xs)]
    cfg = {"sim_params": sim_params(N, dx, 0.4 + h, det_size, 1e-6, 1.0, chunk),
           "use_disk_vector": False, "save_final_u_vectors": False, "dtype": dtype, "elements": els}
    src = write_ones(tmp / "ones.npy", N)
    subs = [{"energy": 46000.0, "source": {"type": "vector", "input_path": src, "z": 0.0},
             "deltabeta_table": [[["Au", 19.32], list(DB_AU_46KEV)], [["Si", 2.34], list(DB_SI_46KEV)]]}]
    return cfg, subs, [math.pi / 2] * 4


def spectrum_weights(energies_kev):
    """Synthetic 70 kVp tungsten spectrum: Kramers (70-E)/E times 3 mm Al with
    mu_Al = 0.5/cm (E/50)^-3, normalised to sum 1."""
    e = np.asarray(energies_kev, dtype=np.float64)
    mu = 0.5 * (e / 50.0) ** -3
    w = (70.0 - e) / e * np.exp(-mu * 0.3)
    w = np.clip(w, 0, None)
    return w / w.sum()


def poly_members(n_energies=64, N=2 ** 24, slices=32):
    """Polychromatic version of the monochromatic chain: energies 20..70 keV uniformly, delta ~ E^-2,
    beta ~ E^-4 scaled from the 25 keV values."""
    e = np.linspace(20.0, 70.0, n_energies)
    w = spectrum_weights(e)
    out = []
    for ek, wk in zip(e, w):
        s = 25.0 / ek
        db = complex(DB_SI_25KEV[0] * s ** 2, DB_SI_25KEV[1] * s ** 4)
        out.append((ek * 1000.0, float(wk), db))
    return out


def sample_case(tmp: Path, N=2 ** 16, dtype="c16", steps=(0.0, 1.3e-7)):
    """Point source at 30 keV, a Si phase grating, then a sample grid (int8,
    6 rows x 40 pixels of 0.5 um: vacuum / Si / Au, 2 um per row) phase
    stepped over `steps`, detector 50 mm behind.  Material boundaries inside
    the grid give interpolated delta/beta (big-wave Sample.apply)."""
    dx = 1e-9
    rng = np.random.default_rng(11)
    grid = rng.integers(0, 3, size=(6, 40)).astype(np.int8)
    grid[:, :3] = 0
    np.save(tmp / "sample_grid.npy", grid)
    z_s = 0.012
    els = [grating(1.0e-6, 0.01, 4e-6, 4, ["Si", 2.34]),
           {"type": "sample", "z_start": z_s, "pixel_size_x": 0.5e-6, "pixel_size_z": 2e-6,
            "grid_path": "sample_grid.npy", "materials": [["Si", 2.34], ["Au", 19.32]],
            "x_positions": [float(s) for s in steps]}]
    z_det = z_s + 6 * 2e-6 + 0.05
    cfg = {"sim_params": sim_params(N, dx, z_det, 40e-6, 1e-6, 1.0),
           "use_disk_vector": False, "save_final_u_vectors": False, "dtype": dtype, "elements": els}
    subs = [{"energy": 30000.0, "source": {"type": "point", "x": 0.0, "z": 0.0},
             "deltabeta_table": [[["Si", 2.34], [5.1e-7, 1.6e-10]], [["Au", 19.32], [3.6e-6, 2.1e-7]]]}]
    return cfg, subs, [math.pi / 2] * 3


def dba_case(N=2 ** 16, dtype="c16", steps=(0.0, 1.05e-6), chunk=None):
    """Point source at 30 keV, a DBA grating (Au, pitches 2.1 and 6.3 um, both
    duty cycles graded over 6 slices, Si substrate) phase stepped over `steps`,
    then a plain Si grating and the detector.  pitch/dx = 2100 and 6300 put
    samples exactly on both conditions' edges."""
    dx = 1e-9
    dba = {"type": "dba_grating", "pitch0": 2.1e-6, "pitch1": 6.3e-6, "dc0": [0.45, 0.55],
           "dc1": [0.6, 0.7], "z_start": 0.01, "thickness": 6e-6, "nr_steps": 6,
           "x_positions": [float(s) for s in steps], "substrate_thickness": 2e-6,
           "mat_a": ["Au", 19.32], "mat_b": None, "mat_substrate": ["Si", 2.34]}
    els = [dba, grating(1.0e-6, 0.02, 4e-6, 4, ["Si", 2.34])]
    z_det = 0.05
    cfg = {"sim_params": sim_params(N, dx, z_det, 40e-6, 1e-6, 1.0, chunk),
           "use_disk_vector": False, "save_final_u_vectors": False, "dtype": dtype, "elements": els}
    subs = [{"energy": 30000.0, "source": {"type": "point", "x": 0.0, "z": 0.0},
             "deltabeta_table": [[["Si", 2.34], [5.1e-7, 1.6e-10]], [["Au", 19.32], [3.6e-6, 2.1e-7]]]}]
    return cfg, subs, [math.pi / 2] * 3


def random_case(tmp: Path, seed: int):
    """A seeded random valid sim dir for cross-checks: N = 2^12..2^15, a plane
    wave or a point source, 1-3 elements drawn from grating / DBA grating /
    sample (one of them phase stepped over 2-3 positions), random substrates,
    chunk sizes, cutoff angles and detector.  Returns (config, sub-configs,
    cutoff angles, sample grid or None)."""
    rng = np.random.default_rng(seed)
    N = int(2 ** rng.integers(12, 16))
    dx = float(rng.choice([1e-9, 2e-9]))
    energy = float(rng.uniform(20e3, 60e3))
    chunk = int(rng.choice([N, N // 4, 1000, 3000]))
    n_el = int(rng.integers(1, 4))
    stepped = int(rng.integers(0, n_el))
    n_steps = int(rng.integers(2, 4))
    mats = [["Si", 2.34], ["Au", 19.32]]
    z = float(rng.uniform(0.002, 0.01))
    els, grid = [], None
    for i in range(n_el):
        kind = str(rng.choice(["grating", "dba", "sample"])) if grid is None else str(rng.choice(["grating", "dba"]))
        pitch = float(rng.uniform(0.3e-6, 1.5e-6))
        xs = [float(rng.uniform(-pitch, pitch)) for _ in range(n_steps if i == stepped else 1)]
        sub = float(rng.choice([0.0, float(rng.uniform(1e-6, 5e-6))]))
        if kind == "grating":
            e = grating(pitch, z, float(rng.uniform(2e-6, 10e-6)), int(rng.integers(1, 5)), mats[int(rng.integers(0, 2))],
                        x_positions=xs, dc=(float(rng.uniform(0.3, 0.7)), float(rng.uniform(0.3, 0.7))),
                        substrate=sub, mat_b=None if rng.random() < 0.5 else mats[0],
                        mat_s=mats[0] if sub else None)
        elif kind == "dba":
            e = {"type": "dba_grating", "pitch0": pitch, "pitch1": pitch * float(rng.choice([2.0, 3.0, 2.5])),
                 "dc0": [float(rng.uniform(0.3, 0.7)), float(rng.uniform(0.3, 0.7))],
                 "dc1": [float(rng.uniform(0.3, 0.7)), float(rng.uniform(0.3, 0.7))], "z_start": z,
                 "thickness": float(rng.uniform(2e-6, 8e-6)), "nr_steps": int(rng.integers(1, 5)), "x_positions": xs,
                 "substrate_thickness": sub, "mat_a": mats[1], "mat_b": None, "mat_substrate": mats[0] if sub else None}
        else:
            rows, cols = int(rng.integers(1, 4)), int(rng.integers(3, 12))
            grid = rng.integers(0, 3, size=(rows, cols)).astype(np.int8)
            e = {"type": "sample", "z_start": z, "pixel_size_x": float(rng.uniform(1e-7, 6e-7)),
                 "pixel_size_z": float(rng.uniform(1e-6, 3e-6)), "grid_path": "grid.npy", "materials": mats,
                 "x_positions": xs}
        els.append(e)
        thick = (e["pixel_size_z"] * grid.shape[0]) if kind == "sample" else e["thickness"] + e["substrate_thickness"]
        z = z + thick + float(rng.uniform(0.0, 0.01))
    z_det = z + float(rng.uniform(0.0, 0.02))
    det = min(N * dx / 2, float(rng.uniform(2e-6, 2e-5)))
    cfg = {"sim_params": sim_params(N, dx, z_det, det, float(rng.choice([2e-7, 5e-7, 1e-6])), 1.0, chunk),
           "use_disk_vector": False, "save_final_u_vectors": False, "dtype": "c16", "elements": els}
    if rng.random() < 0.5:
        src = {"type": "point", "x": float(rng.uniform(-1e-7, 1e-7)), "z": -float(rng.uniform(0.5, 2.0))}
    else:
        src = {"type": "vector", "input_path": write_ones(tmp / "ones.npy", N), "z": 0.0}
    subs = [{"energy": energy, "source": src,
             "deltabeta_table": [[["Si", 2.34], [5.1e-7, 1.6e-10]], [["Au", 19.32], [3.6e-6, 2.1e-7]]]}]
    ang = sorted(float(rng.choice([float(rng.uniform(0.2e-3, 2e-3)), math.pi / 2])) for _ in range(n_el + 1))
    return cfg, subs, ang, grid


def sphere_grid(radius: float, px: float, pz: float, seed: int = 0, pitch: float = 4.2e-6):
    """Sphere phantom as a sample grid: material 1 where the pixel centre
    ((ix + 1/2 - nx/2) px - s, (iz + 1/2 - nz/2) pz) lies within `radius` of
    the centre, 0 elsewhere; the lateral offset s of the centre is uniform
    within one grating pitch from default_rng(seed)."""
    rng = np.random.default_rng(seed)
    s = float(rng.uniform(-pitch / 2, pitch / 2))
    nx = int(round(2 * radius / px)) + 2
    nz = int(round(2 * radius / pz))
    x = (np.arange(nx) + 0.5 - nx / 2) * px - s
    z = (np.arange(nz) + 0.5 - nz / 2) * pz
    return (x[None, :] ** 2 + z[:, None] ** 2 <= radius * radius).astype(np.int8)


def talbot_lau_phantom(tmp: Path, N=2 ** 26, slices=64, det_size=None, dtype="c8", pz=1e-6, seed=0):
    """The Talbot-Lau chain with its optional sample: a sphere of radius
    50 um, delta = 3e-7, beta = 0, on 1 um x pixels and `pz` z pixels,
    starting 0.1 m behind G1 (100 extra propagation steps at pz = 1 um)."""
    cfg, subs, ang = talbot_lau(tmp, N=N, slices=slices, det_size=det_size, dtype=dtype)
    grid = sphere_grid(50e-6, 1e-6, pz, seed)
    np.save(tmp / "sphere.npy", grid)
    sample = {"type": "sample", "z_start": 0.3, "pixel_size_x": 1e-6, "pixel_size_z": pz,
              "grid_path": "sphere.npy", "materials": [["sphere", 1.0]], "x_positions": [0.0]}
    cfg["elements"].insert(2, sample)
    subs[0]["deltabeta_table"].append([["sphere", 1.0], [3e-7, 0.0]])
    return cfg, subs, [math.pi / 2] * 5
