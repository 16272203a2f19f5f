"""Sim-dir producer: writes config.yaml / computed.yaml / %08d/subconfig.yaml.

Follows big-wave's setup rules (big-wave/multisim.py:282-521): point sources
drawn from one numpy RandomState (energies first, then x positions,
multisim.py:358-379), the Nyquist check of the analytic source
(propagation.py:31-45), cutoff angles from the detector / element geometry
(propagation.py:178-276) and the reflection check on N (multisim.py:480-487).
Delta/beta values are supplied by the caller (a table or a function of
(material, density, energy)) instead of being looked up in nist_lookup.
"""
from __future__ import annotations

import copy
import math
import shutil
from pathlib import Path
from typing import Callable, Optional, Sequence, Tuple, Union

import numpy as np
import yaml

from .core import energy_to_wavelength

DeltaBeta = Union[Callable[[str, float, float], Tuple[float, float]], dict]


def grid_density_check(dz: float, x_source: float, N: int, dx: float, wl: float) -> None:
    def dist(idx: int) -> float:
        return math.sqrt(((N / 2 - idx) * dx + abs(x_source)) ** 2 + dz ** 2)
    if not abs(dist(0) - dist(1)) <= wl / 2:
        raise ValueError("Nyquist not satisfied, dx too large for the point source")


def cutoff_angles(detector_size: float, dx: float, energy_range, z_source: float,
                  z_starts: Sequence[float], heights: Sequence[float], z_detector: float,
                  max_x: float):
    """Cutoff angle per free-space section and the max |x| reached at each
    element start / the detector (big-wave/propagation.py:178-276)."""
    wl_lo = energy_to_wavelength(energy_range[0])
    wl_hi = energy_to_wavelength(energy_range[1])
    direct = math.atan((detector_size / 2 + max_x) / z_detector)
    if 2 / (math.sin(direct) / wl_hi) < dx:
        raise ValueError("grid too coarse: the highest energy does not cover the detector")
    g_emin = math.tan(math.asin(wl_lo * 2 / dx)) if wl_lo * 2 / dx < 1 else math.inf
    grads, xs = [], []
    cx, cz = max_x, z_source
    for nz in list(z_starts) + [z_detector]:
        if cz == z_detector:
            raise ValueError("an element has the same z coordinate as the detector")
        grads.append(min((detector_size / 2 + cx) / (z_detector - cz), g_emin))
        cx += grads[-1] * (nz - cz)
        xs.append(float(cx))
        cz = nz
    if z_starts and abs(z_detector - (z_starts[-1] + heights[-1])) < 1e-8:
        grads[-1] = grads[-2]
        xs[-1] = xs[-2] + grads[-1] * heights[-1]
    return [float(math.atan(g)) for g in grads], xs


def _materials(elements) -> list:
    """Deduplicated materials in element order (big-wave/optical_element.py:513-545)."""
    mats = []

    def push(m):
        if m is not None and [m[0], float(m[1])] not in mats:
            mats.append([m[0], float(m[1])])

    for el in elements:
        if el["type"] == "sample":
            for m in el["materials"]:
                push(m)
        else:
            for k in ("mat_a", "mat_b", "mat_substrate"):
                push(el.get(k))
    return mats


def _thickness(el: dict, config_dir: Optional[Path]) -> float:
    """get_thickness() of a grating, DBA grating or sample element
    (big-wave/optical_element.py:132, 245, 420)."""
    if el["type"] == "sample":
        if config_dir is None:
            raise ValueError("a sample element needs config_dir to find its grid")
        grid = np.load(Path(config_dir) / el["grid_path"])
        return float(el["pixel_size_z"]) * grid.shape[0]
    return float(el["thickness"]) + float(el.get("substrate_thickness", 0.0))


def _table(materials, energy: float, deltabeta: DeltaBeta) -> list:
    rows = []
    for name, dens in materials:
        if callable(deltabeta):
            d, b = deltabeta(name, dens, energy)
        else:
            d, b = deltabeta[(name, dens)]
        rows.append([[name, float(dens)], [float(d), float(b)]])
    return rows


def write_simdir(sim_dir: Union[str, Path], config: dict, sources: Sequence[dict],
                 cutoff: Sequence[float], computed_extra: Optional[dict] = None) -> Path:
    """Write a sim dir from a config dict (big-wave schema) and per-source
    sub-configs ({"source": ..., "energy": ..., "deltabeta_table": ...})."""
    sim_dir = Path(sim_dir)
    sim_dir.mkdir(parents=True, exist_ok=True)
    dump = dict(default_flow_style=None, sort_keys=False)
    (sim_dir / "config.yaml").write_text(yaml.safe_dump(config, **dump))
    comp = {"cutoff_angles": [float(a) for a in cutoff]}
    comp.update(computed_extra or {})
    (sim_dir / "computed.yaml").write_text(yaml.safe_dump(comp, **dump))
    for i, sub in enumerate(sources):
        d = sim_dir / f"{i:08d}"
        d.mkdir(exist_ok=True)
        (d / "subconfig.yaml").write_text(yaml.safe_dump(sub, **dump))
    return sim_dir


def setup_simulation(config: dict, sim_dir: Union[str, Path], deltabeta: DeltaBeta,
                     config_dir: Optional[Union[str, Path]] = None) -> Path:
    """Point-source multisource setup (big-wave/multisim.py:282-521, "points"
    branch) into `sim_dir`.  Sample grids are read relative to `config_dir`
    and copied into the sim dir under their base names (multisim.py:489-499)."""
    cfg = copy.deepcopy(config)
    sp = cfg["sim_params"]
    N, dx = int(sp["N"]), float(sp["dx"])
    if not N * dx >= float(sp["detector_size"]):
        raise ValueError("Detector is bigger than simulation space")
    els = cfg["elements"]
    ms = cfg["multisource"]
    if ms["type"] != "points":
        raise ValueError("only the 'points' multisource is supported")
    n = int(ms["nr_source_points"])
    e_rng = (float(ms["energy_range"][0]), float(ms["energy_range"][1]))
    x_rng = (float(ms["x_range"][0]), float(ms["x_range"][1]))
    z_src = float(ms["z"])
    first_z = float(els[0]["z_start"]) if els else float(sp["z_detector"])
    wl_hi = energy_to_wavelength(e_rng[1])
    grid_density_check(first_z - z_src, x_rng[0], N, dx, wl_hi)
    grid_density_check(first_z - z_src, x_rng[1], N, dx, wl_hi)
    rs = np.random.RandomState(int(ms["seed"]))
    energies = rs.rand(n) * (e_rng[1] - e_rng[0]) + e_rng[0]
    xs = rs.rand(n) * (x_rng[1] - x_rng[0]) + x_rng[0]
    mats = _materials(els)
    subs = [{"source": {"type": "point", "x": float(xs[i]), "z": z_src},
             "energy": float(energies[i]),
             "deltabeta_table": _table(mats, float(energies[i]), deltabeta)} for i in range(n)]
    heights = [_thickness(e, config_dir) for e in els]
    angles, max_x_list = cutoff_angles(float(sp["detector_size"]), dx, e_rng, z_src,
                                       [float(e["z_start"]) for e in els], heights,
                                       float(sp["z_detector"]), max(abs(x_rng[0]), abs(x_rng[1])))
    required_N = int(np.ceil(max_x_list[-1] / dx) * 2)
    if N < required_N:
        raise ValueError(f"Reflections at the boundary might occur: N={N}, required_N={required_N}")
    points = [{"x": float(xs[i]), "z": z_src, "energy": float(energies[i])} for i in range(n)]
    grids = []
    for el in els:
        if el["type"] == "sample":
            grids.append(Path(config_dir) / el["grid_path"])
            el["grid_path"] = Path(el["grid_path"]).name
    Path(sim_dir).mkdir(parents=True, exist_ok=True)
    for gf in grids:
        shutil.copyfile(gf, Path(sim_dir) / gf.name)
    return write_simdir(sim_dir, cfg, subs, angles,
                        {"max_x": max_x_list[-1], "energy_range": list(e_rng), "source_points": points})
