"""Sim-dir producer: writes config.yaml / computed.yaml / %08d/subconfig.yaml.

Follows big-wave's setup rules (big-wave/multisim.py:166-521): point sources
drawn from one numpy RandomState (energies first, uniform or from a spectrum,
then x positions, multisim.py:338-379), continuation from a base simulation's
saved u vectors (multisim.py:404-462), the save_and_exit cut
(multisim.py:166-192), the Nyquist check of the analytic source
(propagation.py:31-45), cutoff angles from the detector / element geometry
(propagation.py:178-276) and the reflection check on N (multisim.py:480-487).
Delta/beta values come from ravesim.xray (the Chantler tables, as big-wave's
nist_lookup call at multisim.py:387 gives them) unless the caller supplies a
table or a function of (material, density, energy).
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


def _table(materials, energy: float, deltabeta: Optional[DeltaBeta]) -> list:
    rows = []
    if deltabeta is None:
        from .xray import xray_delta_beta as deltabeta
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


def load_spectrum(path: Union[str, Path]) -> Tuple[np.ndarray, np.ndarray]:
    """(energy, pdf) arrays of a spectrum file (big-wave/multisim.py:245-247
    reads the `energy` and `pdf` datasets of an HDF5 file).  HDF5 needs h5py;
    an .npz with the same two keys is read with numpy."""
    path = Path(path)
    if path.suffix == ".npz":
        with np.load(path, allow_pickle=False) as z:
            return np.asarray(z["energy"]), np.asarray(z["pdf"])
    import h5py  # the HDF5 form of big-wave's spectrum files
    with h5py.File(path, "r") as f:
        return np.array(f["energy"]), np.array(f["pdf"])


def energies_from_spectrum(energy: np.ndarray, pdf: np.ndarray, rs: np.random.RandomState,
                           energy_range, n: int) -> np.ndarray:
    """Energies drawn from a tabulated spectrum cropped to energy_range: a bin
    by rs.choice with p = pdf / sum(pdf), then a uniform offset within the bin
    spacing (big-wave/multisim.py:250-279, same RandomState draws)."""
    mask = np.logical_and(energy >= energy_range[0], energy < energy_range[1])
    se, si = energy[mask], pdf[mask]
    if len(se) < 2:
        raise ValueError("Not enough energy points of the spectrum lie within the given range")
    bases = rs.choice(se, size=n, replace=True, p=si / np.sum(si))
    offsets = rs.rand(n) * (se[1] - se[0])
    return bases + offsets


def _check_inputs(sp: dict, z_src: float, els: list, heights: Sequence[float],
                  angles: Sequence[float]) -> None:
    """big-wave/multisim.py:195-242 (element validity itself is checked again
    when a source is run)."""
    tol = 1e-8
    if len(angles) != len(els) + 1:
        raise ValueError("cutoff angles must have len(elements) + 1 entries")
    if els and not z_src <= float(els[0]["z_start"]):
        raise ValueError("The source must not be at a greater z-coordinate than the first optical element")
    steps = None
    cur = z_src
    z_det = float(sp["z_detector"])
    for i, (el, h) in enumerate(zip(els, heights)):
        n = len(el.get("x_positions", [0]))
        if n > 1:
            if steps is not None and steps != n:
                raise ValueError("If multiple elements have phase stepping enabled, then all elements "
                                 "must have the same number of steps.")
            steps = n
        z0 = float(el["z_start"])
        if not cur <= z0 + tol:
            raise ValueError("Overlapping optical elements are not supported")
        cur = z0 + h
        if el["type"] == "save_and_exit":
            if i == len(els) - 1:
                if z_det != z0:
                    raise ValueError("A save_and_exit element at the end must have the same z coordinate "
                                     "as the detector")
            elif float(els[i + 1]["z_start"]) != z0:
                raise ValueError("A save_and_exit element must have the same z coordinate as the next "
                                 "optical element")
    if not cur <= z_det + tol:
        raise ValueError("Detector must be after the last optical element")
    for a, b in zip(angles[:-1], angles[1:]):
        if not a <= b:
            raise ValueError("cutoff angles must be increasing")
    for a in angles:
        if not 0 <= a <= math.pi / 2:
            raise ValueError("cutoff angles must be between 0 and pi/2")


def _reduce_for_save_and_exit(cfg: dict, max_x_list: Sequence[float]) -> float:
    """Cut the elements from a save_and_exit marker on, move the detector to
    it and save the final u vectors (big-wave/multisim.py:166-192).  Returns
    the max |x| at the (possibly moved) detector."""
    max_x = max_x_list[-1]
    kept = []
    for i, el in enumerate(cfg["elements"]):
        if el["type"] == "save_and_exit":
            max_x = max_x_list[i]
            cfg["sim_params"]["z_detector"] = el["z_start"]
            cfg["save_final_u_vectors"] = True
            cfg["elements"] = kept
            break
        kept.append(el)
    return max_x


def setup_simulation(config: dict, sim_dir: Union[str, Path], deltabeta: Optional[DeltaBeta] = None,
                     config_dir: Optional[Union[str, Path]] = None,
                     spectrum: Optional[Tuple[np.ndarray, np.ndarray]] = None) -> Path:
    """Set up `sim_dir` from a big-wave config (big-wave/multisim.py:282-521).

    multisource "points": energies uniform in energy_range, or drawn from a
    spectrum (multisource["spectrum"], a file relative to `config_dir`, or the
    `spectrum` argument as (energy, pdf) arrays), then x positions, from one
    RandomState(seed) (seed -1: a random seed).  multisource "vectors": one
    vector source per source of `base_sim_dir`, reading its
    u_{input_u_index:04}.npy at the base detector's z.  A save_and_exit element
    ends the setup there and turns on save_final_u_vectors.  Sample grids are
    read relative to `config_dir` and copied into the sim dir.  `deltabeta`
    None: every material's delta and beta at each source energy from
    ravesim.xray.xray_delta_beta (big-wave/multisim.py:387).
    """
    import random

    cfg = copy.deepcopy(config)
    sp = cfg["sim_params"]
    N, dx = int(sp["N"]), float(sp["dx"])
    if not N * dx >= float(sp["detector_size"]):
        raise ValueError("Detector is bigger than simulation space")
    els = cfg["elements"]
    ms = cfg["multisource"]
    mats = _materials(els)
    heights = [_thickness(e, config_dir) if e["type"] != "save_and_exit" else 0.0 for e in els]
    if ms["type"] == "points":
        n = int(ms["nr_source_points"])
        e_rng = (float(ms["energy_range"][0]), float(ms["energy_range"][1]))
        x_rng = (float(ms["x_range"][0]), float(ms["x_range"][1]))
        z_src = float(ms["z"])
        seed = int(ms["seed"])
        if seed == -1:
            seed = random.randint(0, 1000000)
        first_z = float(els[0]["z_start"]) if els else float(sp["z_detector"])
        wl_hi = energy_to_wavelength(e_rng[1])
        grid_density_check(first_z - z_src, x_rng[0], N, dx, wl_hi)
        grid_density_check(first_z - z_src, x_rng[1], N, dx, wl_hi)
        rs = np.random.RandomState(seed)
        if "spectrum" in ms or spectrum is not None:
            se, si = spectrum if spectrum is not None else load_spectrum(Path(config_dir) / ms["spectrum"])
            energies = energies_from_spectrum(np.asarray(se), np.asarray(si), rs, e_rng, n)
        else:
            energies = rs.rand(n) * (e_rng[1] - e_rng[0]) + e_rng[0]
        xs = rs.rand(n) * (x_rng[1] - x_rng[0]) + x_rng[0]
        subs = [{"source": {"type": "point", "x": float(xs[i]), "z": z_src},
                 "energy": float(energies[i]),
                 "deltabeta_table": _table(mats, float(energies[i]), deltabeta)} for i in range(n)]
        points = [{"x": float(xs[i]), "z": z_src, "energy": float(energies[i])} for i in range(n)]
        max_x0 = max(abs(x_rng[0]), abs(x_rng[1]))
    elif ms["type"] == "vectors":
        base = Path(ms["base_sim_dir"])
        idx = int(ms["input_u_index"])
        base_cfg = yaml.safe_load((base / "config.yaml").read_text())
        bsp = base_cfg["sim_params"]
        z_src = float(bsp["z_detector"])
        if not (int(bsp["N"]) == N and float(bsp["dx"]) == dx):
            raise ValueError(f"simulation N and dx must match those of the base simulation in {base}")
        if not base_cfg["save_final_u_vectors"]:
            raise ValueError("The base simulation must have save_final_u_vectors set to true or have a "
                             "save_and_exit element")
        base_comp = yaml.safe_load((base / "computed.yaml").read_text())
        e_rng = (float(base_comp["energy_range"][0]), float(base_comp["energy_range"][1]))
        max_x0 = float(base_comp["max_x"])
        n = len([d for d in base.glob("[0-9]" * 8) if d.is_dir()])
        subs = []
        for i in range(n):
            bsub = base / f"{i:08d}"
            energy = float(yaml.safe_load((bsub / "subconfig.yaml").read_text())["energy"])
            subs.append({"source": {"type": "vector", "input_path": str(bsub / f"u_{idx:04}.npy"),
                                    "z": z_src},
                         "energy": energy, "deltabeta_table": _table(mats, energy, deltabeta)})
        points = base_comp["source_points"]
    else:
        raise ValueError(f"Unknown multisource type {ms['type']}.")
    angles, max_x_list = cutoff_angles(float(sp["detector_size"]), dx, e_rng, z_src,
                                       [float(e["z_start"]) for e in els], heights,
                                       float(sp["z_detector"]), max_x0)
    required_N = int(np.ceil(max_x_list[-1] / dx) * 2)
    if N < required_N:
        raise ValueError(f"Reflections at the boundary might occur: N={N}, required_N={required_N}")
    grids = []
    for el in els:
        if el["type"] == "sample":
            grids.append(Path(config_dir) / el["grid_path"])
            el["grid_path"] = Path(el["grid_path"]).name
    _check_inputs(sp, z_src, els, heights, angles)  # the full setup, before the reduction
    max_x = _reduce_for_save_and_exit(cfg, max_x_list)
    Path(sim_dir).mkdir(parents=True, exist_ok=True)
    for gf in grids:
        shutil.copyfile(gf, Path(sim_dir) / gf.name)
    return write_simdir(sim_dir, cfg, subs, angles,
                        {"max_x": max_x, "energy_range": list(e_rng), "source_points": points})
