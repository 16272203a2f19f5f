"""Golden detector outputs of the reduced benchmark configurations, computed
with the complex128 oracle (oracle/bigwave.py) and stored as small .npz files.

Inputs are rebuilt deterministically by tests/simcases.py; each fixture keeps
the parameters it was made with and the float64 detector rows.

Usage:  python tests/golden/make_fixtures.py
"""
import math
import sys
import tempfile
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(HERE.parent))

from oracle import bigwave as bw  # noqa: E402
import simcases  # noqa: E402


def simdir_rows(case_fn, **kw):
    import yaml
    with tempfile.TemporaryDirectory() as t:
        t = Path(t)
        cfg, subs, ang = case_fn(t, **kw)
        sd = t / "sim"
        (sd / "00000000").mkdir(parents=True)
        (sd / "config.yaml").write_text(yaml.safe_dump(cfg))
        (sd / "computed.yaml").write_text(yaml.safe_dump({"cutoff_angles": ang}))
        (sd / "00000000" / "subconfig.yaml").write_text(yaml.safe_dump(subs[0]))
        return bw.run_simdir(sd, 0, dtype=np.complex128)


def poly_rows(n_energies, N, slices):
    """Weighted polychromatic sum over the monochromatic chain (plane wave, one Si phase grating and a hop)."""
    dx, h = 2e-9, 30e-6
    total = None
    for e_ev, w, db in simcases.poly_members(n_energies, N, slices):
        wl = bw.convert_energy_wavelength(e_ev)
        p = bw.SimParams(N=N, dx=dx, z_detector=h + 0.1, detector_size=1e-4, detector_pixel_size_x=1e-6,
                         detector_pixel_size_y=1.0, wl=wl, chunk_size=N)
        g = bw.Grating(pitch=4.2e-6, dc=(0.5, 0.5), z_start=0.0, thickness=h, nr_steps=slices,
                       x_positions=np.array([0.0]), substrate_thickness=0.0, db_a=db)
        rows = bw.run_simulation(p, bw.VectorSource(np.ones(N, dtype=np.complex128), 0.0), [g],
                                 [math.pi / 2] * 2)
        total = w * rows[0] if total is None else total + w * rows[0]
    return total[None, :]


def main():
    fx = {}
    fx["mono_n16"] = dict(rows=simdir_rows(simcases.mono_chain, N=2 ** 16, slices=32, det_size=1e-4),
                          N=2 ** 16, slices=32, det_size=1e-4)
    fx["talbot_lau_n18_steps8"] = dict(rows=simdir_rows(simcases.talbot_lau, N=2 ** 18, slices=8, det_size=1e-4,
                                                  g2_steps=8),
                                 N=2 ** 18, slices=8, det_size=1e-4, g2_steps=8)
    fx["poly_n16_e8"] = dict(rows=poly_rows(8, 2 ** 16, 16), N=2 ** 16, slices=16, n_energies=8)
    for name, d in fx.items():
        np.savez(HERE / f"{name}.npz", **{k: np.asarray(v) for k, v in d.items()})
        print(name, d["rows"].shape)


if __name__ == "__main__":
    main()
