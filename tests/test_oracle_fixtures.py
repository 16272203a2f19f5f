"""The committed fixtures are reproduced by the oracle (pins the files)."""
import sys
from pathlib import Path

import numpy as np

sys.path.insert(0, str(Path(__file__).resolve().parent / "golden"))
import make_fixtures  # noqa: E402
import simcases  # noqa: E402


def test_mono_fixture_reproduced(golden_dir):
    fx = np.load(golden_dir / "mono_n16.npz")
    rows = make_fixtures.simdir_rows(simcases.mono_chain, N=int(fx["N"]), slices=int(fx["slices"]),
                                     det_size=float(fx["det_size"]))
    assert np.array_equal(rows, fx["rows"])


def test_poly_fixture_reproduced(golden_dir):
    fx = np.load(golden_dir / "poly_n16_e8.npz")
    rows = make_fixtures.poly_rows(int(fx["n_energies"]), int(fx["N"]), int(fx["slices"]))
    assert np.array_equal(rows, fx["rows"])
