"""ravesim.xray: the product's delta/beta (big-wave/multisim.py:387 ->
nist_lookup/nist_lookup/xraydb_plugin.py:134-180).

Pinned by the oracle's restatement for Si and Au over 20-70 keV (the oracle
is itself pinned by big-wave's end-to-end golden vector), by big-wave's
subconfig.yaml example, by the table rows of the golden fixture, by
chemparser.py's own docstring examples, and by what the formula fixes:
delta and beta are linear in the density and in the composition."""
import json

import numpy as np
import pytest

from oracle import xray as oxr


def test_tables_match_the_golden_fixture_rows(golden_dir):
    from ravesim import xray
    fix = json.loads((golden_dir / "chantler_si_au.json").read_text())
    _, cols = xray._tables()
    for sym in ("Si", "Au"):
        assert xray.element(sym) == (fix[sym]["Z"], fix[sym]["molar_mass"])
        for c in ("energy", "f1", "f2", "mu_photo", "mu_total"):
            assert np.array_equal(cols[f"{sym}.{c}"], np.array(fix[sym][c]))
    assert len(xray._tables()[0]) == 92


@pytest.mark.parametrize("mat,rho", [("Si", 2.34), ("Au", 19.32)])
def test_equals_oracle_over_20_to_70_kev(mat, rho):
    from ravesim import xray
    for e in np.linspace(20e3, 70e3, 26):
        assert xray.xray_delta_beta(mat, rho, e) == oxr.xray_delta_beta(mat, rho, e)


def test_subconfig_example_at_46_kev():
    # big-wave/subconfig.yaml:9-12 lists Si/Au deltabeta at 46 keV to 5 digits
    from ravesim import xray
    for (mat, rho), (d, b) in {("Si", 2.34): (2.2888e-7, 2.5847e-10),
                               ("Au", 19.32): (1.4981e-6, 3.6174e-8)}.items():
        de, be = xray.xray_delta_beta(mat, rho, 46000.0)
        assert abs(de / d - 1) < 1e-4 and abs(be / b - 1) < 1e-4


def test_formula_grammar():
    from ravesim import xray
    # nist_lookup/nist_lookup/chemparser.py docstring examples
    assert xray.composition("Mn(SO4)2(H2O)7") == {"Mn": 1, "S": 2.0, "O": 15.0, "H": 14.0}
    assert xray.composition("Zn1.e-5Fe3O4") == {"Zn": 1e-05, "Fe": 3.0, "O": 4.0}
    assert xray.composition("CH3CH2OH") == {"C": 2, "H": 6, "O": 1}
    for bad in ("", "Xy", "2H", "(H2O", "H2O)", "h2o"):
        with pytest.raises(ValueError):
            xray.composition(bad)


def test_linear_in_density_and_composition():
    from ravesim import xray
    e = 33000.0
    d1, b1 = xray.xray_delta_beta("SiO2", 2.2, e)
    d2, b2 = xray.xray_delta_beta("SiO2", 4.4, e)
    assert (d2, b2) == (2 * d1, 2 * b1)
    # a compound is the mass-weighted sum of its elements at unit density:
    # delta(SiO2, rho) = rho / M(SiO2) * (M(Si) delta(Si, 1) + 2 M(O) delta(O, 1))
    mSi, mO = xray.element("Si")[1], xray.element("O")[1]
    M = mSi + 2 * mO
    dSi, bSi = xray.xray_delta_beta("Si", 1.0, e)
    dO, bO = xray.xray_delta_beta("O", 1.0, e)
    assert d1 == pytest.approx(2.2 / M * (mSi * dSi + 2 * mO * dO), rel=1e-13)
    assert b1 == pytest.approx(2.2 / M * (mSi * bSi + 2 * mO * bO), rel=1e-13)
    # (H2O)2 is H2O
    assert xray.xray_delta_beta("(H2O)2", 1.0, e) == pytest.approx(xray.xray_delta_beta("H2O", 1.0, e), rel=1e-15)


def test_setup_simulation_fills_the_table_itself(golden_dir, tmp_path):
    import yaml
    from ravesim.simdir import setup_simulation
    d = json.loads((golden_dir / "bigwave_snapshot.json").read_text())
    sd = setup_simulation(dict(d["known_config"]), tmp_path / "sim")
    sub = yaml.safe_load((sd / "00000000" / "subconfig.yaml").read_text())
    e = sub["energy"]
    for (mat, rho), db in sub["deltabeta_table"]:
        assert tuple(db) == oxr.xray_delta_beta(mat, rho, e)
