"""delta/beta of a material from Chantler form factors (test oracle).

TEST INFRASTRUCTURE ONLY -- see oracle/__init__.py.  Restates
nist_lookup/nist_lookup/xraydb_plugin.py:134-180 (`xray_delta_beta`) and the
Chantler interpolation of nist_lookup/nist_lookup/xraydb.py:371-403
(`_getChantler`): f1 by scipy UnivariateSpline(s=1) over the table window
[-5, +6] around the energy, f2 and mu by log-log linear interpolation.  The
table rows come from the JSON fixture tests/golden/chantler_si_au.json
(extracted from the reference's SQLite data file by
tests/golden/make_chantler_fixture.py).  Physical constants are taken from
scipy.constants, as nist_lookup/nist_lookup/physical_constants.py does.

Pinned by the end-to-end golden vector of big-wave/test.py:853-960, which uses
Si (2.34 g/cm^3) and Au (19.32 g/cm^3) near 46 keV.
"""
from __future__ import annotations

import json
import re
from functools import lru_cache
from pathlib import Path

import numpy as np
import scipy.constants as consts
from scipy.interpolate import UnivariateSpline

FIXTURE = Path(__file__).resolve().parent.parent / "tests" / "golden" / "chantler_si_au.json"

AVOGADRO = consts.Avogadro
PLANCK_HC = 1.0e10 * consts.Planck * consts.c / consts.e  # eV * Angstrom
R_ELECTRON_CM = 100.0 * consts.physical_constants["classical electron radius"][0]


@lru_cache(maxsize=None)
def _table() -> dict:
    return json.loads(FIXTURE.read_text())


def parse_formula(formula: str) -> dict:
    """Element -> count for formulas like 'Si', 'H2O', 'Mg(SO4)2'
    (nist_lookup/nist_lookup/chemparser.py semantics, counts as floats)."""
    tokens = re.findall(r"[A-Z][a-z]*|[0-9]+\.?[0-9]*(?:[eE][-+]?[0-9]+)?|[()]", formula)
    if "".join(tokens) != formula:
        raise ValueError(f"cannot parse formula {formula!r}")
    pos = 0

    def seq() -> dict:
        nonlocal pos
        out: dict = {}
        while pos < len(tokens) and tokens[pos] != ")":
            tok = tokens[pos]
            pos += 1
            if tok == "(":
                inner = seq()
                if pos >= len(tokens) or tokens[pos] != ")":
                    raise ValueError("expected right paren")
                pos += 1
            elif tok[0].isalpha():
                inner = {tok: 1}
            else:
                raise ValueError(f"unexpected number in {formula!r}")
            mult = 1.0
            if pos < len(tokens) and tokens[pos][0].isdigit():
                mult = float(tokens[pos])
                pos += 1
            for k, v in inner.items():
                out[k] = out.get(k, 0) + v * mult
        return out

    res = seq()
    if pos != len(tokens):
        raise ValueError("unbalanced parentheses")
    return res


def chantler(symbol: str, energy: float, column: str) -> float:
    """Chantler f1 / f2 / mu_photo / mu_total at one energy (eV);
    nist_lookup/nist_lookup/xraydb.py:371-403."""
    row = _table()[symbol]
    te_all = np.array(row["energy"])
    e = np.array([energy], dtype=np.float64)
    emin, emax = e.min(), e.max()
    nemin = max(0, -5 + max(np.where(te_all <= emin)[0]))
    nemax = min(len(te_all), 6 + max(np.where(te_all <= emax)[0]))
    region = np.arange(nemin, nemax)
    te = te_all[region]
    ty = np.array(row[column])[region]
    if column == "f1":
        out = UnivariateSpline(te, ty, s=1)(e)
    else:
        out = np.exp(np.interp(np.log(e), np.log(te), np.log(ty)))
    return out[0]


def xray_delta_beta(material: str, density: float, energy: float):
    """(delta, beta) of a material at `energy` eV;
    nist_lookup/nist_lookup/xraydb_plugin.py:134-180 (total cross-section)."""
    lamb_cm = 1.0e-8 * PLANCK_HC / energy
    total_mass, delta, beta = 0.0, 0.0, 0.0
    for sym, number in parse_formula(material).items():
        row = _table()[sym]
        f1 = chantler(sym, energy, "f1") + row["Z"]
        f2 = chantler(sym, energy, "f2")
        mu_photo = chantler(sym, energy, "mu_photo")
        mu_total = chantler(sym, energy, "mu_total")
        weight = density * number * AVOGADRO
        delta += weight * f1
        beta += weight * f2 * (mu_total / mu_photo)
        total_mass += number * row["molar_mass"]
    scale = lamb_cm * lamb_cm * R_ELECTRON_CM / (2 * np.pi * total_mass)
    return delta * scale, beta * scale
