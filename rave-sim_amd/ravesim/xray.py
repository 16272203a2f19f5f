"""Delta and beta of the X-ray refractive index n = 1 - delta - i beta.

big-wave fills every source's `deltabeta_table` with nist_lookup's
`xray_delta_beta(material, density, energy)` (big-wave/multisim.py:387 via
optical_element.py:549-564).  This module computes the same quantity for the
sim-dir producer (`ravesim.simdir.setup_simulation`):

- the material is a chemical formula with the grammar of
  nist_lookup/nist_lookup/chemparser.py (element symbols, parentheses,
  integer / decimal / exponent counts);
- per element, the Chantler f1 at the energy comes from a smoothing spline
  (scipy UnivariateSpline, s = 1) through the tabulated points from five
  below to six above the energy, f2, mu_photo and mu_total from linear
  interpolation of their logarithms (nist_lookup/nist_lookup/xraydb.py:
  371-403);
- delta and beta are the density-weighted sums of (f1 + Z) and
  f2 mu_total / mu_photo, scaled by lambda^2 r_e / (2 pi M)
  (nist_lookup/nist_lookup/xraydb_plugin.py:134-180, total cross-section);
- constants come from scipy.constants (nist_lookup/nist_lookup/
  physical_constants.py).

The tables of all 92 elements are read from data/chantler.npz, extracted
from nist_lookup's xrayref.db by data/make_chantler_table.py.
"""
from __future__ import annotations

import re
from functools import lru_cache
from pathlib import Path
from typing import Dict, Tuple

import numpy as np
import scipy.constants as _k
from scipy.interpolate import UnivariateSpline

TABLE = Path(__file__).resolve().parent / "data" / "chantler.npz"

_AVOGADRO = _k.Avogadro
_HC_EV_ANGSTROM = 1.0e10 * _k.Planck * _k.c / _k.e
_R_E_CM = 100.0 * _k.physical_constants["classical electron radius"][0]

# element symbols chemparser.py accepts (its ELEMENTS list)
_SYMBOLS = frozenset((
    "Ac Ag Al Am Ar As At Au B Ba Be Bi Bk Br C Ca Cd Ce Cf Cl Cm Co Cr Cs Cu Dy Er Es Eu F Fe Fm "
    "Fr Ga Gd Ge H He Hf Hg Ho I In Ir K Kr La Li Lr Lu Md Mg Mn Mo N Na Nb Nd Ne Ni No Np O Os P "
    "Pa Pb Pd Pm Po Pr Pt Pu Ra Rb Re Rh Rn Ru S Sb Sc Se Si Sm Sn Sr Ta Tb Tc Te Th Ti Tl Tm U "
    "Unh Unp Unq Uns V W Xe Y Yb Zn Zr").split())

_TOKEN = re.compile(r"(?P<name>[A-Z][a-z]*)|(?P<num>[0-9]+\.?[0-9]*(?:[eE][-+]?[0-9]+)?)|(?P<open>\()|(?P<close>\))")


def composition(formula: str) -> Dict[str, float]:
    """Atoms per formula unit, e.g. 'Mn(SO4)2(H2O)7' -> {Mn: 1, S: 2, O: 15, H: 14}.

    A count multiplies the element or parenthesised group right before it.
    Raises ValueError on an unknown symbol, a count with nothing to multiply,
    unbalanced parentheses or an empty formula."""
    groups = [{}]          # open groups, innermost last
    operand = None         # element or closed group waiting for its count
    pos = 0

    def settle(count: float) -> None:
        nonlocal operand
        if operand is not None:
            for sym, n in operand.items():
                groups[-1][sym] = groups[-1].get(sym, 0) + n * count
            operand = None

    while pos < len(formula):
        m = _TOKEN.match(formula, pos)
        if m is None:
            raise ValueError(f"unrecognized element or number in {formula!r} at {pos}")
        pos = m.end()
        if m.group("num") is not None:
            if operand is None:
                raise ValueError(f"count without an element or group in {formula!r}")
            settle(float(m.group("num")))
            continue
        settle(1)
        if m.group("name") is not None:
            if m.group("name") not in _SYMBOLS:
                raise ValueError(f"{m.group('name')!r} is not an element symbol")
            operand = {m.group("name"): 1}
        elif m.group("open") is not None:
            groups.append({})
        else:
            if len(groups) == 1:
                raise ValueError(f"unbalanced ')' in {formula!r}")
            operand = groups.pop()
    settle(1)
    if len(groups) != 1:
        raise ValueError(f"expected right paren in {formula!r}")
    if not groups[0]:
        raise ValueError("empty formula")
    return groups[0]


@lru_cache(maxsize=1)
def _tables():
    with np.load(TABLE, allow_pickle=False) as z:
        meta = {str(s): (int(zz), float(mm)) for s, zz, mm in zip(z["symbols"], z["Z"], z["molar_mass"])}
        cols = {k: z[k] for k in z.files if "." in k}
    return meta, cols


def element(symbol: str) -> Tuple[int, float]:
    """(atomic number, molar mass in g/mol) of an element."""
    meta, _ = _tables()
    if symbol not in meta:
        raise ValueError(f"no Chantler table for {symbol!r}")
    return meta[symbol]


def chantler(symbol: str, energy: float, column: str) -> float:
    """Chantler `column` (f1, f2, mu_photo or mu_total) of an element at
    `energy` eV, interpolated from the tabulated points around it."""
    element(symbol)
    _, cols = _tables()
    te = cols[f"{symbol}.energy"]
    ty = cols[f"{symbol}.{column}"]
    below = np.flatnonzero(te <= energy)
    if below.size == 0:
        raise ValueError(f"{energy} eV is below the {symbol} table")
    last = int(below[-1])
    lo, hi = max(0, last - 5), min(len(te), last + 6)
    x, y = te[lo:hi], ty[lo:hi]
    e = np.array([float(energy)])
    if column == "f1":
        return float(UnivariateSpline(x, y, s=1)(e)[0])
    return float(np.exp(np.interp(np.log(e), np.log(x), np.log(y)))[0])


def xray_delta_beta(material: str, density: float, energy: float) -> Tuple[float, float]:
    """(delta, beta) of `material` (a formula) at `density` g/cm^3 and
    `energy` eV, beta from the total cross-section."""
    wl_cm = 1.0e-8 * _HC_EV_ANGSTROM / float(energy)
    mass = delta = beta = 0.0
    for sym, n in composition(material).items():
        z, molar = element(sym)
        w = float(density) * n * _AVOGADRO
        delta += w * (chantler(sym, energy, "f1") + z)
        beta += w * chantler(sym, energy, "f2") * (chantler(sym, energy, "mu_total") /
                                                   chantler(sym, energy, "mu_photo"))
        mass += n * molar
    s = wl_cm * wl_cm * _R_E_CM / (2 * np.pi * mass)
    return delta * s, beta * s
