"""Extract the Chantler form-factor rows for the elements used by the golden
end-to-end case into a small JSON fixture.

Reads big-wave's material database (`nist_lookup/nist_lookup/xrayref.db`, an
SQLite data file) with the standard-library sqlite3 module; no reference code is
imported or executed.  The fixture is what the oracle's delta/beta restatement
(`oracle/xray.py`) reads, so neither the oracle nor any test needs the reference
tree at run time.

Usage:  python tests/golden/make_chantler_fixture.py
"""
import json
import sqlite3
from pathlib import Path

DB = "/root/reference/nist_lookup/nist_lookup/xrayref.db"
ELEMENTS = ("Si", "Au")
OUT = Path(__file__).with_name("chantler_si_au.json")


def main() -> None:
    con = sqlite3.connect(f"file:{DB}?mode=ro", uri=True)
    out = {}
    for sym in ELEMENTS:
        z, mass = con.execute(
            "select atomic_number, molar_mass from elements where element=?", (sym,)
        ).fetchone()
        cols = con.execute(
            "select energy, f1, f2, mu_photo, mu_total from Chantler where element=?",
            (sym,),
        ).fetchone()
        names = ("energy", "f1", "f2", "mu_photo", "mu_total")
        out[sym] = {"Z": int(z), "molar_mass": float(mass)}
        for name, txt in zip(names, cols):
            out[sym][name] = json.loads(txt)
    OUT.write_text(json.dumps(out, indent=0))
    print(f"wrote {OUT}")


if __name__ == "__main__":
    main()
