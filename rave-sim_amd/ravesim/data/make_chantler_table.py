"""Write chantler.npz: the Chantler form-factor tables of all elements that
nist_lookup ships, in the arrays ravesim.xray reads.

Reads nist_lookup's material database (nist_lookup/nist_lookup/xrayref.db, an
SQLite data file) with the standard-library sqlite3 module; nothing of
nist_lookup is imported or executed.  For each element the table holds the
energies (eV) and the f1, f2, mu_photo and mu_total columns of its Chantler
row, parsed from their JSON text as Python floats, plus its atomic number and
molar mass from the `elements` table.

Usage:  python make_chantler_table.py [path/to/xrayref.db]
"""
import json
import sqlite3
import sys
from pathlib import Path

import numpy as np

DB = "/root/reference/nist_lookup/nist_lookup/xrayref.db"
OUT = Path(__file__).with_name("chantler.npz")
COLUMNS = ("energy", "f1", "f2", "mu_photo", "mu_total")


def main(db: str = DB) -> None:
    con = sqlite3.connect(f"file:{db}?mode=ro", uri=True)
    mass = {sym: (int(z), float(m)) for sym, z, m in
            con.execute("select element, atomic_number, molar_mass from elements")}
    arrays, symbols, zs, masses = {}, [], [], []
    for row in con.execute("select element, " + ", ".join(COLUMNS) + " from Chantler order by id"):
        sym = row[0]
        symbols.append(sym)
        zs.append(mass[sym][0])
        masses.append(mass[sym][1])
        for name, txt in zip(COLUMNS, row[1:]):
            arrays[f"{sym}.{name}"] = np.asarray(json.loads(txt), dtype=np.float64)
    np.savez_compressed(OUT, symbols=np.array(symbols), Z=np.array(zs, dtype=np.int64),
                        molar_mass=np.array(masses, dtype=np.float64), **arrays)
    print(f"wrote {OUT}: {len(symbols)} elements")


if __name__ == "__main__":
    main(*sys.argv[1:])
