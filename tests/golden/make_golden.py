"""Writes the golden fixtures of the parity tests from the reference's data.

Run in the build container (it reads /root/reference, which the GPU host
does not have); the outputs are committed:

* ``structures.npz``: sphere sets (centres in milli-angstrom as int32, radius
  as float32 value) of structures whose reference meshes are recorded in the
  reference's validation campaign. Selection as the campaign's PQR inputs:
  first model, alternate location of highest occupancy, no water, no
  hydrogen; Bondi radii
  (reference radii/tables.rs:25-41) read back as f32 as the PQR path does
  (reference radii/mod.rs:193-196).
* ``expected.json``: per (entry, spacing) the reference's recorded vertex,
  triangle, component and cavity counts, area (all printed digits) and volume
  of the rows of campaigns ``acc`` (host-c) and ``acc110`` (x86 host-a) of
  docs/validation/compare/runs.csv, plus the SHA-256 of the reference's Marching Cubes
  table (crates/sesmesh/src/mesh/mc_table.rs), computed here as data.
"""
import csv
import gzip
import hashlib
import json
import re
from pathlib import Path

import numpy as np

REF = Path("/root/reference")
OUT = Path(__file__).resolve().parent

# Bondi radii of the elements present in these entries (radii/tables.rs).
BONDI = {"C": 1.70, "N": 1.55, "O": 1.52, "S": 1.80, "P": 1.80, "MG": 1.73, "K": 2.75, "NA": 2.27,
         "CL": 1.75, "ZN": 1.39, "CA": 2.31, "SE": 1.90}

ENTRIES = {
    "1CRN": REF / "tests/data/fixtures/1CRN.pdb.gz",
    "1EJG": REF / "tests/data/fixtures/1EJG.pdb.gz",
    "1UBQ": REF / "tests/data/fixtures/1UBQ.pdb.gz",
    "3FX2": REF / "tests/data/fixtures/3FX2.pdb.gz",
    "2PTN": REF / "tests/data/fixtures/2PTN.pdb.gz",
    "1AON": REF / "data/1aon.pdb",
}
WATER = {"HOH", "WAT", "DOD", "H2O"}


def read_pdb(path):
    """Atoms of the first model. Alternate locations: per residue the letter
    of highest total occupancy, the first seen on ties (reference
    structure/select.rs:298-345); then no water and no hydrogen."""
    op = gzip.open if str(path).endswith(".gz") else open
    atoms = []
    with op(path, "rt") as f:
        for line in f:
            if line.startswith("ENDMDL"):
                break
            if line.startswith(("ATOM", "HETATM")):
                atoms.append(line)
    occ = {}
    for line in atoms:
        if line[16] != " ":
            key = (line[21], line[22:26], line[26])
            letters = occ.setdefault(key, [])
            for e in letters:
                if e[0] == line[16]:
                    e[1] += float(line[54:60])
                    break
            else:
                letters.append([line[16], float(line[54:60])])
    chosen = {}
    for key, letters in occ.items():
        best = None
        for c, o in letters:
            if best is None or not best[1] >= o:
                best = (c, o)
        chosen[key] = best[0]
    rows = []
    for line in atoms:
        if line[16] != " " and chosen[(line[21], line[22:26], line[26])] != line[16]:
            continue
        if line[17:20].strip() in WATER:
            continue
        el = line[76:78].strip().upper()
        if el in ("H", "D"):
            continue
        x, y, z = (float(line[c:c + 8]) for c in (30, 38, 46))
        rows.append((round(x * 1000), round(y * 1000), round(z * 1000), BONDI[el]))
    return rows


def mc_table_sha():
    txt = (REF / "crates/sesmesh/src/mesh/mc_table.rs").read_text()
    off = txt.split("TRI_OFFSET")[1].split("];")[0].split("= [")[1]
    offsets = [int(x) for x in re.findall(r"\d+", off)]
    body = txt.split("TRI_EDGES")[1].split("= [", 1)[1]
    body = "\n".join(l.split("//")[0] for l in body.splitlines())
    edges = [int(x) for x in re.findall(r"\d+", body)]
    data = np.array(offsets, dtype="<u2").tobytes() + bytes(edges)
    return hashlib.sha256(data).hexdigest(), len(edges) // 3


def main():
    arrays = {}
    for name, path in ENTRIES.items():
        rows = read_pdb(path)
        a = np.array(rows, dtype=np.float64)
        arrays[name + "_xyz"] = a[:, :3].astype(np.int32)
        arrays[name + "_r"] = a[:, 3].astype(np.float32)
    np.savez_compressed(OUT / "structures.npz", **arrays)

    expected = {"runs": []}
    with open(REF / "docs/validation/compare/runs.csv") as f:
        for r in csv.DictReader(f):
            if r["tool"] != "sesmesh" or r["entry"] not in ENTRIES:
                continue
            if (r["campaign"], r["host"]) not in (("acc", "host-c"), ("acc110", "host-a")):
                continue
            if float(r["spacing"]) not in (0.5, 0.25):
                continue
            expected["runs"].append({
                "entry": r["entry"], "atoms": int(r["atoms"]), "spacing": float(r["spacing"]),
                "vertices": int(r["vertices"]), "triangles": int(r["triangles"]),
                "components": int(r["components"]), "cavities": int(r["cavities"]),
                "area": float(r["reported_area"]), "volume": float(r["volume"]),
                "campaign": r["campaign"], "host": r["host"],
                "source": "docs/validation/compare/runs.csv"})
    sha, n = mc_table_sha()
    expected["mc_table"] = {"sha256": sha, "triangles": n,
                            "layout": "TRI_OFFSET[257] as little-endian u16, then TRI_EDGES as u8 triples",
                            "source": "crates/sesmesh/src/mesh/mc_table.rs"}
    (OUT / "expected.json").write_text(json.dumps(expected, indent=1) + "\n")
    for name in ENTRIES:
        print(name, len(arrays[name + "_r"]))


if __name__ == "__main__":
    main()
