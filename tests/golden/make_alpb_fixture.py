"""Collect the reference's GFN-FF + ALPB solvation check data into one JSON file.

Source (data only): test_cases/sqm_reference/reference_data/gfnff_alpb_xtb.ref.json
(total energies of xtb 6.7.1 --gfnff --alpb SOLVENT, gsolv state) and the
geometries test_cases/sqm_reference/molecules/<name>.xyz that the reference's
test_cases/sqm_reference/validate_gfnff_solvation.py feeds to curcuma
(tolerance 1e-7 Eh, test_cases/sqm_reference/CMakeLists.txt:332-356).
Run once in a container that has the reference tree:

    python tests/golden/make_alpb_fixture.py /root/reference
"""
from __future__ import annotations

import json
import sys
from pathlib import Path

SYMBOLS = ["X", "H", "He", "Li", "Be", "B", "C", "N", "O", "F", "Ne"]
MOLS = ["H2O", "NH3", "CH4", "HCN", "C6H6", "acetic_acid_dimer", "caffeine"]
SOLVENTS = ["water", "dmso", "acetone", "chloroform"]


def main(ref_root: str) -> None:
    base = Path(ref_root) / "test_cases" / "sqm_reference"
    ref = json.loads((base / "reference_data" / "gfnff_alpb_xtb.ref.json").read_text())
    out = {"source": "test_cases/sqm_reference/reference_data/gfnff_alpb_xtb.ref.json",
           "generator": ref.get("_generator"), "tolerance_Eh": 1e-7, "molecules": {}, "energies": {}}
    for m in MOLS:
        lines = (base / "molecules" / f"{m}.xyz").read_text().splitlines()
        n = int(lines[0].split()[0])
        Z, xyz = [], []
        for ln in lines[2:2 + n]:
            f = ln.split()
            Z.append(SYMBOLS.index(f[0]))
            xyz.append([float(v) for v in f[1:4]])
        out["molecules"][m] = {"Z": Z, "xyz_ang": xyz}
        for s in SOLVENTS:
            out["energies"][f"{m}_{s}"] = ref["data"][f"{m}_{s}"]
    dst = Path(__file__).parent / "fixtures" / "alpb_xtb.json"
    dst.write_text(json.dumps(out, indent=1))
    print(dst, len(out["energies"]))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
