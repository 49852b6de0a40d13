"""Convert the reference's Fortran-derived GFN-FF fixtures into compact JSON.

Source: test_cases/reference_data/*.ref.json of the reference tree (generated
there by the Fortran GFN-FF "gfnff-analyzer-parser"; consumed by the
reference's test_cases/test_gfnff_validation.cpp).  Only data is carried over:
element, Bohr coordinates, per-term energies, per-atom per-term gradients,
EEQ charges, CN/hybridisation/topology charges and the bond/angle/torsion
parameter tables.  Run once in a container that has the reference tree:

    python tests/golden/make_fixtures.py /root/reference
"""
from __future__ import annotations

import json
import sys
from pathlib import Path

SYMBOLS = ["X", "H", "He", "Li", "Be", "B", "C", "N", "O", "F", "Ne", "Na", "Mg", "Al", "Si", "P", "S", "Cl",
           "Ar", "K", "Ca", "Sc", "Ti", "V", "Cr", "Mn", "Fe", "Co", "Ni", "Cu", "Zn", "Ga", "Ge", "As", "Se",
           "Br", "Kr"]

NAMES = ["HH", "OH", "HCl", "H2O", "water", "HCN", "O3", "CH4", "methane", "CH3OH", "CH3OCH3", "C6H6",
         "acetic_acid_dimer", "caffeine", "triose", "complex", "polymer"]


def convert(src: Path) -> dict:
    d = json.loads(src.read_text())
    atoms = d["molecule"]["atoms"]
    n = len(atoms)
    out = {
        "name": d["molecule"]["name"],
        "source": "test_cases/reference_data/" + src.name,
        "Z": [SYMBOLS.index(a["element"]) for a in atoms],
        "xyz_bohr": [[a["x"], a["y"], a["z"]] for a in atoms],
        "energy_components": d["energy_components"],
        "energy_charges": d.get("energy_charges"),
        "gradient_norm": d.get("gradients", {}).get("norm"),
    }
    topo = d.get("topology", {})
    out["topology"] = {k: topo.get(k) for k in ("coordination_numbers", "hybridizations", "charges")}
    comps = {}
    gd = d.get("gradient_decomposition", {})
    for i in range(n):
        for comp, v in gd.get(str(i), {}).items():
            comps.setdefault(comp, [[0.0, 0.0, 0.0] for _ in range(n)])[i] = [v["x"], v["y"], v["z"]]
    out["gradient_components"] = comps
    out["bonds"] = [[b["atoms"][0], b["atoms"][1], b["r0"], b["kbond"], b["prefactor"]] for b in d.get("bonds", [])]
    out["angles"] = [[a["atoms"][0], a["atoms"][1], a["atoms"][2], a["phi0"], a["fc"]] for a in d.get("angles", [])]
    out["torsions"] = [[t["atoms"][0], t["atoms"][1], t["atoms"][2], t["atoms"][3], t["n"], t["phi0"], t["V"]]
                       for t in d.get("torsions", [])]
    return out


def main(ref_root: str) -> None:
    src_dir = Path(ref_root) / "test_cases" / "reference_data"
    dst = Path(__file__).resolve().parent / "fixtures"
    dst.mkdir(exist_ok=True)
    for name in NAMES:
        out = convert(src_dir / f"{name}.ref.json")
        (dst / f"{name}.json").write_text(json.dumps(out, separators=(",", ":")))
        print(name, len(out["Z"]))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
