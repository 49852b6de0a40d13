"""The C-ABI library loads and exports every symbol declared in include/*.h;
error semantics that need no GPU."""
import re
import subprocess
from pathlib import Path

import numpy as np

from conftest import ROOT
from curcuma_amd import GFNFFMethod, Mol
from curcuma_amd._lib import LIB_PATH, load_library


def declared_symbols():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        for m in re.finditer(r"^\s*[A-Za-z_][\w\s\*]*?\b(gfnff_[a-z_0-9]+)\s*\(", h.read_text(), re.M):
            names.add(m.group(1))
    return names


def test_exports_all_declared_symbols():
    load_library()
    out = subprocess.run(["nm", "-D", "--defined-only", str(LIB_PATH)], capture_output=True, text=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if line.strip()}
    missing = declared_symbols() - exported
    assert not missing, missing
    assert len(declared_symbols()) >= 15


def test_errors_without_molecule():
    m = GFNFFMethod()
    e = m.calculateEnergy(True)
    assert e == 0.0 and m.hasError() and "no molecule" in m.getErrorMessage()


def test_bad_element_is_reported():
    m = GFNFFMethod()
    assert not m.setMolecule(Mol(np.array([1, 120]), np.zeros((2, 3))))
    assert m.hasError()


def test_nan_geometry_is_reported():
    m = GFNFFMethod()
    x = np.array([[0.0, 0.0, 0.0], [0.0, 0.0, np.nan]])
    assert not m.setMolecule(Mol(np.array([1, 1]), x))
    assert "non-finite" in m.getErrorMessage()


def test_unsupported_systems_are_refused_loudly():
    """A charged system of exactly two bonded fragments (the reference
    distributes the charge by an extra fragment EEQ that is not built here) is
    refused with a message instead of returning wrong numbers."""
    from curcuma_amd.workloads import water_cluster
    Z, x = water_cluster(2, seed=1)
    m = GFNFFMethod()
    assert not m.setMolecule(Mol(Z, x, charge=1))
    assert "two-fragment" in m.getErrorMessage()


def test_pi_system_topology():
    """Benzene: one pi system over the six carbons, hydrogens outside it."""
    from conftest import BOHR, load_fixture
    fx = load_fixture("C6H6")
    Z = np.array(fx["Z"])
    m = GFNFFMethod()
    assert m.setMolecule(Mol(Z, np.array(fx["xyz_bohr"]) * BOHR)), m.getErrorMessage()
    hyb, pi, bonds = m.topology()
    assert set(pi[Z == 6]) == {1} and set(pi[Z == 1]) == {0}
    assert set(hyb[Z == 6]) == {2}
    assert m.params_view().ninv == 6


def test_topology_of_water_and_alkane():
    from curcuma_amd.workloads import alkane, water_cluster
    Z, x = water_cluster(8, seed=1)
    m = GFNFFMethod()
    assert m.setMolecule(Mol(Z, x))
    hyb, pi, bonds = m.topology()
    assert len(bonds) == 16 and pi.max() == 0
    assert set(hyb[Z == 8]) == {3} and set(hyb[Z == 1]) == {0}
    v = m.params_view()
    assert v.nfrag == 8
    Z, x = alkane(10)
    assert m.setMolecule(Mol(Z, x))
    hyb, pi, bonds = m.topology()
    assert len(bonds) == len(Z) - 1 and m.params_view().nfrag == 1
