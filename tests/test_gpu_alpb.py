"""ALPB solvation on the GPU against the CPU oracle (same parameter set, same
Bohr coordinates), through the C-ABI, and against the reference's xtb
energies.  FP64 kernels on both precision paths.

Bars: energy and every term |dE| <= 1e-9 Eh, charges 1e-9 e, gradient
components 1e-9 Eh/Bohr against the oracle (FP64 path); the xtb fixture's
own 1e-7 Eh.  The mixed path's pair terms are FP32, so its bar is the
north star's (1e-6 Eh/atom, gradient RMS 1e-5 Eh/Bohr)."""
import json

import numpy as np
import pytest

from conftest import BOHR, FIXTURES, gpu_available
from curcuma_amd import GFNFFMethod, Mol
from curcuma_amd.workloads import water_cluster
from test_alpb import ammonium

pytestmark = pytest.mark.gpu

if not gpu_available():
    pytest.skip("no HIP device", allow_module_level=True)

ALPB = json.loads((FIXTURES / "alpb_xtb.json").read_text())


def run(Z, xyz_ang, solvent, oracle, charge=0, params=None, precision="fp64"):
    p = {"solvent": solvent}
    p.update(params or {})
    m = GFNFFMethod(p, precision=precision)
    assert m.setMolecule(Mol(np.asarray(Z), np.asarray(xyz_ang), charge)), m.getErrorMessage()
    e = m.calculateEnergy(True)
    assert not m.hasError(), m.getErrorMessage()
    ref = oracle.eval_alpb(m.params_view(), m.solvation_view(), np.asarray(xyz_ang) / BOHR, nthreads=8)
    return m, e, ref


def check(m, e, ref, etol=1e-9, gtol=1e-9, qtol=1e-9):
    dg = np.abs(m.getGradient() - ref["grad"]).max()
    dq = np.abs(m.getCharges() - ref["q"]).max()
    parts = m.solvationEnergy()
    dp = max(abs(parts[k] - ref["parts"][i]) for i, k in enumerate(("gborn", "ghb", "gsasa", "gshift")))
    print(f"E_gpu={e:.12f} E_cpu={ref['E']:.12f} dE={e - ref['E']:.2e} dg={dg:.2e} dq={dq:.2e} dparts={dp:.2e}",
          flush=True)
    assert abs(e - ref["E"]) <= etol
    assert np.abs(m._terms - ref["terms"]).max() <= etol
    assert dp <= etol
    assert dg <= gtol
    assert dq <= qtol


@pytest.mark.parametrize("case", sorted(ALPB["energies"]))
def test_fixture_cases(case, oracle):
    mol, solvent = case.rsplit("_", 1)
    d = ALPB["molecules"][mol]
    print(case, flush=True)
    m, e, ref = run(d["Z"], d["xyz_ang"], solvent, oracle)
    check(m, e, ref)
    assert abs(e - ALPB["energies"][case]) <= ALPB["tolerance_Eh"]


@pytest.mark.parametrize("solver", [0, 1])
def test_solvers_and_charged(solver, oracle):
    """PCG (packed tiles) and Cholesky (row-major) storage of the reduced
    matrix both take the Born part; a cation adds it to the right-hand side
    and makes the shape term non-zero."""
    Z, xyz = ammonium()
    m, e, ref = run(Z, xyz, "water", oracle, charge=1, params={"eeq_solver": solver})
    check(m, e, ref)
    assert abs(m.getCharges().sum() - 1.0) <= 1e-10


@pytest.mark.parametrize("nmol", [40, 200])
def test_water_clusters(nmol, oracle):
    """120 and 600 atoms: several 64-blocks of the reduced matrix, surface
    neighbour lists of tens of atoms, two fragments per cluster pair."""
    Z, xyz = water_cluster(nmol, seed=3)
    print(f"water cluster {nmol}", flush=True)
    m, e, ref = run(Z, xyz, "dmso", oracle)
    check(m, e, ref, etol=1e-9 * max(1.0, len(Z) / 100), gtol=1e-9)


def test_mixed_precision(oracle):
    d = ALPB["molecules"]["caffeine"]
    m, e, ref = run(d["Z"], d["xyz_ang"], "water", oracle, precision="mixed")
    n = len(d["Z"])
    assert abs(e - ref["E"]) / n <= 1e-6
    assert np.sqrt(np.mean((m.getGradient() - ref["grad"]) ** 2)) <= 1e-5


def test_device_path_and_repeat(oracle):
    """Device-resident evaluation twice and a geometry change: no state
    leaks between evaluations (Born radii, areas and lists are rebuilt)."""
    import torch
    d = ALPB["molecules"]["acetic_acid_dimer"]
    m, e, ref = run(d["Z"], d["xyz_ang"], "acetone", oracle)
    x = torch.tensor(np.asarray(d["xyz_ang"]) / BOHR, dtype=torch.float64, device="cuda").reshape(-1)
    out = torch.zeros(16, dtype=torch.float64, device="cuda")
    g = torch.zeros_like(x)
    s = torch.cuda.current_stream()
    for _ in range(2):
        m.energyGradientDevice(x.data_ptr(), out.data_ptr(), g.data_ptr(), s.cuda_stream)
        torch.cuda.synchronize()
        assert abs(out[0].item() - ref["E"]) <= 1e-9
        assert np.abs(g.cpu().numpy().reshape(-1, 3) - ref["grad"]).max() <= 1e-9
    x2 = np.asarray(d["xyz_ang"]) + 0.01 * np.sin(np.arange(3 * len(d["Z"]))).reshape(-1, 3)
    assert m.updateGeometry(x2)
    e2 = m.calculateEnergy(True)
    ref2 = oracle.eval_alpb(m.params_view(), m.solvation_view(), x2 / BOHR)
    assert abs(e2 - ref2["E"]) <= 1e-9
    assert np.abs(m.getGradient() - ref2["grad"]).max() <= 1e-9


def test_gas_phase_unchanged(oracle):
    """Switching the solvent off restores the gas-phase evaluation exactly."""
    d = ALPB["molecules"]["caffeine"]
    m = GFNFFMethod()
    assert m.setMolecule(Mol(np.array(d["Z"]), np.array(d["xyz_ang"])))
    e0 = m.calculateEnergy(True)
    g0 = m.getGradient().copy()
    m.setSolvent("water")
    e1 = m.calculateEnergy(True)
    assert e1 < e0 and m._terms[13] != 0.0
    m.setSolvent("none")
    e2 = m.calculateEnergy(True)
    assert e2 == e0 and np.array_equal(m.getGradient(), g0) and m._terms[13] == 0.0
