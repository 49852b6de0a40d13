"""Mixed-precision path (FP32 pair terms with FP64 row sums; FP32 MFMA
factorisation of the reduced EEQ matrix with FP64 iterative refinement)
against the FP64 path and the CPU oracle on the same inputs.

Bar (the accuracy at which the mixed-precision throughput is quoted):
|dE| <= 1e-6 Eh per atom and gradient RMS <= 1e-5 Eh/Bohr versus the FP64
oracle; refined EEQ charges within 1e-8 e of the FP64 charges.
"""
import numpy as np
import pytest

from conftest import BOHR, gpu_available
from curcuma_amd import GFNFFMethod, Mol
from curcuma_amd.workloads import alkane, glycine_peptide, water_cluster

pytestmark = pytest.mark.gpu

if not gpu_available():
    pytest.skip("no HIP device", allow_module_level=True)


def evaluate(Z, xyz, precision, solver=0):
    m = GFNFFMethod(precision=precision)
    m.setParameters({"eeq_solver": solver})
    assert m.setMolecule(Mol(Z, xyz)), m.getErrorMessage()
    e = m.calculateEnergy(True)
    assert not m.hasError(), m.getErrorMessage()
    return m, e


@pytest.mark.parametrize("solver", [0, 1])
@pytest.mark.parametrize("case", ["water27", "alkane40", "water512", "gly20"])
def test_mixed_against_fp64_and_oracle(case, solver, oracle):
    if case == "water27":
        Z, xyz = water_cluster(27, seed=5)
    elif case == "alkane40":
        Z, xyz = alkane(40, sigma=0.02, seed=1)
    elif case == "gly20":
        Z, xyz = glycine_peptide(20, sigma=0.03, seed=6)
    else:
        Z, xyz = water_cluster(512, seed=3)
    n = len(Z)
    m64, e64 = evaluate(Z, xyz, "fp64", solver)
    m32, e32 = evaluate(Z, xyz, "mixed", solver)
    ref = oracle.eval(m64.params_view(), np.asarray(xyz) * (1.0 / BOHR))
    g64, g32 = m64.getGradient(), m32.getGradient()
    de_atom = abs(e32 - ref["E"]) / n
    grms = float(np.sqrt(np.mean((g32 - ref["grad"]) ** 2)))
    dq = float(np.abs(m32.getCharges() - m64.getCharges()).max())
    print(f"{case}: N={n} E64={e64:.10f} Emix={e32:.10f} |dE|/N={de_atom:.2e} gradRMS={grms:.2e} max|dq|={dq:.2e}")
    assert abs(e64 - ref["E"]) <= 1e-9 * max(1.0, abs(ref["E"]))
    assert de_atom <= 1e-6
    assert grms <= 1e-5
    assert dq <= 1e-8
    for k in range(len(m32._terms)):
        assert abs(m32._terms[k] - ref["terms"][k]) <= 1e-6 * n, k
