"""Batched evaluation of independent molecules against one-at-a-time
evaluation through the single-molecule path and against the CPU oracle.

Every molecule of the batch must reproduce its single-molecule energy and
gradient (the batch only changes the launch decomposition; FP64 sums are
reordered, so agreement is to rounding: 1e-9 Eh, 1e-9 Eh/Bohr)."""
import numpy as np
import pytest

from conftest import BOHR, gpu_available
from curcuma_amd import GFNFFBatch, GFNFFMethod, Mol
from curcuma_amd.workloads import alkane_batch, mixed_batch, water_cluster

pytestmark = pytest.mark.gpu

if not gpu_available():
    pytest.skip("no HIP device", allow_module_level=True)


def single(Z, xyz, precision="fp64"):
    m = GFNFFMethod(precision=precision)
    assert m.setMolecule(Mol(Z, xyz)), m.getErrorMessage()
    e = m.calculateEnergy(True)
    assert not m.hasError(), m.getErrorMessage()
    return e, m.getGradient(), m


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_alkane_batch_matches_single_molecules(precision, oracle):
    off, Z, xyz = alkane_batch(nmol=12, nmin=20, nmax=400, seed=31)
    b = GFNFFBatch(precision=precision)
    assert b.setBatch(off, Z, xyz), b.getErrorMessage()
    etot = b.calculateEnergy(True)
    assert not b.hasError(), b.getErrorMessage()
    em = b.getEnergies()
    g = b.getGradient()
    assert abs(etot - em.sum()) <= 1e-9 * max(1.0, abs(etot))
    tol = 1e-9 if precision == "fp64" else 1e-6
    for k in range(len(off) - 1):
        a0, a1 = off[k], off[k + 1]
        e1, g1, m1 = single(Z[a0:a1], xyz[a0:a1], precision)
        de = abs(em[k] - e1)
        dg = float(np.abs(g[a0:a1] - g1).max())
        print(f"{precision} mol {k}: N={a1 - a0} E_batch={em[k]:.10f} E_single={e1:.10f} dE={de:.2e} dg={dg:.2e}")
        assert de <= tol * max(1.0, abs(e1) / 10)
        assert dg <= tol
        if k == 0 and precision == "fp64":
            ref = oracle.eval(m1.params_view(), np.asarray(xyz[a0:a1]) * (1.0 / BOHR))
            assert abs(em[k] - ref["E"]) <= 1e-9 * max(1.0, abs(ref["E"]))


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_alkane_peptide_batch(precision, oracle):
    """Alkanes and glycine oligopeptides (pi systems, inversions, hydrogen-bond
    cases 2/3) in one batch, each molecule against its single evaluation."""
    off, Z, xyz = mixed_batch(nmol=10, nmin=20, nmax=200, seed=17)
    b = GFNFFBatch(precision=precision)
    assert b.setBatch(off, Z, xyz), b.getErrorMessage()
    b.calculateEnergy(True)
    assert not b.hasError(), b.getErrorMessage()
    em, g = b.getEnergies(), b.getGradient()
    tol = 1e-9 if precision == "fp64" else 1e-6
    npep = 0
    for k in range(len(off) - 1):
        a0, a1 = off[k], off[k + 1]
        e1, g1, m1 = single(Z[a0:a1], xyz[a0:a1], precision)
        npep += int(m1.params_view().ninv > 0)
        assert abs(em[k] - e1) <= tol * max(1.0, abs(e1) / 10), k
        assert np.abs(g[a0:a1] - g1).max() <= tol, k
        if precision == "fp64" and m1.params_view().ninv > 0:
            ref = oracle.eval(m1.params_view(), np.asarray(xyz[a0:a1]) * (1.0 / BOHR))
            assert abs(em[k] - ref["E"]) <= 1e-9 * max(1.0, abs(ref["E"]))
    assert npep > 0


def test_mixed_chemistry_batch():
    """Waters and alkanes in one batch (different element sets per molecule)."""
    Zw, xw = water_cluster(8, seed=4)
    off1, Za, xa = alkane_batch(nmol=3, nmin=20, nmax=80, seed=8)
    Z = np.concatenate([Zw, Za])
    xyz = np.vstack([xw, xa])
    off = np.concatenate([[0, len(Zw)], off1[1:] + len(Zw)])
    b = GFNFFBatch()
    assert b.setBatch(off, Z, xyz), b.getErrorMessage()
    b.calculateEnergy(True)
    assert not b.hasError(), b.getErrorMessage()
    em = b.getEnergies()
    for k in range(len(off) - 1):
        e1, g1, _ = single(Z[off[k]:off[k + 1]], xyz[off[k]:off[k + 1]])
        assert abs(em[k] - e1) <= 1e-9 * max(1.0, abs(e1))
        assert np.abs(b.getGradient()[off[k]:off[k + 1]] - g1).max() <= 1e-9
