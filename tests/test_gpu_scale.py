"""The benchmark configurations at full size on the GPU.

* 80 parallel C80H162 chains (19,360 atoms, 80 fragments) on the mixed path
  with the FP32 MFMA Cholesky of the 19,280-variable reduced EEQ system and
  FP64 iterative refinement: the constrained EEQ residual of the refined
  charges, recomputed by the CPU oracle in FP64, is below 1e-10 relative;
  every chain's charge sums to zero; energy per atom and gradient RMS against
  the FP64 path (PCG to 1e-13) within the accuracy bar; a 4-chain bundle of
  the same geometry recipe against the CPU oracle on both solvers.
* The force-matching parameter adjoint over a batch of 256 conformers
  (60-200 atoms, alkanes and glycine oligopeptides, sigma 0.05 A, seed 5)
  in one call, against central finite differences (h = 1e-5) of the FP64
  loss sum_conformers |g(theta) - g_ref|^2, g_ref at per-element scales
  perturbed by 2 % (seed 6).
* The 3000-H2O water cluster (9000 atoms, 3000 fragments) on the FP64 path:
  charges by the constrained EEQ residual, the rest against the threaded
  CPU oracle given those charges.
"""
import os

import numpy as np
import pytest

from conftest import BOHR, gpu_available
from curcuma_amd import GFNFFBatch, GFNFFMethod, Mol
from curcuma_amd.workloads import alkane_bundle, mixed_batch, water_cluster

pytestmark = pytest.mark.gpu

if not gpu_available():
    pytest.skip("no HIP device", allow_module_level=True)

NTHR = max(1, min(len(os.sched_getaffinity(0)), int(os.environ.get("OMP_NUM_THREADS", "16") or 16)))


def evaluate(Z, xyz, precision, solver):
    m = GFNFFMethod(precision=precision)
    m.setParameters({"eeq_solver": solver})
    assert m.setMolecule(Mol(Z, xyz)), m.getErrorMessage()
    e = m.calculateEnergy(True)
    assert not m.hasError(), m.getErrorMessage()
    return m, e


def test_bundle_mixed_cholesky_with_refinement(oracle):
    Z, xyz = alkane_bundle(80, 80, seed=4)
    n = len(Z)
    assert n == 19360
    x = xyz * (1.0 / BOHR)
    print(f"bundle: {n} atoms, FP64 PCG", flush=True)
    m64, e64 = evaluate(Z, xyz, "fp64", 0)
    print(f"bundle: mixed, FP32 MFMA Cholesky + FP64 refinement", flush=True)
    mmx, emx = evaluate(Z, xyz, "mixed", 1)
    v = mmx.params_view()
    assert v.nfrag == 80
    q, q64 = mmx.getCharges(), m64.getCharges()
    rel, qerr = oracle.eeq_residual(v, x, mmx.getCN(), q, NTHR)
    rel64, _ = oracle.eeq_residual(m64.params_view(), x, m64.getCN(), q64, NTHR)
    g, g64 = mmx.getGradient(), m64.getGradient()
    de = abs(emx - e64) / n
    grms = float(np.sqrt(np.mean((g - g64) ** 2)))
    dq = float(np.abs(q - q64).max())
    frag = np.ctypeslib.as_array(v.frag, shape=(n,))
    fsum = np.abs(np.bincount(frag, weights=q)).max()
    print(f"bundle: E64={e64:.10f} Emix={emx:.10f} |dE|/N={de:.2e} gradRMS={grms:.2e} max|dq|={dq:.2e} "
          f"residual mixed={rel:.2e} fp64={rel64:.2e} max|sum q_chain|={fsum:.2e}")
    assert rel < 1e-10
    assert rel64 < 1e-12
    assert qerr < 1e-10 and fsum < 1e-10
    assert de <= 1e-6
    assert grms <= 1e-5
    assert dq <= 1e-8


@pytest.mark.parametrize("precision,solver", [("fp64", 0), ("fp64", 1), ("mixed", 1)])
def test_bundle_subset_against_oracle(precision, solver, oracle):
    Z, xyz = alkane_bundle(4, 80, seed=4)
    n = len(Z)
    m, e = evaluate(Z, xyz, precision, solver)
    ref = oracle.eval_mt(m.params_view(), xyz * (1.0 / BOHR), NTHR)
    de = abs(e - ref["E"]) / n
    gerr = float(np.abs(m.getGradient() - ref["grad"]).max())
    print(f"bundle4 {precision}/{solver}: N={n} E={e:.12f} oracle={ref['E']:.12f} |dE|/N={de:.2e} max|dg|={gerr:.2e}")
    if precision == "fp64":
        assert abs(e - ref["E"]) <= 1e-9 * max(1.0, abs(ref["E"]))
        assert gerr <= 1e-9
    else:
        assert de <= 1e-6 and float(np.sqrt(np.mean((m.getGradient() - ref["grad"]) ** 2))) <= 1e-5


def test_force_matching_adjoint_256_conformers():
    # sizes drawn in [66, 200] and rounded down to the family (3n+2 alkanes,
    # 7n+3 peptides) stay within 60..200 atoms
    off, Z, xyz = mixed_batch(256, 66, 200, sigma=0.05, seed=5)
    sizes = np.diff(off)
    assert sizes.min() >= 60 and sizes.max() <= 200
    b = GFNFFBatch(precision="fp64")
    assert b.setBatch(off, Z, xyz), b.getErrorMessage()
    npar = b.paramCount()
    print(f"adjoint batch: {len(off) - 1} conformers, {len(Z)} atoms, {npar} parameters", flush=True)

    def grad_at(theta):
        b.setParamScales(theta)
        b.calculateEnergy(True)
        assert not b.hasError(), b.getErrorMessage()
        return b.getGradient().copy()

    theta_ref = 1.0 + 0.02 * np.random.default_rng(6).normal(size=npar)
    g_ref = grad_at(theta_ref)
    theta = np.ones(npar)
    g = grad_at(theta)
    d_adj = b.paramAdjoint(2.0 * (g - g_ref))

    def loss(t):
        return float(((grad_at(t) - g_ref) ** 2).sum())

    h = 1e-5
    d_fd = np.zeros(npar)
    for k in range(npar):
        tp, tm = theta.copy(), theta.copy()
        tp[k] += h
        tm[k] -= h
        d_fd[k] = (loss(tp) - loss(tm)) / (2 * h)
    scale = np.abs(d_fd).max()
    err = np.abs(d_adj - d_fd).max() / scale
    print(f"adjoint batch: L={loss(theta):.6e} adjoint={d_adj} fd={d_fd} max rel err={err:.2e}")
    assert err < 1e-5


def test_water_3000_fp64_against_threaded_oracle(oracle):
    """The EEQ charges are pinned by the constrained residual (O(N^2) on the
    CPU, independent of any solver); every other part of the evaluation by
    the threaded oracle given those charges (its O(N^3) dense solve of the
    12,000-row bordered system is left out).  The Coulomb energy is
    stationary in q at the solution, so the charge error enters the energy
    at second order."""
    Z, xyz = water_cluster(3000, seed=2)
    n = len(Z)
    print(f"water_3000: {n} atoms, FP64 PCG", flush=True)
    m, e = evaluate(Z, xyz, "fp64", 0)
    x = xyz * (1.0 / BOHR)
    v = m.params_view()
    q = m.getCharges()
    rel, qerr = oracle.eeq_residual(v, x, m.getCN(), q, NTHR)
    print(f"water_3000: EEQ residual {rel:.2e}, fragment charge sums {qerr:.2e}; threaded oracle with {NTHR} "
          f"threads", flush=True)
    ref = oracle.eval_mt(v, x, NTHR, q_in=q)
    gerr = float(np.abs(m.getGradient() - ref["grad"]).max())
    cerr = float(np.abs(m.getCN() - ref["cn"]).max())
    print(f"water_3000: E={e:.10f} oracle={ref['E']:.10f} dE={abs(e - ref['E']):.2e} max|dg|={gerr:.2e} "
          f"max|dCN|={cerr:.2e}")
    assert rel < 1e-12 and qerr < 1e-12
    assert abs(e - ref["E"]) <= 1e-9 * abs(ref["E"])
    for k in range(len(m._terms)):
        assert abs(m._terms[k] - ref["terms"][k]) <= 1e-9 * max(1.0, abs(ref["terms"][k])), k
    assert gerr <= 1e-9
    assert cerr <= 1e-12
