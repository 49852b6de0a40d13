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


@pytest.mark.parametrize("batch_fp32", [0, 1])
@pytest.mark.parametrize("batch_solver", [0, 1, 2, 3, 4, 5, 6])
def test_mixed_batch_charges_refined(batch_solver, batch_fp32):
    """With batch_fp32 = 1 the mixed path factors each molecule's reduced EEQ
    system in FP32 on MFMA and refines in FP64 against the FP64 matrix
    (batch_solver 2, 4 and 5 always factor in FP64): its charges equal the FP64
    path's to 1e-8 e, whichever per-molecule solve form runs."""
    off, Z, xyz = mixed_batch(nmol=16, nmin=20, nmax=400, seed=23)
    q = {}
    for prec in ("fp64", "mixed"):
        b = GFNFFBatch(precision=prec, parameters={"batch_solver": batch_solver, "batch_fp32": batch_fp32})
        assert b.setBatch(off, Z, xyz), b.getErrorMessage()
        b.calculateEnergy(True)
        assert not b.hasError(), b.getErrorMessage()
        q[prec] = b.getCharges()
    dq = float(np.abs(q["mixed"] - q["fp64"]).max())
    print(f"batch_solver {batch_solver} batch_fp32 {batch_fp32}: max|q_mixed - q_fp64| = {dq:.2e}")
    assert dq <= 1e-8


def test_batch_solvers_agree(oracle):
    """Every per-molecule solve form (right-looking in two register variants,
    the whole-workgroup tile steps, the stepwise batch factorisation and the
    left-looking factorisation with the matrix formed in place, per molecule
    and by column blocks over the batch) gives the
    same FP64 charges, energies and gradient to rounding, including a
    two-fragment molecule (direct four-term entries) and padded systems of
    every size class up to 448; molecule 0 against the oracle."""
    off, Z, xyz = mixed_batch(nmol=14, nmin=20, nmax=400, seed=41)
    Zw, xw = water_cluster(2, seed=9)  # two fragments in one molecule
    Z = np.concatenate([Z, Zw])
    xyz = np.vstack([xyz, xw + np.array([200.0, 0.0, 0.0])])
    off = np.concatenate([off, [len(Z)]])
    res = {}
    for bs in (0, 1, 2, 3, 4, 5, 6):
        b = GFNFFBatch(precision="fp64", parameters={"batch_solver": bs})
        assert b.setBatch(off, Z, xyz), b.getErrorMessage()
        b.calculateEnergy(True)
        assert not b.hasError(), b.getErrorMessage()
        res[bs] = (b.getCharges(), b.getEnergies(), b.getGradient())
        if bs == 5:
            a0, a1 = off[0], off[1]
            _, _, m1 = single(Z[a0:a1], xyz[a0:a1])
            ref = oracle.eval(m1.params_view(), np.asarray(xyz[a0:a1]) * (1.0 / BOHR))
            assert abs(res[5][1][0] - ref["E"]) <= 1e-9 * max(1.0, abs(ref["E"]))
            assert np.abs(res[5][0][a0:a1] - ref["q"]).max() <= 1e-10
    q0, e0, g0 = res[0]
    for bs in (1, 2, 3, 4, 5, 6):
        q, e, g = res[bs]
        dq, de, dg = np.abs(q - q0).max(), np.abs(e - e0).max(), np.abs(g - g0).max()
        print(f"batch_solver {bs}: max|dq|={dq:.2e} max|dE|={de:.2e} max|dg|={dg:.2e}")
        assert dq <= 1e-11 and de <= 1e-9 and dg <= 1e-10


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_pair_split_agrees(precision):
    """pair_split = 1 evaluates dispersion and repulsion before the EEQ solve
    and the Coulomb term after it, into a second set of tile partials: the
    same energies and gradient to rounding, for a batch and for a single
    molecule.  On the mixed path each pass rounds its own FP32 pair force
    before the FP64 sums, so the gradient agrees to FP32 rounding (1e-6)."""
    off, Z, xyz = mixed_batch(nmol=8, nmin=20, nmax=300, seed=5)
    res = {}
    for split in (0, 1):
        b = GFNFFBatch(precision=precision, parameters={"pair_split": split})
        assert b.setBatch(off, Z, xyz), b.getErrorMessage()
        b.calculateEnergy(True)
        assert not b.hasError(), b.getErrorMessage()
        res[split] = (b.getEnergies(), b.getGradient())
    de = float(np.abs(res[1][0] - res[0][0]).max())
    dg = float(np.abs(res[1][1] - res[0][1]).max())
    gtol = 1e-11 if precision == "fp64" else 1e-6
    print(f"{precision} batch: max|dE|={de:.2e} max|dg|={dg:.2e}")
    assert de <= 1e-10 and dg <= gtol
    Zs, xs = water_cluster(64, seed=3)
    out = {}
    for split in (0, 1):
        m = GFNFFMethod(precision=precision)
        m.setParameters({"pair_split": split})
        assert m.setMolecule(Mol(Zs, xs)), m.getErrorMessage()
        out[split] = (m.calculateEnergy(True), m.getGradient())
        assert not m.hasError(), m.getErrorMessage()
    assert abs(out[1][0] - out[0][0]) <= 1e-10 and np.abs(out[1][1] - out[0][1]).max() <= gtol


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_batch_streams_identical(precision):
    """batch_streams 2 (the factorisation in two halves on two streams) runs
    the same arithmetic per molecule: charges, energies and gradient equal bit
    for bit to the one-stream run (eager launches; graph capture is refused
    for it)."""
    off, Z, xyz = mixed_batch(nmol=31, nmin=20, nmax=400, seed=5)
    res = {}
    for bs in (1, 2):
        b = GFNFFBatch(precision=precision, parameters={"batch_streams": bs})
        assert b.setBatch(off, Z, xyz), b.getErrorMessage()
        b.calculateEnergy(True)
        assert not b.hasError(), b.getErrorMessage()
        res[bs] = (b.getCharges(), b.getEnergies(), b.getGradient())
    for k in range(3):
        assert np.array_equal(res[1][k], res[2][k]), k


@pytest.mark.skip(reason="graph capture with batch_streams 2 crashed inside the HIP runtime (host segfault in "
                         "gfnff_graph_capture); the cause is not established, so the engine refuses that capture")
def test_batch_streams_graph():
    """Graph replays of the two-stream factorisation equal its eager run."""
    import torch
    off, Z, xyz = mixed_batch(nmol=31, nmin=20, nmax=400, seed=5)
    b = GFNFFBatch(precision="mixed", parameters={"batch_streams": 2})
    assert b.setBatch(off, Z, xyz), b.getErrorMessage()
    b.calculateEnergy(True)
    g0 = b.getGradient()
    x = torch.tensor(np.asarray(xyz) / BOHR, dtype=torch.float64, device="cuda").reshape(-1)
    out = torch.zeros(16, dtype=torch.float64, device="cuda")
    g = torch.zeros_like(x)
    st = torch.cuda.Stream()
    b.captureGraph(x.data_ptr(), out.data_ptr(), g.data_ptr(), st.cuda_stream)
    for _ in range(3):
        b.launchGraph(st.cuda_stream)
        st.synchronize()
        assert np.array_equal(g.cpu().numpy().reshape(-1, 3), g0)
        torch.cuda.synchronize()


def test_batch_streams_capture_refused():
    """The engine refuses the capture that crashed, with a message."""
    import torch
    off, Z, xyz = mixed_batch(nmol=4, nmin=20, nmax=100, seed=5)
    b = GFNFFBatch(precision="mixed", parameters={"batch_streams": 2})
    assert b.setBatch(off, Z, xyz), b.getErrorMessage()
    x = torch.tensor(np.asarray(xyz) / BOHR, dtype=torch.float64, device="cuda").reshape(-1)
    out = torch.zeros(16, dtype=torch.float64, device="cuda")
    g = torch.zeros_like(x)
    st = torch.cuda.Stream()
    with pytest.raises(RuntimeError, match="batch_streams 2"):
        b.captureGraph(x.data_ptr(), out.data_ptr(), g.data_ptr(), st.cuda_stream)
