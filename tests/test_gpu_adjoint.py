"""Parameter adjoint (force matching) against central finite differences.

Loss L(theta) = sum |g(theta) - g_ref|^2 with g_ref evaluated at per-element
scales perturbed by 2 % (seeded); dL/dtheta from the adjoint kernel at
theta = 1 is compared with central differences (h = 1e-5) of the FP64 loss.
Also checks the torch.autograd.Function wrapper gives the same gradient.
"""
import numpy as np
import pytest

from conftest import gpu_available
from curcuma_amd import GFNFFMethod, Mol
from curcuma_amd.workloads import alkane, water_cluster

pytestmark = pytest.mark.gpu

if not gpu_available():
    pytest.skip("no HIP device", allow_module_level=True)


def grad_at(m, theta):
    m.setParamScales(theta)
    m.calculateEnergy(True)
    assert not m.hasError(), m.getErrorMessage()
    return m.getGradient().copy()


@pytest.mark.parametrize("case", ["alkane30", "water27", "complex"])
def test_adjoint_matches_finite_differences(case):
    """complex (H, C, N, O; 231 atoms) holds case-4 hydrogen bonds, whose
    lone-pair distance must not follow the fitted repz scales."""
    if case == "alkane30":
        Z, xyz = alkane(30, sigma=0.05, seed=5)
    elif case == "complex":
        from conftest import BOHR, load_fixture
        fx = load_fixture("complex")
        Z, xyz = np.array(fx["Z"]), np.array(fx["xyz_bohr"]) * BOHR
    else:
        Z, xyz = water_cluster(27, seed=7)
    m = GFNFFMethod()
    assert m.setMolecule(Mol(Z, xyz)), m.getErrorMessage()
    npar = m.paramCount()
    rng = np.random.default_rng(6)
    theta_ref = 1.0 + 0.02 * rng.normal(size=npar)
    g_ref = grad_at(m, theta_ref)
    theta = np.ones(npar)
    g = grad_at(m, theta)
    w = 2.0 * (g - g_ref)
    d_adj = m.paramAdjoint(w)

    def loss(t):
        return float(((grad_at(m, t) - g_ref) ** 2).sum())

    h = 1e-5
    d_fd = np.zeros(npar)
    for k in range(npar):
        tp, tm = theta.copy(), theta.copy()
        tp[k] += h
        tm[k] -= h
        d_fd[k] = (loss(tp) - loss(tm)) / (2 * h)
    scale = np.abs(d_fd).max()
    err = np.abs(d_adj - d_fd).max() / scale
    print(f"{case}: L={loss(theta):.6e} dL/dtheta adjoint={d_adj} fd={d_fd} max rel err={err:.2e}")
    assert err < 1e-5


def test_torch_function_gradient():
    torch = pytest.importorskip("torch")
    from curcuma_amd.fitting import force_matching_loss
    Z, xyz = alkane(20, sigma=0.05, seed=9)
    m = GFNFFMethod()
    assert m.setMolecule(Mol(Z, xyz)), m.getErrorMessage()
    npar = m.paramCount()
    theta_ref = 1.0 + 0.02 * np.random.default_rng(3).normal(size=npar)
    g_ref = grad_at(m, theta_ref)
    theta = torch.ones(npar, dtype=torch.float64, requires_grad=True)
    L = force_matching_loss(theta, [m], [g_ref])
    L.backward()
    g = grad_at(m, np.ones(npar))
    d_direct = m.paramAdjoint(2.0 * (g - g_ref))
    assert np.allclose(theta.grad.numpy(), d_direct, rtol=1e-12, atol=0)
