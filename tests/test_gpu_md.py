"""Device-resident molecular dynamics (gfnff_md_steps, curcuma_amd.md.MDRun).

* Parity: the device velocity-Verlet trajectory against the same integrator
  on the host in numpy (the reference's SimpleMD::Verlet update,
  simplemd.cpp:2580-2673) driving the engine through the host-buffer entry
  point once per step.  Both evaluate the same kernels, so the trajectories
  agree to rounding after 20 steps.
* NVE energy conservation over 400 steps of 0.5 fs.
* The Berendsen bath (SimpleMD::Berendson, simplemd.cpp:3967-3975) pulls a
  600 K start towards 300 K.
* Hydrogen-bond re-detection between chunks once the displacement threshold
  (here lowered to 0.02 Bohr) is crossed, with the run continuing.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")  # before the engine library (one HIP runtime for both)

from conftest import BOHR, gpu_available  # noqa: E402
from curcuma_amd import GFNFFMethod, Mol  # noqa: E402
from curcuma_amd.md import AMU_ME, KB_EH, MDRun, masses_of, maxwell_boltzmann  # noqa: E402
from curcuma_amd.workloads import alkane, water_cluster  # noqa: E402

pytestmark = pytest.mark.gpu

if not gpu_available():
    pytest.skip("no HIP device", allow_module_level=True)

FS_AU = 41.34137314


def host_verlet(m, x0, v0, masses, dt_fs, nsteps):
    """Velocity Verlet on the host with the engine's host-buffer evaluation."""
    dt = dt_fs * FS_AU
    im = 1.0 / (np.repeat(masses, 3).reshape(-1, 3) * AMU_ME)
    x, v = x0.copy(), v0.copy()
    assert m.updateGeometry(x * BOHR)
    m.calculateEnergy(True)
    g = m.getGradient()
    out = []
    for _ in range(nsteps):
        a = g * im
        x = x + dt * v - 0.5 * a * dt * dt
        v = v - 0.5 * dt * a
        assert m.updateGeometry(x * BOHR)
        e = m.calculateEnergy(True)
        g = m.getGradient()
        v = v - 0.5 * dt * g * im
        ek = 0.5 * float((v * v / im).sum())
        out.append((e, ek))
    return x, v, np.array(out)


def test_md_matches_host_verlet():
    Z, xyz = alkane(20, sigma=0.02, seed=3)
    masses = masses_of(Z)
    x0 = xyz * (1.0 / BOHR)
    v0 = maxwell_boltzmann(masses, 300.0, seed=4)
    m_dev = GFNFFMethod()
    assert m_dev.setMolecule(Mol(Z, xyz)), m_dev.getErrorMessage()
    run = MDRun(m_dev, x0, v0, masses, dt_fs=0.5)
    rec = run.run(20, chunk=7)
    m_host = GFNFFMethod()
    assert m_host.setMolecule(Mol(Z, xyz)), m_host.getErrorMessage()
    xh, vh, hrec = host_verlet(m_host, x0, v0, masses, 0.5, 20)
    dx = float(np.abs(run.positions() - xh).max())
    dv = float(np.abs(run.velocities() - vh).max())
    de = float(np.abs(rec[:, 0] - hrec[:, 0]).max())
    dk = float(np.abs(rec[:, 1] - hrec[:, 1]).max())
    print(f"md parity: max|dx|={dx:.2e} Bohr max|dv|={dv:.2e} max|dEpot|={de:.2e} max|dEkin|={dk:.2e}")
    assert dx <= 1e-9 and dv <= 1e-12 and de <= 1e-9 and dk <= 1e-9
    dof = 3 * len(Z) - 6
    assert np.allclose(rec[:, 2], 2.0 * rec[:, 1] / (KB_EH * dof), rtol=1e-12)
    assert np.allclose(rec[:, 3], rec[:, 0] + rec[:, 1], rtol=0, atol=1e-12)


def test_md_nve_energy_conservation():
    Z, xyz = alkane(20, sigma=0.02, seed=3)
    masses = masses_of(Z)
    m = GFNFFMethod()
    assert m.setMolecule(Mol(Z, xyz)), m.getErrorMessage()
    run = MDRun(m, xyz * (1.0 / BOHR), maxwell_boltzmann(masses, 300.0, seed=5), masses, dt_fs=0.5)
    rec = run.run(400, chunk=100)
    drift = float(np.abs(rec[:, 3] - rec[0, 3]).max())
    print(f"nve: Etot0={rec[0, 3]:.8f} max|Etot-Etot0|={drift:.2e} Eh, <T>={rec[:, 2].mean():.1f} K")
    assert np.isfinite(rec).all()
    assert drift < 5e-4


def test_md_berendsen_bath():
    Z, xyz = alkane(20, sigma=0.02, seed=3)
    masses = masses_of(Z)
    m = GFNFFMethod()
    assert m.setMolecule(Mol(Z, xyz)), m.getErrorMessage()
    run = MDRun(m, xyz * (1.0 / BOHR), maxwell_boltzmann(masses, 900.0, seed=6), masses, dt_fs=0.5, T0=300.0,
                tau_fs=10.0)
    rec = run.run(600, chunk=200)
    late = rec[-200:, 2].mean()
    print(f"berendsen: T first={rec[0, 2]:.1f} K, mean of the last 200 steps={late:.1f} K")
    assert 200.0 < late < 400.0


def test_md_hb_redetection_between_chunks(oracle):
    Z, xyz = water_cluster(27, seed=5)
    masses = masses_of(Z)
    m = GFNFFMethod()
    m.setParameters({"hb_update_rmsd_bohr": 0.02})
    assert m.setMolecule(Mol(Z, xyz)), m.getErrorMessage()
    run = MDRun(m, xyz * (1.0 / BOHR), maxwell_boltzmann(masses, 300.0, seed=7), masses, dt_fs=0.5)
    rec = run.run(300, chunk=25)
    print(f"water_27: {run.redetections} re-detections, Etot drift {np.abs(rec[:, 3] - rec[0, 3]).max():.2e}")
    assert run.redetections >= 1
    assert np.isfinite(rec).all()
    # the final state evaluated on the device equals the oracle on the re-detected lists
    x = run.positions()
    ref = oracle.eval(m.params_view(), x)
    e = run.e.cpu().numpy()
    run.m.energyGradientDevice(run.x.data_ptr(), run.e.data_ptr(), run.g.data_ptr(), run.stream.cuda_stream)
    torch.cuda.synchronize()
    e = run.e.cpu().numpy()
    assert abs(e[0] - ref["E"]) <= 1e-9 * abs(ref["E"])
