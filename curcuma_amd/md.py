"""Device-resident molecular dynamics over the engine (gfnff_md_steps).

The reference's SimpleMD calls the energy interface once per step and keeps
coordinates, velocities and gradient on the host (simplemd.cpp:2580-2673,
3866-3890).  Here they stay in HBM: each call of ``MDRun.run`` queues a chunk
of velocity-Verlet steps (one device-resident evaluation each, plus a few
update kernels) and records Epot, Ekin, T and Etot per step on the device.
Between chunks the driver reads the status word; when the molecule has moved
past the hydrogen/halogen-bond re-detection threshold (sqrt(sum dx^2)/N >
0.3 Bohr since the last detection, gfnff_method.cpp:2243-2251) it reads the
coordinates back once, lets ``updateGeometry`` re-detect the lists (the
engine is rebuilt when they changed) and continues.

Units: coordinates in Bohr, velocities in Bohr per atomic time unit, masses
in amu, time step in fs, temperature in K.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import NTERMS, fptr

# standard atomic weights (amu) of the elements the workloads use
ATOMIC_MASS = {1: 1.00794, 5: 10.811, 6: 12.011, 7: 14.00674, 8: 15.9994, 9: 18.9984, 14: 28.0855, 15: 30.97376,
               16: 32.066, 17: 35.4527, 35: 79.904, 53: 126.90447}
BOHR_ANGSTROM = 0.529177210903
KB_EH = 3.166811e-6          # Hartree per Kelvin
AMU_ME = 1822.8884850        # electron masses per amu
FLAG_HB_STALE = 8


def masses_of(Z) -> np.ndarray:
    try:
        return np.array([ATOMIC_MASS[int(z)] for z in Z], dtype=np.float64)
    except KeyError as e:
        raise ValueError(f"no mass for element {e}; pass masses explicitly") from None


def maxwell_boltzmann(masses_amu, T, seed=0) -> np.ndarray:
    """Velocities (Bohr per atomic time unit) drawn at temperature T, centre-of-mass motion removed."""
    rng = np.random.default_rng(seed)
    m = np.asarray(masses_amu, dtype=np.float64) * AMU_ME
    v = rng.normal(size=(len(m), 3)) * np.sqrt(KB_EH * T / m)[:, None]
    v -= (m[:, None] * v).sum(0) / m.sum()
    return v


class MDRun:
    """Velocity-Verlet MD of one handle's molecule with device-resident state.

    torch is used only for device memory and the stream (plumbing)."""

    def __init__(self, method, xyz_bohr, vel, masses_amu, dt_fs, T0=0.0, tau_fs=100.0, device="cuda:0"):
        import torch
        self.torch = torch
        self.m = method
        self.n = len(masses_amu)
        self.masses = np.ascontiguousarray(masses_amu, dtype=np.float64)
        self.dt_fs, self.T0, self.tau_fs = float(dt_fs), float(T0), float(tau_fs)
        dev = torch.device(device)
        self.x = torch.tensor(np.asarray(xyz_bohr, dtype=np.float64).reshape(-1, 3), device=dev).contiguous()
        self.v = torch.tensor(np.asarray(vel, dtype=np.float64).reshape(-1, 3), device=dev).contiguous()
        self.g = torch.zeros((self.n, 3), dtype=torch.float64, device=dev)
        self.e = torch.zeros(NTERMS + 1, dtype=torch.float64, device=dev)
        self.stream = torch.cuda.Stream()
        self.redetections = 0
        self.m.energyGradientDevice(self.x.data_ptr(), self.e.data_ptr(), self.g.data_ptr(), self.stream.cuda_stream)

    def run(self, nsteps: int, chunk: int = 50) -> np.ndarray:
        """nsteps steps in chunks; returns the (nsteps, 4) record Epot, Ekin, T, Etot."""
        torch = self.torch
        rec = torch.zeros((max(nsteps, 1), 4), dtype=torch.float64, device=self.x.device)
        done = 0
        lib = self.m._lib
        while done < nsteps:
            k = min(chunk, nsteps - done)
            st = lib.gfnff_md_steps(self.m.handle, C.c_void_p(self.x.data_ptr()), C.c_void_p(self.v.data_ptr()),
                                    C.c_void_p(self.g.data_ptr()), fptr(self.masses), self.dt_fs, self.T0,
                                    self.tau_fs, k, C.c_void_p(rec[done].data_ptr()),
                                    C.c_void_p(self.stream.cuda_stream))
            if st != 0:
                raise RuntimeError(self.m.getErrorMessage())
            done += k
            flags = self.m.deviceStatus(clear=True)  # synchronises; once per chunk
            if flags & ~FLAG_HB_STALE:
                raise RuntimeError(f"device evaluation failed during MD (flags {flags})")
            if flags & FLAG_HB_STALE:
                self._redetect()
        torch.cuda.synchronize()
        return rec[:nsteps].cpu().numpy()

    def _redetect(self):
        """Re-detects the hydrogen/halogen-bond lists at the current geometry
        (updateGeometry, gfnff_method.cpp:1597-1690); the gradient is refreshed
        when the engine was rebuilt with new lists."""
        self.torch.cuda.synchronize()
        xyz_ang = self.x.cpu().numpy() * BOHR_ANGSTROM
        if not self.m.updateGeometry(xyz_ang):
            raise RuntimeError(self.m.getErrorMessage())
        self.redetections += 1
        self.m.energyGradientDevice(self.x.data_ptr(), self.e.data_ptr(), self.g.data_ptr(), self.stream.cuda_stream)

    def positions(self) -> np.ndarray:
        self.torch.cuda.synchronize()
        return self.x.cpu().numpy()

    def velocities(self) -> np.ndarray:
        self.torch.cuda.synchronize()
        return self.v.cpu().numpy()
