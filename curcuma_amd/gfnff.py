"""Python mirror of the reference's ``ComputationalMethod`` interface for GFN-FF.

The reference's plug-in boundary is ``ComputationalMethod``
(src/core/energy_calculators/computational_method.h:43-360) with the GFN-FF
adapter ``GFNFFComputationalMethod`` (qm_methods/gfnff_method.{h,cpp}).  This
class keeps the same method names, argument meaning and error behaviour
(errors are reported through ``hasError()`` / ``getErrorMessage()``, nothing
raises out of ``calculateEnergy``), and forwards to the C-ABI of the HIP
engine.  It never falls back to a CPU implementation: without a GPU,
``calculateEnergy`` reports a device error.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from ._lib import NTERMS, TERM_NAMES, as_f64, as_i32, fptr, iptr

GFNFF_E = {name: k for k, name in enumerate(TERM_NAMES)}

ANGSTROM_TO_BOHR = 1.0 / 0.529177210903  # src/core/units.h:81-82


@dataclass
class Mol:
    """Subset of the reference's ``Mol`` (src/core/global.h:69-99)."""

    atoms: np.ndarray                      # Z, shape (N,)
    geometry: np.ndarray                   # Angstrom, shape (N, 3), row-major
    charge: int = 0
    bonds: list = field(default_factory=list)   # forced bonds (0-based pairs)

    @property
    def number_atoms(self) -> int:
        return int(len(self.atoms))


class GFNFFMethod:
    """``-method gfnff`` drop-in: same Mol in, energy and N x 3 gradient out."""

    def __init__(self, parameters: dict | None = None, precision: str = "fp64"):
        self._lib = _lib.load_library()
        self._h = self._lib.gfnff_create()
        if not self._h:
            raise MemoryError("gfnff_create failed")
        self._n = 0
        self._energy = 0.0
        self._grad = np.zeros((0, 3))
        self._terms = np.zeros(NTERMS)
        self._error = ""
        if parameters:
            self.setParameters(parameters)
        self.setPrecision(precision)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            self._lib.gfnff_destroy(h)
            self._h = None

    # --- ComputationalMethod interface -------------------------------------------------
    def setMolecule(self, mol: Mol) -> bool:
        Z = as_i32(mol.atoms)
        xyz = as_f64(np.asarray(mol.geometry).reshape(-1, 3))
        fb = as_i32(np.asarray(mol.bonds, dtype=np.int32).reshape(-1, 2)) if len(mol.bonds) else as_i32(np.zeros((0, 2)))
        st = self._lib.gfnff_set_molecule(self._h, len(Z), iptr(Z), fptr(xyz), int(mol.charge),
                                          iptr(fb) if len(fb) else None, len(fb))
        self._n = len(Z) if st == 0 else 0
        self._grad = np.zeros((self._n, 3))
        return st == 0

    def updateGeometry(self, geometry_angstrom) -> bool:
        xyz = as_f64(np.asarray(geometry_angstrom).reshape(-1, 3))
        if xyz.shape[0] != self._n:
            self._error = "updateGeometry: atom count mismatch"
            return False
        return self._lib.gfnff_update_geometry(self._h, fptr(xyz)) == 0

    def calculateEnergy(self, gradient: bool = True) -> float:
        e = C.c_double(0.0)
        g = np.zeros((self._n, 3))
        t = np.zeros(NTERMS)
        st = self._lib.gfnff_energy_gradient(self._h, 1 if gradient else 0, C.byref(e), fptr(g), fptr(t))
        self._energy = float(e.value) if st in (0, 6) else 0.0
        self._grad = g
        self._terms = t
        return self._energy

    def getGradient(self) -> np.ndarray:
        return self._grad.copy()

    def copyGradientTo(self, out: np.ndarray) -> None:
        out[...] = self._grad

    def getCharges(self) -> np.ndarray:
        q = np.zeros(self._n)
        self._lib.gfnff_charges(self._h, fptr(q))
        return q

    def getCN(self) -> np.ndarray:
        cn = np.zeros(self._n)
        self._lib.gfnff_cn(self._h, fptr(cn))
        return cn

    def getEnergyDecomposition(self) -> dict:
        """The reference's ten keys (qm_methods/gfnff_method.cpp:152-185).  As
        there, Torsion holds the primary and extra torsions
        (FFEnergyComponents::dihedral, ff_workspace_gfnff.cpp:348,449); the
        repulsion (getRepulsionEnergy) and the triple-bond torsions
        (getSTorsionEnergy) are not among the keys."""
        t = self._terms
        return {"Bond": float(t[GFNFF_E["bond"]]), "Angle": float(t[GFNFF_E["angle"]]),
                "Torsion": float(t[GFNFF_E["torsion"]]), "Inversion": float(t[GFNFF_E["inversion"]]),
                "Dispersion": float(t[GFNFF_E["dispersion"]]), "Coulomb": float(t[GFNFF_E["coulomb"]]),
                "HBond": float(t[GFNFF_E["hbond"]]), "XBond": float(t[GFNFF_E["xbond"]]),
                "ATM": float(t[GFNFF_E["atm"]]), "BATM": float(t[GFNFF_E["batm"]])}

    # per-term getters (computational_method.h:295-337, GFNFF::*Energy gfnff_method.cpp:10496-10575)
    def getBondEnergy(self) -> float:
        return float(self._terms[GFNFF_E["bond"]])

    def getAngleEnergy(self) -> float:
        return float(self._terms[GFNFF_E["angle"]])

    def getDihedralEnergy(self) -> float:
        return float(self._terms[GFNFF_E["torsion"]])

    def getInversionEnergy(self) -> float:
        return float(self._terms[GFNFF_E["inversion"]])

    def getRepulsionEnergy(self) -> float:
        """Bonded plus non-bonded repulsion (GFNFF::RepulsionEnergy, gfnff_method.cpp:10521-10525)."""
        return float(self._terms[GFNFF_E["bonded_rep"]] + self._terms[GFNFF_E["nonbonded_rep"]])

    def getDispersionEnergy(self) -> float:
        return float(self._terms[GFNFF_E["dispersion"]])

    def getCoulombEnergy(self) -> float:
        return float(self._terms[GFNFF_E["coulomb"]])

    def getSTorsionEnergy(self) -> float:
        """Triple-bond torsions (FFEnergyComponents::stors, ff_workspace_gfnff.cpp:633)."""
        return float(self._terms[GFNFF_E["stors"]])

    def getTerms(self) -> dict:
        """All per-term energies of the last evaluation, by engine term name."""
        return {name: float(v) for name, v in zip(TERM_NAMES, self._terms) if not name.startswith("reserved")}

    def energyGradientDevice(self, xyz_bohr_dev: int, energy_dev: int, grad_dev: int | None, stream: int | None) -> None:
        """Device-resident evaluation (gfnff_energy_gradient_device); pointers and
        the hipStream_t as integers, queued without synchronising."""
        st = self._lib.gfnff_energy_gradient_device(self._h, C.c_void_p(xyz_bohr_dev), C.c_void_p(energy_dev),
                                                    C.c_void_p(grad_dev) if grad_dev else None,
                                                    C.c_void_p(stream) if stream else None)
        if st != 0:
            raise RuntimeError(self.getErrorMessage())

    def captureGraph(self, xyz_bohr_dev: int, energy_dev: int, grad_dev: int | None, stream: int) -> None:
        """Records one device-resident evaluation with these buffers into a HIP
        graph (gfnff_graph_capture); `stream` must be a created stream."""
        st = self._lib.gfnff_graph_capture(self._h, C.c_void_p(xyz_bohr_dev), C.c_void_p(energy_dev),
                                           C.c_void_p(grad_dev) if grad_dev else None, C.c_void_p(stream))
        if st != 0:
            raise RuntimeError(self.getErrorMessage())

    def launchGraph(self, stream: int | None) -> None:
        """Replays the captured evaluation (gfnff_graph_launch) on `stream`."""
        st = self._lib.gfnff_graph_launch(self._h, C.c_void_p(stream) if stream else None)
        if st != 0:
            raise RuntimeError(self.getErrorMessage())

    def deviceStatus(self, clear: bool = True) -> int:
        """GFNFF_FLAG_* bits raised by device-resident evaluations since the last clearing read."""
        f = C.c_int32(0)
        st = self._lib.gfnff_device_status(self._h, C.byref(f), 1 if clear else 0)
        if st != 0:
            raise RuntimeError(self.getErrorMessage())
        return int(f.value)

    def hasError(self) -> bool:
        return bool(self._lib.gfnff_has_error(self._h))

    def getErrorMessage(self) -> str:
        return self._lib.gfnff_last_error(self._h).decode()

    def setParameters(self, parameters: dict) -> None:
        """Engine keys (numbers) and the reference's "solvent" / "solvent_model"
        (gfnff_method.cpp:811-835; GFN-FF always uses ALPB, so solvent_model
        "gbsa" or "alpb" select the same model)."""
        for k, v in parameters.items():
            if k == "solvent":
                self.setSolvent(v)
            elif k == "solvent_model":
                if str(v).lower() not in ("alpb", "gbsa", "gb", "2", "2.0"):
                    raise ValueError(f"unknown solvent_model {v!r}")
            else:
                self._lib.gfnff_set_param(self._h, k.encode(), float(v))

    def setSolvent(self, solvent: str | None) -> None:
        """ALPB implicit solvation (one of the reference's GFN-FF solvents, or
        None / "none" for the gas phase); raises on an unknown name, after
        which the method is in the gas phase, as the reference continues."""
        st = self._lib.gfnff_set_solvent(self._h, (solvent or "").encode())
        if st != 0:
            raise ValueError(self.getErrorMessage())

    def solvationEnergy(self) -> dict:
        """gborn, ghb, gsasa, gshift of the last evaluation (Hartree)."""
        p = np.zeros(4)
        st = self._lib.gfnff_solvation_energy(self._h, fptr(p))
        if st != 0:
            raise RuntimeError(self.getErrorMessage())
        return {"gborn": float(p[0]), "ghb": float(p[1]), "gsasa": float(p[2]), "gshift": float(p[3])}

    def solverIterations(self) -> int:
        """Iterations of the last iterative EEQ solve (0 on the Cholesky path)."""
        return int(self._lib.gfnff_solver_iterations(self._h))

    def setPrecision(self, precision: str) -> None:
        mode = {"fp64": 0, "mixed": 1}[precision]
        self._lib.gfnff_set_precision(self._h, mode)

    def isThreadSafe(self) -> bool:
        return False

    # --- engine extensions ------------------------------------------------------------
    def params_view(self) -> _lib.ParamsView:
        v = _lib.ParamsView()
        st = self._lib.gfnff_get_params(self._h, C.byref(v))
        if st != 0:
            raise RuntimeError("no molecule set")
        return v

    def solvation_view(self) -> _lib.AlpbView:
        v = _lib.AlpbView()
        st = self._lib.gfnff_get_solvation(self._h, C.byref(v))
        if st != 0:
            raise RuntimeError("no molecule set")
        return v

    def topology(self):
        hyb = np.zeros(self._n, dtype=np.int32)
        pi = np.zeros(self._n, dtype=np.int32)
        nb = C.c_int32(0)
        self._lib.gfnff_get_topology(self._h, iptr(hyb), iptr(pi), None, 0, C.byref(nb))
        bonds = np.zeros((nb.value, 2), dtype=np.int32)
        self._lib.gfnff_get_topology(self._h, None, None, iptr(bonds), nb.value, C.byref(nb))
        return hyb, pi, bonds

    def paramCount(self) -> int:
        """Number of fitted per-element parameter scales (3 * number of elements)."""
        return int(self._lib.gfnff_param_count(self._h))

    def setParamScales(self, theta) -> None:
        """Per-element scales of repz, the repulsion exponent and sqrt(Z r4r2)
        (see include/gfnff_mi355x.h); all 1 after setMolecule."""
        t = np.ascontiguousarray(theta, dtype=np.float64)
        st = self._lib.gfnff_param_scales(self._h, fptr(t), len(t))
        if st != 0:
            raise RuntimeError(self.getErrorMessage())

    def paramAdjoint(self, w) -> np.ndarray:
        """sum_i w_i . d g_i / d theta at the last evaluation (w: natoms x 3)."""
        wv = np.ascontiguousarray(w, dtype=np.float64).reshape(-1)
        if wv.size != 3 * self._n:
            raise ValueError("w must have natoms x 3 entries")
        out = np.zeros(self.paramCount())
        st = self._lib.gfnff_param_adjoint(self._h, fptr(wv), fptr(out), len(out))
        if st != 0:
            raise RuntimeError(self.getErrorMessage())
        return out

    @property
    def handle(self):
        return self._h

    @property
    def natoms(self) -> int:
        return self._n


class GFNFFBatch(GFNFFMethod):
    """Batch of independent molecules evaluated together (one launch per kernel
    for the whole batch).  ``setBatch`` replaces ``setMolecule``;
    ``calculateEnergy`` returns the sum over the batch and ``getEnergies`` the
    per-molecule energies; ``getGradient`` covers all atoms."""

    def setBatch(self, atom_off, Z, geometry_angstrom, charges=None) -> bool:
        off = as_i32(atom_off)
        z = as_i32(Z)
        x = as_f64(geometry_angstrom).reshape(-1, 3)
        ch = as_i32(charges) if charges is not None else None
        self._nmol = len(off) - 1
        self._off = off
        st = self._lib.gfnff_set_batch(self._h, self._nmol, iptr(off), iptr(z), fptr(x),
                                       iptr(ch) if ch is not None else None)
        self._n = len(z) if st == 0 else 0
        self._grad = np.zeros((self._n, 3))
        return st == 0

    def getEnergies(self) -> np.ndarray:
        e = np.zeros(self._nmol)
        st = self._lib.gfnff_batch_energies(self._h, fptr(e))
        if st != 0:
            raise RuntimeError(self.getErrorMessage())
        return e

    @property
    def nmol(self) -> int:
        return self._nmol
