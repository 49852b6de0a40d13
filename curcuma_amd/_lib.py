"""ctypes binding of the engine's C-ABI (include/gfnff_mi355x.h)."""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "lib" / "libgfnff_mi355x.so"

NTERMS = 15
TERM_NAMES = ["bond", "angle", "torsion", "inversion", "dispersion", "bonded_rep", "nonbonded_rep", "coulomb",
              "hbond", "xbond", "atm", "batm", "stors", "solvation", "reserved14"]

STATUS = {0: "ok", 1: "args", 2: "unsupported", 3: "eeq", 4: "geometry", 5: "device", 6: "nan"}

_i32p = C.POINTER(C.c_int32)
_f64p = C.POINTER(C.c_double)


class ParamsView(C.Structure):
    """Mirror of gfnff_params_view (include/gfnff_params.h)."""

    _fields_ = [
        ("natoms", C.c_int32), ("nfrag", C.c_int32), ("nelem", C.c_int32), ("total_charge", C.c_int32),
        ("Z", _i32p), ("elem", _i32p), ("frag", _i32p),
        ("chi_base", _f64p), ("gam_eff", _f64p), ("alpeeq", _f64p), ("cnf", _f64p), ("rcov_cn", _f64p),
        ("rcov_d3", _f64p), ("hb_rad", _f64p), ("metal", _i32p), ("qa", _f64p), ("zeta", _f64p),
        ("sqrt_zr4r2", _f64p), ("rep_anb", _f64p), ("repz", _f64p), ("hb_lpd", _f64p),
        ("hb_bas", _f64p), ("hb_aci", _f64p), ("qfrag", _f64p),
        ("present_Z", _i32p), ("nref", _i32p), ("refcn", _f64p), ("c6ref", _f64p),
        ("topo_off", _i32p), ("topo_idx", _i32p), ("topo_dist", _i32p),
        ("nbond", C.c_int32), ("bond_ij", _i32p), ("bond_par", _f64p), ("bond_hbH", _i32p),
        ("nhbcn", C.c_int32), ("hbcn_pair", _i32p),
        ("nangle", C.c_int32), ("angle_ijk", _i32p), ("angle_par", _f64p),
        ("ntors", C.c_int32), ("tors_ijkl", _i32p), ("tors_par", _f64p),
        ("ninv", C.c_int32), ("inv_ijkl", _i32p), ("inv_par", _f64p),
        ("nstors", C.c_int32), ("stors_ijkl", _i32p), ("stors_par", _f64p),
        ("nbrep", C.c_int32), ("brep_ij", _i32p), ("brep_par", _f64p),
        ("nhb", C.c_int32), ("hb_ahb", _i32p), ("hb_case", _i32p), ("hb_par", _f64p),
        ("hb_nbB_off", _i32p), ("hb_nbB_idx", _i32p), ("hb_parentC", _i32p), ("hb_nbC_off", _i32p),
        ("hb_nbC_idx", _i32p),
        ("nxb", C.c_int32), ("xb_axb", _i32p), ("xb_par", _f64p),
        ("natm", C.c_int32), ("atm_ijk", _i32p), ("atm_scale", _f64p),
        ("nbatm", C.c_int32), ("batm_ijk", _i32p), ("batm_par", _f64p),
    ]


_LIB = None


class AlpbView(C.Structure):
    """Mirror of gfnff_alpb_view (include/gfnff_params.h)."""

    _fields_ = [("natoms", C.c_int32), ("hbond", C.c_int32), ("nang", C.c_int32), ("keps", C.c_double),
                ("alpbet", C.c_double), ("born_scale", C.c_double), ("gshift", C.c_double), ("lrcut", C.c_double),
                ("srcut", C.c_double)] + [(k, _f64p) for k in ("vdwr", "rho", "svdw", "vdwsa", "wrp", "trj_lo",
                                                                 "trj_hi", "gamsasa", "hbmag", "grid")]


def load_library(path: Path | None = None) -> C.CDLL:
    """Load the engine library; raises if it has not been built."""
    global _LIB
    if _LIB is not None:
        return _LIB
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise RuntimeError(f"engine library missing: {p} (run __graft_entry__.build())")
    lib = C.CDLL(str(p))
    lib.gfnff_create.restype = C.c_void_p
    lib.gfnff_destroy.argtypes = [C.c_void_p]
    lib.gfnff_set_molecule.argtypes = [C.c_void_p, C.c_int32, _i32p, _f64p, C.c_int32, _i32p, C.c_int32]
    lib.gfnff_update_geometry.argtypes = [C.c_void_p, _f64p]
    lib.gfnff_set_batch.argtypes = [C.c_void_p, C.c_int32, _i32p, _i32p, _f64p, _i32p]
    lib.gfnff_batch_energies.argtypes = [C.c_void_p, _f64p]
    lib.gfnff_batch_energies_device.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
    lib.gfnff_energy_gradient.argtypes = [C.c_void_p, C.c_int32, _f64p, _f64p, _f64p]
    lib.gfnff_energy_gradient_device.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    lib.gfnff_device_status.argtypes = [C.c_void_p, _i32p, C.c_int32]
    lib.gfnff_graph_capture.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    lib.gfnff_graph_launch.argtypes = [C.c_void_p, C.c_void_p]
    lib.gfnff_md_steps.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, _f64p, C.c_double, C.c_double,
                                   C.c_double, C.c_int32, C.c_void_p, C.c_void_p]
    lib.gfnff_charges.argtypes = [C.c_void_p, _f64p]
    lib.gfnff_cn.argtypes = [C.c_void_p, _f64p]
    lib.gfnff_set_param.argtypes = [C.c_void_p, C.c_char_p, C.c_double]
    lib.gfnff_set_precision.argtypes = [C.c_void_p, C.c_int32]
    lib.gfnff_has_error.argtypes = [C.c_void_p]
    lib.gfnff_last_error.argtypes = [C.c_void_p]
    lib.gfnff_last_error.restype = C.c_char_p
    lib.gfnff_get_params.argtypes = [C.c_void_p, C.POINTER(ParamsView)]
    lib.gfnff_set_solvent.argtypes = [C.c_void_p, C.c_char_p]
    lib.gfnff_solvation_energy.argtypes = [C.c_void_p, _f64p]
    lib.gfnff_get_solvation.argtypes = [C.c_void_p, C.POINTER(AlpbView)]
    lib.gfnff_get_topology.argtypes = [C.c_void_p, _i32p, _i32p, _i32p, C.c_int32, _i32p]
    lib.gfnff_device_count.restype = C.c_int
    lib.gfnff_solver_iterations.argtypes = [C.c_void_p]
    lib.gfnff_dense_cholesky.argtypes = [C.c_int32, _f64p, _f64p]
    lib.gfnff_dense_cholesky_f32.argtypes = [C.c_int32, C.c_void_p, C.c_void_p]
    lib.gfnff_dense_cholesky_noalloc.argtypes = [C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    lib.gfnff_param_count.argtypes = [C.c_void_p]
    lib.gfnff_param_scales.argtypes = [C.c_void_p, _f64p, C.c_int32]
    lib.gfnff_param_adjoint.argtypes = [C.c_void_p, _f64p, _f64p, C.c_int32]
    lib.gfnff_cn_neighbours.argtypes = [C.c_void_p, _i32p, _i32p, C.c_int32]
    lib.gfnff_profile_kernel.argtypes = [C.c_void_p, C.c_char_p]
    lib.gfnff_profile_read.argtypes = [C.c_void_p, _f64p, _i32p]
    _LIB = lib
    return lib


def as_i32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.int32)


def as_f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


def iptr(a: np.ndarray):
    return a.ctypes.data_as(_i32p)


def fptr(a: np.ndarray):
    return a.ctypes.data_as(_f64p)
