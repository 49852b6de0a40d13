"""Kernel timing and roofline figures for one GFN-FF evaluation.

Launch times come from HIP events that the engine records around every launch
of one kernel class, on the stream the kernels run on
(gfnff_profile_kernel / gfnff_profile_read).  Algorithmic work per launch
follows the fixed convention in include/roofline_counts.h; the constants are
read from that header so that the engine, the benchmark and the documentation
share one source.
"""
from __future__ import annotations

import ctypes as C
import re
from pathlib import Path

import numpy as np

from ._lib import load_library

HEADER = Path(__file__).resolve().parent.parent / "include" / "roofline_counts.h"

PEAK_FP64 = 78.6e12    # MI355X dense FP64, vector and matrix (FLOP/s)
PEAK_FP32 = 157.3e12   # MI355X FP32, matrix and (packed) vector
PEAK_HBM = 8.0e12      # bytes/s

NB = 64                # Cholesky block size of the engine

# every kernel class the engine can time
CLASSES = ["cn", "eeq_build", "chol_diag", "chol_panel", "chol_update", "trsv", "charges", "pairs", "pairs_aux", "nbrep", "refine",
           "bonds", "angles", "torsions", "brep", "hb", "atm", "batm", "chain", "reduce", "pcg", "batch_solve"]
# classes whose kernels run on the engine's side stream, overlapped with the
# evaluation stream (the reduced EEQ matrix build)
SIDE_STREAM_CLASSES = {"eeq_build"}


def read_counts(path: Path = HEADER) -> dict:
    """The GFNFF_RF_* integers of include/roofline_counts.h."""
    out = {}
    for m in re.finditer(r"#define\s+GFNFF_RF_(\w+)\s+(\d+)", path.read_text()):
        out[m.group(1)] = int(m.group(2))
    return out


def pair_counts(xyz_bohr: np.ndarray, cuts=(100.0, 50.0, 20.0), chunk: int = 1024) -> list[int]:
    """Unordered pairs (i < j) with r <= cut for each cut-off."""
    x = np.ascontiguousarray(xyz_bohr, dtype=np.float64)
    n = len(x)
    c2 = [c * c for c in cuts]
    cnt = [0] * len(cuts)
    for a in range(0, n, chunk):
        blk = x[a:a + chunk]
        d2 = ((blk[:, None, :] - x[None, :, :]) ** 2).sum(-1)
        rows = np.arange(a, a + len(blk))[:, None]
        upper = np.arange(n)[None, :] > rows
        for k, cc in enumerate(c2):
            cnt[k] += int(np.count_nonzero((d2 <= cc) & upper))
    return cnt


class KernelTimer:
    """HIP-event timing of one kernel class inside the engine."""

    def __init__(self, handle, lib=None):
        self.lib = lib or load_library()
        self.h = handle

    def select(self, name: str | None):
        st = self.lib.gfnff_profile_kernel(self.h, (name or "").encode())
        if st != 0:
            raise RuntimeError(self.lib.gfnff_last_error(self.h).decode())

    def read(self) -> tuple[float, int]:
        ms = C.c_double(0.0)
        nl = C.c_int32(0)
        st = self.lib.gfnff_profile_read(self.h, C.byref(ms), C.byref(nl))
        if st != 0:
            raise RuntimeError(self.lib.gfnff_last_error(self.h).decode())
        return ms.value, nl.value


def survey(timer: KernelTimer, step, classes=CLASSES) -> dict:
    """One evaluation per class; returns {class: (ms per eval, launches per eval)}."""
    out = {}
    for c in classes:
        timer.select(c)
        step()
        ms, nl = timer.read()
        if nl:
            out[c] = (ms, nl)
    timer.select(None)
    return out


def chol_update_flops(np_pad: int, nb: int = NB) -> float:
    """Algorithmic flops of all trailing updates of one blocked factorisation."""
    nbk = np_pad // nb
    tot = 0.0
    for k in range(nbk - 1):
        m = (nbk - k - 1) * nb
        tot += float(m) * (m + 1) * nb
    return tot


def work_per_eval(cls: str, info: dict, rc: dict) -> tuple[float, str]:
    """Algorithmic work of one evaluation of kernel class `cls`.

    info: {"xyz_bohr", "np_pad", "n_reduced", "nfrag", "nhb", "nbond", ...}.
    Returns (work, "flop" | "byte")."""
    n = len(info["xyz_bohr"])
    if "mol_off" in info:  # batch of independent molecules
        off = info["mol_off"]
        if cls == "pairs":
            tot = [0, 0, 0]
            for k in range(len(off) - 1):
                c = pair_counts(info["xyz_bohr"][off[k]:off[k + 1]])
                tot = [a + b for a, b in zip(tot, c)]
            nc, nd, nr = tot
            return (nc * rc["COULOMB_PAIR_FLOPS"] + nd * rc["DISP_PAIR_FLOPS"] + nr * rc["REP_PAIR_FLOPS"]), "flop"
        if cls == "batch_solve":  # factorisation + both substitutions per molecule (+ FP64 refinement steps)
            nref = info.get("batch_refine_steps", 0)
            return sum(m ** 3 / rc["CHOL_DIV"] + rc["TRSV_FACTOR"] * m * m + nref * 4.0 * m * m
                       for m in info["mol_nr"]), "flop"
    if cls == "pairs":
        nc, nd, nr = info.get("pairs") or pair_counts(info["xyz_bohr"])
        return (nc * rc["COULOMB_PAIR_FLOPS"] + nd * rc["DISP_PAIR_FLOPS"] + nr * rc["REP_PAIR_FLOPS"]), "flop"
    if cls == "chol_update":
        return chol_update_flops(info["np_pad"]), "flop"
    if cls == "chol_diag":
        nbk = info["np_pad"] // NB
        return nbk * 2.0 * NB ** 3 / rc["CHOL_DIV"], "flop"  # factor + triangular inverse of each block
    if cls == "chol_panel":
        nbk = info["np_pad"] // NB
        return sum(2.0 * (nbk - k - 1) * NB * NB * NB / 2 for k in range(nbk)), "flop"
    if cls == "trsv":
        npd = info["np_pad"]
        return 2.0 * 8.0 * npd * npd / 2, "byte"  # the factor is read once per sweep
    if cls == "pcg":  # the packed symmetric FP64 matrix (64x64 lower blocks) is streamed once per CG iteration
        nbk = info["np_pad"] // NB
        return info.get("pcg_iters", 0) * 8.0 * (nbk * (nbk + 1) // 2) * NB * NB, "byte"
    if cls == "eeq_build":
        npd = info["np_pad"]
        return 8.0 * npd * (npd + NB) / 2, "byte"
    if cls == "hb":
        return info["nhb"] * (rc["HB_PAR_BYTES"] + 3 * rc["ATOM_GATHER_BYTES"]), "byte"
    if cls == "cn":
        return n * (3 * 8 + 8 * 8), "byte"
    if cls == "bonds":
        return info["nbond"] * (rc["BOND_PAR_BYTES"] + 4 * rc["ATOM_GATHER_BYTES"]), "byte"
    raise KeyError(f"no work convention for kernel class {cls}")


# kernel classes that run in FP32 on the mixed-precision path (the batched
# per-molecule solve, class batch_solve, factors in FP32 only with the engine
# key batch_fp32 = 1: info["batch_fp32"])
FP32_CLASSES_MIXED = {"pairs", "chol_update", "chol_diag", "chol_panel", "trsv"}
# classes whose flops run on the matrix cores
MFMA_CLASSES = {"chol_update", "chol_diag", "chol_panel", "batch_solve"}


def roofline(cls: str, ms_total: float, launches: int, evals: int, info: dict, traffic=None,
             precision: str = "fp64") -> dict:
    """The bench's roofline object for `cls` from event times over `evals` evaluations."""
    rc = read_counts()
    work, kind = work_per_eval(cls, info, rc)
    per_launch = work * evals / max(launches, 1)
    avg_s = ms_total * 1e-3 / max(launches, 1)
    achieved = per_launch / avg_s
    if kind == "flop":
        fp32 = precision == "mixed" and (cls in FP32_CLASSES_MIXED or (cls == "batch_solve" and info.get("batch_fp32")))
        peak = PEAK_FP32 if fp32 else PEAK_FP64
        bound = "mfma" if cls in MFMA_CLASSES else "valu"
        unit = "TFLOP/s"
        scale = 1e12
    else:
        peak = PEAK_HBM
        bound = "hbm"
        unit = "GB/s"
        scale = 1e9
    return {"kernel": cls, "bound": bound, "achieved": achieved / scale, "peak": peak / scale, "unit": unit,
            "frac": achieved / peak, "traffic": traffic, "launches": launches,
            "avg_launch_us": avg_s * 1e6, "work_per_launch": per_launch, "work_unit": kind}
