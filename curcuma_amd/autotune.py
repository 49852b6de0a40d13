"""Launch-shape autotuner for the GFN-FF engine, with a JSON cache.

The engine keys tuned here change only how the work is cut into launches,
except batch_fp32, which trades the FP64 factor for an FP32 one refined in
FP64.  The batch-solve variants, the PCG workgroup counts and the gather
widths reorder no sum, so their results are bitwise the same; the two batched
pair decompositions add each atom's pair terms in a different order, so
theirs agree to rounding:
  batch_solver   0..6     per-molecule Cholesky of batches: right-looking with
                          tile phases shared out over wavefronts in two
                          occupancy / operand-prefetch variants, one tile step
                          per workgroup, or one column block at a time over
                          the whole batch; left-looking with the matrix formed
                          inside the factorisation, per molecule or one
                          column block at a time over the batch (6: with the
                          panel product inside the tiles below the diagonal)
  batch_fp32     0 / 1    mixed path: factor in FP64, or in FP32 on MFMA with
                          two FP64 refinement steps (charges agree with the
                          FP64 factor's to about 1e-12 e)
  pair_split     0 / 1    pair terms in one pass after the EEQ solve, or the
                          charge-independent dispersion and repulsion before
                          it (beside the solve) and the Coulomb term after
  batch_pairs    0 / 1    pair terms of batches: each pair once in 64 x 64
                          tiles per molecule, or one thread per row atom
  gather_lanes   16 / 64  lanes per atom in the slot and hydrogen-bond gathers
  pcg_groups     0 / 128 / 192
                          workgroups of the persistent PCG solve (single
                          molecules; 0 = all that fit, at most 256)
A configuration is timed on the device-resident evaluation: warm-up
evaluations, then the median over repetitions of a few back-to-back
evaluations between two HIP events on one stream.  The search is coordinate
descent from the engine's defaults (one key at a time, the best value kept).
The cache maps "device name | workload signature" to the chosen keys; the
signature holds the atom and molecule counts, the element composition, the
precision and the EEQ solver, so a cache entry is reused only for the same
kind of workload on the same kind of device.
"""
from __future__ import annotations

import hashlib
import json
import os
import statistics
from pathlib import Path

import numpy as np

SPACE_BATCH = {"batch_solver": [0, 1, 2, 3, 4, 5, 6], "batch_fp32": [0, 1], "batch_pairs": [0, 1], "pair_split": [0, 1],
               "gather_lanes": [16, 64]}
# the engine's values when a key is not set (gather_lanes 0: by mean list length)
ENGINE_DEFAULTS = {"batch_solver": 5, "batch_fp32": 0, "batch_pairs": 0, "pair_split": 0, "batch_streams": 1,
                   "gather_lanes": 0, "pcg_groups": 0}
SPACE_SINGLE = {"pcg_groups": [0, 128, 192], "pair_split": [0, 1], "gather_lanes": [16, 64]}


def default_cache_path() -> Path:
    env = os.environ.get("CURCUMA_AMD_AUTOTUNE_CACHE")
    if env:
        return Path(env)
    base = os.environ.get("XDG_CACHE_HOME") or str(Path.home() / ".cache")
    return Path(base) / "curcuma_amd" / "autotune.json"


def signature(Z, atom_off, precision: str, solver: int) -> str:
    z = np.asarray(Z, dtype=np.int64)
    comp = np.bincount(z, minlength=87)
    nmol = 1 if atom_off is None else len(atom_off) - 1
    h = hashlib.sha1(comp.tobytes()).hexdigest()[:12]
    return f"n={len(z)};nmol={nmol};elements={h};precision={precision};eeq_solver={solver}"


def load_cache(path: Path | None = None) -> dict:
    p = Path(path) if path else default_cache_path()
    try:
        return json.loads(p.read_text())
    except (OSError, ValueError):
        return {}


def save_cache(cache: dict, path: Path | None = None) -> Path:
    p = Path(path) if path else default_cache_path()
    p.parent.mkdir(parents=True, exist_ok=True)
    tmp = p.with_suffix(".tmp")
    tmp.write_text(json.dumps(cache, indent=1, sort_keys=True))
    os.replace(tmp, p)
    return p


def lookup(device: str, sig: str, path: Path | None = None) -> dict | None:
    return load_cache(path).get(f"{device}|{sig}")


def time_config(torch, make, params: dict, x_dev, e_dev, g_dev, stream, warmup: int = 2, reps: int = 5,
                per_rep: int = 3) -> float:
    """Median ms per evaluation of a fresh engine with `params`."""
    m = make(params)
    sp = stream.cuda_stream
    for _ in range(warmup):
        m.energyGradientDevice(x_dev.data_ptr(), e_dev.data_ptr(), g_dev.data_ptr(), sp)
    times = []
    for _ in range(reps):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(stream)
        for _ in range(per_rep):
            m.energyGradientDevice(x_dev.data_ptr(), e_dev.data_ptr(), g_dev.data_ptr(), sp)
        b.record(stream)
        b.synchronize()
        times.append(a.elapsed_time(b) / per_rep)
    del m
    return statistics.median(times)


def descend(space: dict, measure, base: dict | None = None):
    """Coordinate descent over `space` from `base` with measure(params) -> ms;
    returns (best params, {config label: ms})."""
    best = dict(base or {})
    timings = {}

    def label(p):
        return ",".join(f"{k}={p[k]}" for k in sorted(p)) or "defaults"

    best_ms = measure(best)
    timings[label(best)] = best_ms
    for key, values in space.items():
        for v in values:
            if best.get(key) == v:
                continue
            cand = dict(best)
            cand[key] = v
            ms = measure(cand)
            timings[label(cand)] = ms
            if ms < best_ms:
                best, best_ms = cand, ms
    return best, timings


def tune(torch, make, space: dict, x_dev, e_dev, g_dev, stream, base: dict | None = None, **kw):
    """Times engine configurations on the device and descends over `space`."""
    return descend(space, lambda p: time_config(torch, make, p, x_dev, e_dev, g_dev, stream, **kw), base)
