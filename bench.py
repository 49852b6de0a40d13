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


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=20)
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--workload", default="water_3000")
    p.add_argument("--precision", default="mixed", choices=["fp64", "mixed"],
                   help="path of the headline number; the FP64 path is always timed beside it")
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--cpu-sample", default="water_400")
    p.add_argument("--lookahead", type=int, default=0, help="engine option cholesky_lookahead (0/1)")
    p.add_argument("--solver", type=int, default=0, help="engine option eeq_solver (0 PCG, 1 Cholesky)")
    p.add_argument("--roofline-kernel", default="auto", help="kernel class timed for the roofline object")
    p.add_argument("--copy-calibration-mib", type=int, default=0,
                   help="run one device copy of this size first (HBM counter calibration under rocprofv3 --pmc)")
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
        if time.perf_counter() - t0 > 10.0 and reps >= 3:
            break
    dt = (time.perf_counter() - t0) / reps
    return {"value": len(Z) / dt, "unit": "atom*evals/s", "cores": 1, "kind": "port",
            "sample": f"water cluster of {nmol} H2O ({len(Z)} atoms) from the same generator, "
                      f"{reps} FP64 energy+gradient evaluation(s) of the single-threaded CPU oracle, "
                      f"{dt:.3f} s each"}


WORKLOAD_TEXT = {
    "water_3000": "water_3000: non-periodic water cluster, 3000 H2O (9000 atoms), 3.1 A cubic lattice, "
                  "seeded random orientations, +-0.05 A jitter",
    "alkane_batch_4096": "alkane_batch_4096: 4096 independent all-trans n-alkanes, atom counts uniform in "
                         "[20, 400] rounded to CnH(2n+2), sigma 0.03 A, seed 3, one launch per kernel",
    "mixed_batch_4096": "mixed_batch_4096: 4096 independent molecules, each a seeded random choice of an "
                        "all-trans n-alkane or an extended glycine oligopeptide H-(Gly)n-OH, atom counts "
                        "uniform in [20, 400] rounded down to the family size (3n+2 / 7n+3), sigma 0.03 A, "
                        "seed 3, one launch per kernel",
}


def pmc_traffic(kernel: str, workload: str):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 --pmc
    summary (profiles/pmc_<kernel>_<workload>.md, line 'traffic_bytes_per_launch: X'),
    or None when no such summary exists."""
    f = ROOT / "profiles" / f"pmc_{kernel}_{workload}.md"
    if not f.exists():
        return None
    for line in f.read_text().splitlines():
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":", 1)[1])
    return None


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); "
              "the engine has no multi-GPU path in this build", file=sys.stderr)
        sys.exit(2)
    import torch
    from curcuma_amd import GFNFFBatch, GFNFFMethod, Mol
    from curcuma_amd._lib import NTERMS
    from curcuma_amd.workloads import WORKLOADS, alkane_batch, mixed_batch

    BATCH_WORKLOADS = {"alkane_batch_4096": lambda: alkane_batch(4096, 20, 400, sigma=0.03, seed=3),
                       "mixed_batch_4096": lambda: mixed_batch(4096, 20, 400, sigma=0.03, seed=3)}

    if args.copy_calibration_mib > 0:
        nel = args.copy_calibration_mib * (1 << 20) // 8
        a = torch.ones(nel, dtype=torch.float64, device="cuda:0")
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        del a, b

    batch = args.workload in BATCH_WORKLOADS
    if batch:
        atom_off, Z, xyz = BATCH_WORKLOADS[args.workload]()
    else:
        Z, xyz = WORKLOADS[args.workload]()
        atom_off = None
    n = len(Z)
    dev = torch.device("cuda:0")
    x_dev = torch.tensor(xyz * (1.0 / 0.529177210903), dtype=torch.float64, device=dev).contiguous()
    stream = torch.cuda.current_stream()
    sp = C.c_void_p(stream.cuda_stream)

    def make(precision):
        cls = GFNFFBatch if batch else GFNFFMethod
        m = cls(precision=precision, parameters={"cholesky_lookahead": args.lookahead, "eeq_solver": args.solver})
        t0 = time.perf_counter()
        ok = m.setBatch(atom_off, Z, xyz) if batch else m.setMolecule(Mol(Z, xyz))
        if not ok:
            print("set-up failed: " + m.getErrorMessage(), file=sys.stderr)
            sys.exit(1)
        e_dev = torch.zeros(NTERMS + 1, dtype=torch.float64, device=dev)
        g_dev = torch.zeros((n, 3), dtype=torch.float64, device=dev)
        lib = m._lib

        def step():
            st = lib.gfnff_energy_gradient_device(m.handle, C.c_void_p(x_dev.data_ptr()),
                                                  C.c_void_p(e_dev.data_ptr()), C.c_void_p(g_dev.data_ptr()), sp)
            if st != 0:
                raise RuntimeError(lib.gfnff_last_error(m.handle).decode())
        return m, step, e_dev, g_dev, time.perf_counter() - t0

    from curcuma_amd.profiling import SIDE_STREAM_CLASSES, KernelTimer, roofline, survey

    def timed(m, step, kernel):
        """Warm-up, per-class survey, then the timed region with HIP events on
        the evaluation stream around every launch of `kernel`."""
        for _ in range(args.warmup):
            step()
        torch.cuda.synchronize()
        timer = KernelTimer(m.handle, m._lib)
        per_class = survey(timer, step)
        # the dominant class of the critical path: classes that run on the
        # engine's side stream, overlapped with the evaluation stream, are not
        # candidates (their event times include the sharing)
        cand = [c for c in per_class if c not in SIDE_STREAM_CLASSES] or list(per_class)
        k = kernel if kernel != "auto" else max(cand, key=lambda c: per_class[c][0])
        timer.select(k)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            step()
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        k_ms, k_launches = timer.read()
        timer.select(None)
        return dt, k, k_ms, k_launches, per_class

    # FP64 path (the reference's arithmetic) and the mixed-precision path on the same input
    m64, step64, e64_dev, g64_dev, setup64 = make("fp64")
    dt64, _, _, _, classes64 = timed(m64, step64, "pairs")
    headline = args.precision
    if headline == "mixed":
        mh, steph, eh_dev, gh_dev, setup_h = make("mixed")
        dt, kernel, k_ms, k_launches, per_class = timed(mh, steph, args.roofline_kernel)
    else:
        mh, steph, eh_dev, gh_dev, setup_h = m64, step64, e64_dev, g64_dev, setup64
        dt, kernel, k_ms, k_launches, per_class = timed(mh, steph, args.roofline_kernel)
    ms = dt / args.steps * 1e3
    value = n * args.steps / dt
    value64 = n * args.steps / dt64

    e64 = float(e64_dev[0].item())
    eh = float(eh_dev[0].item())
    g64 = g64_dev.cpu().numpy()
    gh = gh_dev.cpu().numpy()
    accuracy = {"vs": "FP64 GPU path (parity-pinned to the FP64 oracle)",
                "dE_per_atom_Eh": abs(eh - e64) / n,
                "grad_rms_Eh_per_bohr": float(np.sqrt(np.mean((gh - g64) ** 2))),
                "bar": "|dE| <= 1e-6 Eh/atom, grad RMS <= 1e-5 Eh/Bohr"}

    pv = mh.params_view()
    nr = pv.natoms - pv.nfrag
    info = {"xyz_bohr": xyz * (1.0 / 0.529177210903), "np_pad": max(64, -(-nr // 64) * 64), "nhb": pv.nhb,
            "nbond": pv.nbond}
    pcg_iters = mh.solverIterations() if not batch else 0
    info["pcg_iters"] = pcg_iters
    if batch:
        frag = np.ctypeslib.as_array(pv.frag, shape=(n,))
        info["mol_off"] = atom_off
        info["mol_nr"] = [int(atom_off[k + 1] - atom_off[k] - len(np.unique(frag[atom_off[k]:atom_off[k + 1]])))
                          for k in range(len(atom_off) - 1)]
    roof = roofline(kernel, k_ms, k_launches, args.steps, info, traffic=pmc_traffic(kernel, args.workload),
                    precision=headline)
    roof["kernel_ms_per_eval"] = {c: round(v[0], 4) for c, v in sorted(per_class.items(), key=lambda t: -t[1][0])}
    t_setup = setup_h
    e_last = eh
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
        "dtype": "f64" if headline == "fp64" else "f32+f64",
        "data": "synthetic",
        "config": {"workload": WORKLOAD_TEXT.get(args.workload, args.workload),
                   "atoms": n, "precision": headline, "setup_s": round(t_setup, 3), "energy_Eh": e_last,
                   "eeq_solver": ("per-molecule Cholesky" if batch else
                                  ("PCG" if args.solver == 0 else "blocked Cholesky")),
                   "pcg_iterations": pcg_iters},
        "fp64": {"value": value64, "ms_per_step": dt64 / args.steps * 1e3, "energy_Eh": e64,
                 "kernel_ms_per_eval": {c: round(v[0], 4) for c, v in
                                        sorted(classes64.items(), key=lambda t: -t[1][0])}},
        "mixed_over_fp64": value / value64,
        "accuracy": accuracy,
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
