"""Benchmark: GFN-FF energy + gradient throughput on one MI355X.

Headline workload (one "step" = one full energy+gradient evaluation of every
molecule): mixed_batch_4096, 4096 independent molecules (all-trans n-alkanes
and extended glycine oligopeptides, 20-400 atoms each, sigma 0.03 A, seed 3)
on the batched mixed-precision path.  Beside it, in the same run:
  * the same batch on the FP64 path;
  * the plain FP64 single-molecule path on the 3000-H2O water cluster
    (9000 atoms), the denominator of the mixed-batched / plain-FP64 ratio;
  * accuracy of the headline against the FP64 CPU oracle on a sample of the
    batch's molecules, and against the FP64 batch on every atom.
Coordinates, outputs and all parameters are resident in HBM before the timed
region starts; the host only enqueues kernels.

Prints one JSON line (see the repository's bench contract).  Extra objects:
  roofline      dominant kernel, algorithmic work per launch / measured launch time
  cpu_baseline  the threaded CPU oracle timed on this host on a bounded sample
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import statistics
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "GFN-FF E+grad atom·evals/s, mixed vs FP64 on 1 GPU at |ΔE|≤1e-6 Eh/atom"
BOHR = 0.529177210903


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=20)
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--workload", default="mixed_batch_4096")
    p.add_argument("--precision", default="mixed", choices=["fp64", "mixed"],
                   help="path of the headline number; the FP64 path is always timed beside it")
    p.add_argument("--plain-workload", default="water_3000",
                   help="single-molecule workload timed on the plain FP64 path ('' to skip)")
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--cpu-sample-mols", type=int, default=1024,
                   help="molecules of a batch workload timed by the CPU oracle")
    p.add_argument("--cpu-sample", default="water_400", help="CPU sample of a single-molecule workload")
    p.add_argument("--oracle-check-mols", type=int, default=12,
                   help="batch molecules checked against the FP64 oracle (0 to skip)")
    p.add_argument("--lookahead", type=int, default=0, help="engine option cholesky_lookahead (0/1)")
    p.add_argument("--solver", type=int, default=0, help="engine option eeq_solver (0 PCG, 1 Cholesky)")
    p.add_argument("--no-autotune", action="store_true",
                   help="skip the launch-shape autotuner (curcuma_amd/autotune.py) and its JSON cache")
    p.add_argument("--retune", action="store_true", help="tune again even when the cache holds this workload")
    p.add_argument("--no-graph", action="store_true",
                   help="time eager launches for the headline instead of replays of the captured HIP graph")
    p.add_argument("--roofline-kernel", default="auto", help="kernel class timed for the roofline object")
    p.add_argument("--param", action="append", default=[], metavar="KEY=VALUE",
                   help="engine launch-shape key applied over the tuned (or default) ones; repeatable")
    p.add_argument("--copy-calibration-mib", type=int, default=0,
                   help="run one device copy of this size first (HBM counter calibration under rocprofv3 --pmc)")
    return p.parse_args()


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


def batch_workloads():
    from curcuma_amd.workloads import alkane_batch, mixed_batch
    return {"alkane_batch_4096": lambda: alkane_batch(4096, 20, 400, sigma=0.03, seed=3),
            "mixed_batch_4096": lambda: mixed_batch(4096, 20, 400, sigma=0.03, seed=3)}


def load_workload(name):
    """(atom_off or None, Z, xyz[A]) of a named workload."""
    from curcuma_amd.workloads import WORKLOADS
    bw = batch_workloads()
    if name in bw:
        return bw[name]()
    Z, xyz = WORKLOADS[name]()
    return None, Z, xyz


def progress(msg: str) -> None:
    """One progress line on stderr per phase (stdout carries only the JSON line)."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_info():
    model = "unknown"
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    avail = len(os.sched_getaffinity(0))
    env = os.environ.get("OMP_NUM_THREADS")
    threads = min(avail, int(env)) if env and env.isdigit() and int(env) > 0 else avail
    return model, avail, threads


def cpu_baseline(args, atom_off, Z, xyz):
    """Times the threaded CPU oracle (test infrastructure; FFWorkspace-style
    partitions, per-thread accumulators, fixed-order reduce) on a bounded
    sample of the headline workload: 1 warm-up, then the median of 5
    repetitions.  Returns atom*evals/s."""
    from curcuma_amd import GFNFFMethod, Mol
    sys.path.insert(0, str(ROOT / "tests"))
    from conftest import Oracle  # noqa: E402

    model, avail, nthr = cpu_info()
    orc = Oracle()
    jobs = []
    if atom_off is not None:
        nm = min(args.cpu_sample_mols, len(atom_off) - 1)
        for k in range(nm):
            a0, a1 = int(atom_off[k]), int(atom_off[k + 1])
            m = GFNFFMethod()
            if not m.setMolecule(Mol(Z[a0:a1], xyz[a0:a1])):
                return None
            jobs.append((m, m.params_view(), xyz[a0:a1] * (1.0 / BOHR)))
        sample = f"the first {nm} molecules of the same batch ({sum(len(j[2]) for j in jobs)} atoms)"
    else:
        from curcuma_amd.workloads import water_cluster
        nmol = int(args.cpu_sample.split("_")[1])
        Zs, xs = water_cluster(nmol, seed=2)
        m = GFNFFMethod()
        if not m.setMolecule(Mol(Zs, xs)):
            return None
        jobs.append((m, m.params_view(), xs * (1.0 / BOHR)))
        sample = f"water cluster of {nmol} H2O ({len(Zs)} atoms) from the same generator"
    natoms = sum(len(j[2]) for j in jobs)

    def one_pass():
        t0 = time.perf_counter()
        for _, v, x in jobs:
            orc.eval_mt(v, x, nthr, grad=True)
        return time.perf_counter() - t0

    one_pass()  # warm-up
    reps = [one_pass() for _ in range(5)]
    med = statistics.median(reps)
    return {"value": natoms / med, "unit": "atom*evals/s", "cores": nthr, "kind": "port",
            "sample": f"{sample}, FP64 energy+gradient of the threaded CPU oracle with {nthr} threads "
                      f"(one molecule at a time, threads inside the evaluation); 1 warm-up + median of 5 "
                      f"passes, {med:.3f} s per pass",
            "cpu_model": model, "cpus_available": avail}


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


class Run:
    """One engine instance on one workload and precision, with its
    device-resident evaluation step."""

    def __init__(self, args, torch, atom_off, Z, xyz, precision, params=None):
        from curcuma_amd import GFNFFBatch, GFNFFMethod, Mol
        from curcuma_amd._lib import NTERMS
        self.batch = atom_off is not None
        self.n = len(Z)
        dev = torch.device("cuda:0")
        self.x_dev = torch.tensor(xyz * (1.0 / BOHR), dtype=torch.float64, device=dev).contiguous()
        self.stream = torch.cuda.Stream()  # a created stream: graph capture needs one
        self.sp = C.c_void_p(self.stream.cuda_stream)
        cls = GFNFFBatch if self.batch else GFNFFMethod
        self.params = dict(params or {})
        self.m = cls(precision=precision,
                     parameters={"cholesky_lookahead": args.lookahead, "eeq_solver": args.solver, **self.params})
        t0 = time.perf_counter()
        ok = self.m.setBatch(atom_off, Z, xyz) if self.batch else self.m.setMolecule(Mol(Z, xyz))
        if not ok:
            print("set-up failed: " + self.m.getErrorMessage(), file=sys.stderr)
            sys.exit(1)
        self.setup_s = time.perf_counter() - t0
        self.e_dev = torch.zeros(NTERMS + 1, dtype=torch.float64, device=dev)
        self.g_dev = torch.zeros((self.n, 3), dtype=torch.float64, device=dev)

    def step(self):
        lib = self.m._lib
        st = lib.gfnff_energy_gradient_device(self.m.handle, C.c_void_p(self.x_dev.data_ptr()),
                                              C.c_void_p(self.e_dev.data_ptr()), C.c_void_p(self.g_dev.data_ptr()),
                                              self.sp)
        if st != 0:
            raise RuntimeError(lib.gfnff_last_error(self.m.handle).decode())

    def set_params(self, params):
        self.params = dict(params)
        self.m.setParameters(self.params)

    def timed_graph(self, args, torch):
        """Captures one evaluation into a HIP graph and times `steps` replays."""
        self.step()
        torch.cuda.synchronize()
        self.m.captureGraph(self.x_dev.data_ptr(), self.e_dev.data_ptr(), self.g_dev.data_ptr(), self.stream.cuda_stream)
        for _ in range(args.warmup):
            self.m.launchGraph(self.stream.cuda_stream)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            self.m.launchGraph(self.stream.cuda_stream)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        self.graph_value = self.n * args.steps / dt
        self.graph_ms = dt / args.steps * 1e3
        return self

    def timed(self, args, torch, kernel):
        """Warm-up, per-class survey, then the timed region with HIP events on
        the evaluation stream around every launch of `kernel`."""
        from curcuma_amd.profiling import SIDE_STREAM_CLASSES, KernelTimer, survey
        for _ in range(args.warmup):
            self.step()
        torch.cuda.synchronize()
        timer = KernelTimer(self.m.handle, self.m._lib)
        per_class = survey(timer, self.step)
        # the dominant class of the critical path: classes that run on the
        # engine's side stream, overlapped with the evaluation stream, are not
        # candidates (their event times include the sharing)
        cand = [c for c in per_class if c not in SIDE_STREAM_CLASSES] or list(per_class)
        k = kernel if kernel != "auto" else max(cand, key=lambda c: per_class[c][0])
        timer.select(k)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            self.step()
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        k_ms, k_launches = timer.read()
        timer.select(None)
        self.dt = dt
        self.value = self.n * args.steps / dt
        self.ms = dt / args.steps * 1e3
        self.kernel, self.k_ms, self.k_launches, self.per_class = k, k_ms, k_launches, per_class
        return self

    def energy(self):
        return float(self.e_dev[0].item())

    def grad(self):
        return self.g_dev.cpu().numpy()

    def classes(self):
        return {c: round(v[0], 4) for c, v in sorted(self.per_class.items(), key=lambda t: -t[1][0])}


def autotune_run(args, torch, atom_off, Z, xyz):
    """Launch-shape keys of the headline path: from the JSON cache, or tuned
    now (outside every timed region) and stored there."""
    from curcuma_amd import autotune
    device = torch.cuda.get_device_name(0)
    sig = autotune.signature(Z, atom_off, args.precision, args.solver)
    space = autotune.SPACE_BATCH if atom_off is not None else autotune.SPACE_SINGLE
    path = autotune.default_cache_path()
    if not args.retune:
        hit = autotune.lookup(device, sig, path)
        if hit is not None:
            return {"params": hit["params"], "from_cache": True, "cache": str(path), "timings_ms": hit.get("timings_ms")}
    run = Run(args, torch, atom_off, Z, xyz, args.precision)
    defaults = {k: autotune.ENGINE_DEFAULTS[k] for k in space}

    def make(p):
        run.set_params({**defaults, **p})
        return run.m

    best, timings = autotune.tune(torch, make, space, run.x_dev, run.e_dev, run.g_dev, run.stream, base=defaults)
    cache = autotune.load_cache(path)
    cache[f"{device}|{sig}"] = {"params": best, "timings_ms": timings}
    try:
        autotune.save_cache(cache, path)
    except OSError as ex:
        print(f"autotune cache not written: {ex}", file=sys.stderr)
    del run
    return {"params": best, "from_cache": False, "cache": str(path), "timings_ms": timings}


def oracle_check(head, atom_off, Z, xyz, nmol):
    """Headline batch against the FP64 CPU oracle on `nmol` molecules spread
    over the batch (per-molecule energies and gradient blocks)."""
    from curcuma_amd import GFNFFMethod, Mol
    sys.path.insert(0, str(ROOT / "tests"))
    from conftest import Oracle  # noqa: E402
    _, _, nthr = cpu_info()
    orc = Oracle()
    e_mol = head.m.getEnergies()
    g = head.grad()
    total = len(atom_off) - 1
    pick = sorted(set(np.linspace(0, total - 1, nmol).round().astype(int).tolist()))
    de, rms = [], []
    for k in pick:
        a0, a1 = int(atom_off[k]), int(atom_off[k + 1])
        m = GFNFFMethod()
        if not m.setMolecule(Mol(Z[a0:a1], xyz[a0:a1])):
            raise RuntimeError(m.getErrorMessage())
        ref = orc.eval_mt(m.params_view(), xyz[a0:a1] * (1.0 / BOHR), nthr)
        de.append(abs(e_mol[k] - ref["E"]) / (a1 - a0))
        rms.append(float(np.sqrt(np.mean((g[a0:a1] - ref["grad"]) ** 2))))
    return {"vs": f"FP64 CPU oracle, {len(pick)} molecules spread over the batch (indices {pick[0]}..{pick[-1]})",
            "dE_per_atom_Eh_max": max(de), "grad_rms_Eh_per_bohr_max": max(rms)}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); "
              "the engine has no multi-GPU path in this build", file=sys.stderr)
        sys.exit(2)
    import torch
    from curcuma_amd.profiling import roofline

    if args.copy_calibration_mib > 0:
        nel = args.copy_calibration_mib * (1 << 20) // 8
        a = torch.ones(nel, dtype=torch.float64, device="cuda:0")
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        del a, b

    atom_off, Z, xyz = load_workload(args.workload)
    batch = atom_off is not None
    n = len(Z)

    # FP64 path (the reference's arithmetic) and the headline path on the same input
    progress(f"workload {args.workload}: {n} atoms")
    tuned = None
    if not args.no_autotune:
        progress("autotune")
        tuned = autotune_run(args, torch, atom_off, Z, xyz)
    params = dict(tuned["params"]) if tuned else {}
    overrides = {}
    for kv in args.param:
        k, _, v = kv.partition("=")
        overrides[k.strip()] = float(v) if "." in v else int(v)
    params.update(overrides)
    progress("FP64 path")
    r64 = Run(args, torch, atom_off, Z, xyz, "fp64", params if args.precision == "fp64" else (overrides or None))
    r64.timed(args, torch, args.roofline_kernel if args.precision == "fp64" else "pairs")
    head = r64
    if args.precision == "mixed":
        progress("mixed path")
        head = Run(args, torch, atom_off, Z, xyz, "mixed", params).timed(args, torch, args.roofline_kernel)
    eager_value, eager_ms = head.value, head.ms
    if not args.no_graph:
        progress("graph replay")
        head.timed_graph(args, torch)
        head.value, head.ms = head.graph_value, head.graph_ms

    g64, gh = r64.grad(), head.grad()
    accuracy = {"mixed_vs_fp64_gpu": {"dE_per_atom_Eh": abs(head.energy() - r64.energy()) / n,
                                      "grad_rms_Eh_per_bohr": float(np.sqrt(np.mean((gh - g64) ** 2)))},
                "bar": "|dE| <= 1e-6 Eh/atom, grad RMS <= 1e-5 Eh/Bohr"}
    if batch and args.oracle_check_mols > 0:
        progress("oracle check")
        accuracy["oracle"] = oracle_check(head, atom_off, Z, xyz, args.oracle_check_mols)

    pv = head.m.params_view()
    nr = pv.natoms - pv.nfrag
    info = {"xyz_bohr": xyz * (1.0 / BOHR), "np_pad": max(64, -(-nr // 64) * 64), "nhb": pv.nhb,
            "nbond": pv.nbond}
    pcg_iters = head.m.solverIterations() if not batch else 0
    info["pcg_iters"] = pcg_iters
    if batch:
        frag = np.ctypeslib.as_array(pv.frag, shape=(n,))
        # the engine factors in FP32 only in the forms that build the matrix first (0, 1, 3)
        fp32_solve = (args.precision == "mixed" and head.params.get("batch_fp32", 0) == 1
                      and head.params.get("batch_solver", 5) in (0, 1, 3))
        info["batch_fp32"] = fp32_solve
        info["batch_refine_steps"] = 2 if fp32_solve else 0  # BSOLVE_NREF of the engine
        info["mol_off"] = atom_off
        info["mol_nr"] = [int(atom_off[k + 1] - atom_off[k] - len(np.unique(frag[atom_off[k]:atom_off[k + 1]])))
                          for k in range(len(atom_off) - 1)]
    roof = roofline(head.kernel, head.k_ms, head.k_launches, args.steps, info,
                    traffic=pmc_traffic(head.kernel, args.workload), precision=args.precision)
    roof["kernel_ms_per_eval"] = head.classes()

    plain = None
    if batch and args.plain_workload:
        progress(f"plain FP64 path on {args.plain_workload}")
        p_off, pZ, pxyz = load_workload(args.plain_workload)
        rp = Run(args, torch, p_off, pZ, pxyz, "fp64").timed(args, torch, "pairs")
        plain = {"workload": WORKLOAD_TEXT.get(args.plain_workload, args.plain_workload), "atoms": len(pZ),
                 "precision": "fp64", "eeq_solver": "PCG" if args.solver == 0 else "blocked Cholesky",
                 "value": rp.value, "ms_per_step": rp.ms, "energy_Eh": rp.energy(),
                 "kernel_ms_per_eval": rp.classes()}
        del rp

    cpu = None
    if not args.no_cpu_baseline:
        try:
            progress("CPU baseline")
            cpu = cpu_baseline(args, atom_off, Z, xyz)
        except Exception as ex:
            print(f"cpu baseline failed: {ex}", file=sys.stderr)

    out = {
        "metric": METRIC,
        "value": head.value,
        "unit": "atom*evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head.ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32+f64",
        "data": "synthetic",
        "config": {"workload": WORKLOAD_TEXT.get(args.workload, args.workload),
                   "molecules": (len(atom_off) - 1) if batch else 1,
                   "atoms": n, "precision": args.precision, "setup_s": round(head.setup_s, 3),
                   "energy_Eh": head.energy(),
                   "eeq_solver": ("per-molecule Cholesky" if batch else
                                  ("PCG" if args.solver == 0 else "blocked Cholesky")),
                   "pcg_iterations": pcg_iters},
        "launch": {"mode": "eager" if args.no_graph else "HIP graph replay",
                   "graph_ms_per_step": None if args.no_graph else head.graph_ms,
                   "eager_ms_per_step": eager_ms, "eager_value": eager_value,
                   "engine_params": head.params, "autotune": tuned},
        "fp64": {"value": r64.value, "ms_per_step": r64.ms, "energy_Eh": r64.energy(),
                 "kernel_ms_per_eval": r64.classes()},
        "mixed_over_fp64": head.value / r64.value,
        "plain_fp64": plain,
        "mixed_batched_over_plain_fp64": (head.value / plain["value"]) if plain else None,
        "accuracy": accuracy,
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
