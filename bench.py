"""Benchmark: mixed-precision softened all-pairs KDK on one MI355X.

Workload: the two-disk collision at N = 65,536 (32,768 per galaxy: central
mass 1.0 + disk mass 0.1), eps = 0.02, dt = 5e-4, direct all-pairs forces
with FP32 pairs on FP32 offsets from FP64 tile centres, two-sum
accumulation and FP64 state; one step = one kick-drift-kick leapfrog step
(one force evaluation, N(N-1) pair interactions) with the particles
resident in HBM.

Prints one JSON line.  value = Gpair-interactions/s over the timed steps;
roofline = the dominant force kernel (FP32 VALU, 20 flop per pair) against
the FP32 vector peak; cpu_baseline = the FP64 oracle (the reference's
targets-form arithmetic) on all host cores for one full force evaluation.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "galaxy-collision-simulation_amd"))

FLOP_PER_PAIR = 20.0          # GPU Gems 3 ch. 31 convention
FP32_PEAK_TFLOPS = 157.3      # MI355X FP32 vector
HBM_PEAK_GBS = 8000.0
METRIC = "Gpair-interactions/s + %FP32 peak, mixed vs FP64 GPU @<=1e-5 rel force err"


def _traffic(kernel: str):
    """HBM bytes per launch from the committed rocprofv3 PMC summary."""
    path = ROOT / "profiles" / "traffic.txt"
    if not path.exists():
        return None
    for line in path.read_text().splitlines():
        parts = line.split()
        if len(parts) >= 2 and parts[0] == kernel:
            try:
                return float(parts[1])
            except ValueError:
                return None
    return None


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=65536)
    ap.add_argument("--precision", default="mixed", choices=["mixed", "fp64"])
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 GPU comparison")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for {args.gpus}); "
              "multi-GPU sharding is described in DESIGN.md, not launched.", file=sys.stderr)
        return 2

    import gcsim
    from gcsim import generators

    if not gcsim.solver_available():
        print("bench.py: no gfx950 device visible", file=sys.stderr)
        return 1

    n, eps, dt = args.n, 0.02, 5e-4
    pos, vel, mass = generators.config1(n=n)
    pairs = float(n) * float(n - 1)

    eng = gcsim.Engine("hip-direct", args.precision, softening=eps)
    eng.upload(pos, vel, mass)
    eng.compute_accelerations()
    _, _, a_step0 = eng.download()
    eng.step(dt, args.warmup)
    eng.synchronize()
    t0 = time.perf_counter()
    eng.step(dt, args.steps)
    eng.synchronize()
    wall = time.perf_counter() - t0
    ms_per_step = wall * 1e3 / args.steps
    value = pairs * args.steps / wall / 1e9

    # dominant kernel: HIP events around back-to-back launches on its stream
    k_ms, k_pairs = eng.time_force_kernel(reps=20)
    achieved = FLOP_PER_PAIR * k_pairs / (k_ms * 1e-3) / 1e12
    ipt, jsplit = eng.launch_shape()
    diag = eng.diagnostics()
    eng.close()

    kernel_name = "direct_mixed_kernel" if args.precision == "mixed" else "direct_fp64_kernel"
    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "Gpair-interactions/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if args.precision == "mixed" else "f64",
        "data": "synthetic",
        "config": {
            "workload": "two-disk collision, direct all-pairs KDK, N=65536, eps=0.02, dt=5e-4",
            "n": n, "eps": eps, "dt": dt, "precision": args.precision,
            "state": "fp64 resident in HBM", "targets_per_thread": ipt, "jsplit": jsplit,
            "parallelism": "single GPU",
        },
        "roofline": {
            "bound": "valu",
            "kernel": kernel_name,
            "achieved": round(achieved, 3),
            "peak": FP32_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": round(achieved / FP32_PEAK_TFLOPS, 4),
            "traffic": _traffic(kernel_name),
            "kernel_ms": round(k_ms, 4),
            "flop_per_pair": FLOP_PER_PAIR,
        },
        "energy_after_steps": diag["total_energy"],
    }

    if not args.no_fp64 and args.precision == "mixed":
        e64 = gcsim.Engine("hip-direct", "fp64", softening=eps)
        e64.upload(pos, vel, mass)
        e64.compute_accelerations()
        _, _, a64 = e64.download()
        f_ms, _ = e64.time_force_kernel(reps=3)
        e64.close()
        err = np.linalg.norm(a_step0 - a64, axis=1) / np.maximum(np.linalg.norm(a64, axis=1), 1e-12)
        fp64_rate = pairs / (f_ms * 1e-3) / 1e9
        out["fp64_gpu"] = {"kernel_ms": round(f_ms, 3), "Gpair_per_s": round(fp64_rate, 3),
                           "mixed_over_fp64": round((k_pairs / (k_ms * 1e-3) / 1e9) / fp64_rate, 3)}
        out["accuracy_vs_fp64"] = {"max_rel_force_err": float(err.max()),
                                   "median_rel_force_err": float(np.median(err))}

    if not args.no_cpu and int(os.environ.get("RANK", "0")) == 0:
        sys.path.insert(0, str(ROOT / "oracle"))
        import oracle  # CPU baseline leg only

        threads = oracle.direct_targets_omp_threads(0)
        t0 = time.perf_counter()
        a_cpu = oracle.direct_targets(pos, mass, 1.0, eps, threads=threads)
        cpu_s = time.perf_counter() - t0
        out["cpu_baseline"] = {
            "value": round(pairs / cpu_s / 1e9, 4),
            "unit": "Gpair-interactions/s",
            "cores": threads,
            "kind": "port",
            "sample": f"one full FP64 force evaluation at N={n} (targets form, "
                      f"{pairs:.3e} pairs) on {threads} OpenMP threads, {cpu_s:.2f} s",
        }
        if "accuracy_vs_fp64" in out:
            err = (np.linalg.norm(a_step0 - a_cpu, axis=1)
                   / np.maximum(np.linalg.norm(a_cpu, axis=1), 1e-12))
            out["accuracy_vs_cpu_oracle"] = {"max_rel_force_err": float(err.max()),
                                             "median_rel_force_err": float(np.median(err))}

    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
