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
FP64_PEAK_TFLOPS = 78.6       # MI355X FP64 vector
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


def _rel_err(a, ref):
    return np.linalg.norm(a - ref, axis=1) / np.maximum(np.linalg.norm(ref, axis=1), 1e-12)


def _oracle():
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle  # accuracy checker and CPU baseline only

    return oracle


def workload_fmm(args):
    """Plummer sphere N = 2^20, leaves <= 64, p = 4, theta 0.5, eps 0.01, 10 KDK
    steps at dt 1e-3 with FP32 evaluation and FP64 state; accuracy on 4096
    sampled targets against FP64 direct sums."""
    import gcsim

    n, eps, dt, steps = 1 << 20, 0.01, 1e-3, 10
    pos, vel, mass = gcsim.generators.plummer(n, seed=1)
    eng = gcsim.Engine("hip-fmm", "mixed", softening=eps, theta=0.5, leaf_capacity=64,
                       expansion_order=4)
    eng.upload(pos, vel, mass)
    eng.compute_accelerations()
    _, _, a0 = eng.download()
    st0 = eng.fmm_stats()
    eng.synchronize()
    t0 = time.perf_counter()
    eng.step(dt, steps)
    eng.synchronize()
    wall = time.perf_counter() - t0
    st = eng.fmm_stats()
    k_ms, _ = eng.time_force_kernel(reps=3)
    # a target range of N/8 (one rank's share under the reference's
    # ownership split): whole build, the range's targets evaluated
    eng.compute_accelerations(0, n // 8)
    eng.synchronize()
    range_st = eng.fmm_stats()
    eng.close()
    flops = 20.0 * st["p2p_pairs"] + 440.0 * st["m2p_evaluations"]
    out = {
        "workload": "fmm: Plummer N=2^20, leaf<=64, p=4, theta=0.5, eps=0.01, dt=1e-3, 10 KDK",
        "ms_per_step": wall * 1e3 / steps,
        "eval_kernel_ms": k_ms,
        "build_ms": st["build_ms"], "eval_ms": st["eval_ms"],
        "p2p_pairs": st["p2p_pairs"], "m2p_evaluations": st["m2p_evaluations"],
        "Gp2p_per_s_eval": st["p2p_pairs"] / (k_ms * 1e-3) / 1e9,
        "Gm2p_per_s_eval": st["m2p_evaluations"] / (k_ms * 1e-3) / 1e9,
        "tflops_eval_20_440": flops / (k_ms * 1e-3) / 1e12,
        "stats": {k: st0[k] for k in ("node_count", "leaf_count", "far_interactions",
                                      "near_interactions", "max_depth")},
        "range_eighth": {"build_ms": range_st["build_ms"], "eval_ms": range_st["eval_ms"],
                         "eval_over_full": range_st["eval_ms"] / st["eval_ms"]},
    }
    if args.no_cpu:  # GPU legs only (profiling runs)
        return out
    O = _oracle()
    threads = O.direct_targets_omp_threads(0)
    sample = 4096
    direct = O.direct_targets(pos, mass, 1.0, eps, 0, sample, threads=threads)
    err = _rel_err(a0[:sample], direct[:sample])
    tc = time.perf_counter()
    acc_cpu, _ = O.fmm(pos, mass, 1.0, eps, 0.5, 64, 32, 4, begin=0, end=sample)
    cpu_s = time.perf_counter() - tc
    err_vs_cpu_fmm = _rel_err(a0[:sample], acc_cpu[:sample])
    out.update({
        "accuracy_vs_fp64_direct_4096": {"max": float(err.max()), "median": float(np.median(err)),
                                         "p95": float(np.percentile(err, 95))},
        "fp32_vs_oracle_fmm_4096_max": float(err_vs_cpu_fmm.max()),
        "cpu_baseline": {"kind": "port", "cores": threads,
                         "sample": f"oracle FMM, full tree + lists, {sample} targets",
                         "seconds": cpu_s,
                         "extrapolated_s_per_evaluation": cpu_s * n / sample},
    })
    return out


def workload_tree(args):
    """Barnes-Hut (the reference's tree solver) on the FMM workload's Plummer
    sphere N = 2^20 with the tree's defaults (theta 0.6, leaves <= 16, p = 4),
    eps 0.01: 3 KDK steps at dt 1e-3, FP32 evaluation and FP64 state;
    accuracy on 4096 sampled targets against FP64 direct sums and the
    oracle's tree."""
    import gcsim

    n, eps, dt, steps = 1 << 20, 0.01, 1e-3, 3
    pos, vel, mass = gcsim.generators.plummer(n, seed=1)
    eng = gcsim.Engine("hip-tree", "mixed", softening=eps, theta=0.6, leaf_capacity=16,
                       expansion_order=4)
    eng.upload(pos, vel, mass)
    eng.compute_accelerations()
    _, _, a0 = eng.download()
    st0 = eng.fmm_stats()
    eng.synchronize()
    t0 = time.perf_counter()
    eng.step(dt, steps)
    eng.synchronize()
    wall = time.perf_counter() - t0
    st = eng.fmm_stats()
    k_ms, _ = eng.time_force_kernel(reps=2)
    eng.close()
    flops = 20.0 * st0["p2p_pairs"] + 440.0 * st0["m2p_evaluations"]
    out = {
        "workload": "tree: Plummer N=2^20, leaf<=16, p=4, theta=0.6, eps=0.01, dt=1e-3, 3 KDK",
        "ms_per_step": wall * 1e3 / steps,
        "eval_kernel_ms": k_ms,
        "build_ms": st["build_ms"], "eval_ms": st["eval_ms"],
        "p2p_pairs": st0["p2p_pairs"], "m2p_evaluations": st0["m2p_evaluations"],
        "Gp2p_per_s_eval": st0["p2p_pairs"] / (k_ms * 1e-3) / 1e9,
        "Gm2p_per_s_eval": st0["m2p_evaluations"] / (k_ms * 1e-3) / 1e9,
        "tflops_eval_20_440": flops / (k_ms * 1e-3) / 1e12,
        "stats": {k: st0[k] for k in ("node_count", "leaf_count", "max_depth")},
    }
    if args.no_cpu:  # GPU legs only (profiling runs)
        return out
    O = _oracle()
    threads = O.direct_targets_omp_threads(0)
    sample = 4096
    direct = O.direct_targets(pos, mass, 1.0, eps, 0, sample, threads=threads)
    err = _rel_err(a0[:sample], direct[:sample])
    tc = time.perf_counter()
    acc_cpu, _ = O.tree(pos, mass, 1.0, eps, 0.6, 16, 32, 4, begin=0, end=sample)
    cpu_s = time.perf_counter() - tc
    err_vs_cpu = _rel_err(a0[:sample], acc_cpu[:sample])
    out.update({
        "accuracy_vs_fp64_direct_4096": {"max": float(err.max()), "median": float(np.median(err)),
                                         "p95": float(np.percentile(err, 95))},
        "fp32_vs_oracle_tree_4096_max": float(err_vs_cpu.max()),
        "cpu_baseline": {"kind": "port", "cores": threads,
                         "sample": f"oracle tree, full build, {sample} targets",
                         "seconds": cpu_s,
                         "extrapolated_s_per_evaluation": cpu_s * n / sample},
    })
    return out


def workload_batch(args, steps=None):
    """256 collisions, N log-uniform in [512, 8192], eps 0.05, dt 1e-3: one
    persistent launch per step, hipGraph 10-step chunks, autotuned schedule;
    against the same runs stepped one at a time."""
    import gcsim
    from gcsim.batch import Batch

    runs = gcsim.generators.collisions(256, seed=2026)
    parts = [gcsim.generators.collision_particles(r, softening=0.05) for r in runs]
    counts = [len(p[2]) for p in parts]
    pairs = float(sum(c * (c - 1) for c in counts))
    data = tuple(np.concatenate([p[i] for p in parts]) for i in range(3))
    if steps is None:
        steps = max(10, args.steps // 10 * 10)
    b = Batch(counts, softening=0.05)
    b.upload(*data)
    chunk, blocks = b.autotune()
    b.compute_accelerations()
    b.step(1e-3, 10)
    b.synchronize()
    t0 = time.perf_counter()
    b.step(1e-3, steps)
    b.synchronize()
    wall = time.perf_counter() - t0
    k_ms, _ = b.time_force_kernel(5)
    b.close()
    # one at a time: every run through a resident single-run engine
    single_steps = 5
    eng = gcsim.Engine("hip-direct", "mixed", softening=0.05)
    t_single = 0.0
    for p in parts:
        eng.upload(*p)
        eng.compute_accelerations()
        eng.synchronize()
        t1 = time.perf_counter()
        eng.step(1e-3, single_steps)
        eng.synchronize()
        t_single += time.perf_counter() - t1
    eng.close()
    batched = pairs * steps / wall / 1e9
    single = pairs * single_steps / t_single / 1e9
    return {
        "workload": "batch: 256 two-disk collisions, N in [512, 8192] log-uniform, eps=0.05, dt=1e-3",
        "runs": len(counts), "particles": sum(counts), "pairs_per_step": pairs,
        "schedule": {"chunk_tiles": chunk, "blocks": blocks, "graph_steps": 10},
        "batched_Gpair_per_s": batched, "ms_per_step": wall * 1e3 / steps,
        "force_kernel_ms": k_ms,
        "force_kernel_frac_fp32_peak": FLOP_PER_PAIR * pairs / (k_ms * 1e-3) / 1e12 / FP32_PEAK_TFLOPS,
        "one_at_a_time_Gpair_per_s": single, "speedup": batched / single,
    }


def workload_adjoint(args):
    """Two-disk N = 16,384, 200 KDK steps at dt 1e-3, checkpoints every 20:
    forward + reversed-leapfrog adjoint, FP32 forces/VJP with FP64 state."""
    import torch

    import gcsim
    from gcsim import adjoint

    n, steps, dt = 16384, 200, 1e-3
    pos, vel, mass = gcsim.generators.two_disk(n // 2, softening=0.05, seed=42)
    dev = torch.device("cuda:0")
    x0, v0, m = (torch.tensor(a, dtype=torch.float64, device=dev) for a in (pos, vel, mass))
    g = torch.Generator(device=dev).manual_seed(8)
    vt = v0 + 0.005 * torch.randn(v0.shape, dtype=torch.float64, device=dev, generator=g)
    out = {"workload": "adjoint: two-disk N=16384, 200 KDK steps, dt=1e-3, checkpoints 20"}
    for prec in ("mixed", "fp64"):
        eng = gcsim.Engine("hip-direct", prec, softening=0.05)
        with torch.no_grad():  # targets from the perturbed-velocity run
            target, _ = adjoint.trajectory(x0, vt, m, dt, steps, engine=eng)
        # one untimed forward + backward first (the engine's VJP buffers and
        # step graphs are set up there), then the timed one
        for _ in range(2):
            va, ma = v0.clone().requires_grad_(), m.clone().requires_grad_()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            xT, _ = adjoint.trajectory(x0, va, ma, dt, steps, engine=eng, checkpoint_every=20)
            loss = ((xT - target) ** 2).sum(1).mean()
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            loss.backward()
            torch.cuda.synchronize()
            t2 = time.perf_counter()
        out[prec] = {"forward_s": t1 - t0, "backward_s": t2 - t1, "loss": loss.item(),
                     "grad_v_norm": va.grad.norm().item(), "grad_m_norm": ma.grad.norm().item()}
        out[prec + "_grad"] = (va.grad.detach().clone(), ma.grad.detach().clone())
        eng.close()
    gv64, gm64 = out.pop("fp64_grad")
    gv32, gm32 = out.pop("mixed_grad")
    out["mixed_vs_fp64_grad_rel"] = {
        "v": ((gv32 - gv64).norm() / gv64.norm()).item(),
        "m": ((gm32 - gm64).norm() / gm64.norm()).item()}
    # VJP kernel throughput on the two-disk N = 65,536 state (47 flop per
    # directed pair), mixed and FP64
    p2, v2, m2 = gcsim.generators.config1()
    n2 = len(m2)
    lam = torch.randn((3, n2), dtype=torch.float64, device=dev, generator=g)
    xbar = torch.empty((3, n2), dtype=torch.float64, device=dev)
    mbar = torch.empty((n2,), dtype=torch.float64, device=dev)
    for prec, reps in (("mixed", 10), ("fp64", 2)):
        eng = gcsim.Engine("hip-direct", prec, softening=0.02)
        eng.upload(p2, v2, m2)
        torch.cuda.synchronize()
        eng.force_vjp(lam.data_ptr(), xbar.data_ptr(), mbar.data_ptr())
        eng.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            eng.force_vjp(lam.data_ptr(), xbar.data_ptr(), mbar.data_ptr())
        eng.synchronize()
        sec = (time.perf_counter() - t0) / reps
        eng.close()
        rate = n2 * (n2 - 1) / sec
        out["vjp_n65536_" + prec] = {"ms": sec * 1e3, "Gpair_per_s": rate / 1e9,
                                     "tflops_47": 47.0 * rate / 1e12,
                                     "frac_fp32_peak": 47.0 * rate / 157.3e12}
    return out


def workload_drift(args):
    """Relative energy drift over 1000 KDK steps of the N = 65,536 two-disk
    collision, FP64 and mixed, and the per-step force error of mixed vs FP64."""
    import gcsim

    n, eps, dt, steps, every = 65536, 0.02, 5e-4, 1000, 50
    pos, vel, mass = gcsim.generators.config1(n=n)
    out = {"workload": "drift: two-disk N=65536, eps=0.02, dt=5e-4, 1000 KDK steps"}
    engines = {p: gcsim.Engine("hip-direct", p, softening=eps) for p in ("fp64", "mixed")}
    for e in engines.values():
        e.upload(pos, vel, mass)
        e.compute_accelerations()
    series = {p: [] for p in engines}
    errs = []
    t0 = time.perf_counter()
    for k in range(0, steps + 1, every):
        if k:
            for e in engines.values():
                e.step(dt, every)
        for p, e in engines.items():
            series[p].append(e.diagnostics()["total_energy"])
        # force error of mixed against FP64 on the mixed trajectory's state
        pm, vm, am = engines["mixed"].download()
        probe = gcsim.Engine("hip-direct", "fp64", softening=eps)
        probe.upload(pm, vm, mass)
        probe.compute_accelerations()
        a64 = probe.download()[2]
        probe.close()
        err = _rel_err(am, a64)
        errs.append((float(err.max()), float(np.median(err))))
        print(f"drift step {k}", file=sys.stderr, flush=True)
    wall = time.perf_counter() - t0
    for p, s in series.items():
        s = np.array(s)
        rel = np.abs(s - s[0]) / abs(s[0])
        out[p] = {"max_rel_energy_drift": float(rel.max()), "final_rel_energy_drift": float(rel[-1])}
    out["mixed_over_fp64_drift"] = (out["mixed"]["max_rel_energy_drift"]
                                    / max(out["fp64"]["max_rel_energy_drift"], 1e-300))
    out["mixed_force_err_vs_fp64"] = {"max_over_steps": max(e[0] for e in errs),
                                      "median_of_medians": float(np.median([e[1] for e in errs]))}
    out["wall_s"] = wall
    return out


def workload_output(args):
    """Run output from resident state: the reference's run loop
    (cpp/core/main.cpp:163-184) on the N = 65,536 two-disk collision, mixed
    direct, dt 5e-4, 200 KDK steps with snapshot_every 10 (21 snapshots and
    diagnostics rows in the reference's CSV formats), against the same 200
    steps without output.  cpu_baseline: the reference's SnapshotWriter
    compiled from its source (oracle/_ref) writing one snapshot of the same
    state."""
    import shutil
    import tempfile

    import gcsim
    from gcsim.writer import SnapshotWriter, output_steps, run_with_output

    n, eps, dt, steps, every = 65536, 0.02, 5e-4, 200, 10
    pos, vel, mass = gcsim.generators.config1(n=n)
    groups = (np.arange(n) >= n // 2).astype(np.uint32)
    root = Path(tempfile.mkdtemp(prefix="gc_output_"))
    out = {"workload": "output: two-disk N=65536 mixed direct, dt=5e-4, 200 KDK steps, "
                       "snapshot_every=10 (21 snapshots + diagnostics rows)"}
    try:
        eng = gcsim.Engine("hip-direct", "mixed", softening=eps)
        eng.upload(pos, vel, mass)
        eng.compute_accelerations()
        eng.step(dt, 20)  # warm-up
        eng.synchronize()
        t0 = time.perf_counter()
        eng.step(dt, steps)
        eng.synchronize()
        bare = time.perf_counter() - t0
        eng.upload(pos, vel, mass)
        eng.compute_accelerations()
        eng.synchronize()
        t0 = time.perf_counter()
        with SnapshotWriter(root / "async") as w:
            w.write_metadata(name="output_bench", solver="hip-direct", particle_count=n,
                             steps=steps, dt=dt, snapshot_every=every, softening=eps)
            run_with_output(eng, w, steps, dt, every, group_ids=groups)
            t_issue = time.perf_counter() - t0
        with_out = time.perf_counter() - t0
        nout = len(output_steps(steps, every))
        snap_bytes = (root / "async" / "snapshot_000010.csv").stat().st_size
        # one diagnostics evaluation, for the split of the output cost
        eng.synchronize()
        t0 = time.perf_counter()
        eng.diagnostics()
        diag_s = time.perf_counter() - t0
        # formatting alone: one snapshot from host arrays
        p, v, a = eng.download()
        t0 = time.perf_counter()
        with SnapshotWriter(root / "host") as w:
            w.write_snapshot(0, 0.0, p, v, a, mass, group_ids=groups)
        fmt_s = time.perf_counter() - t0
        eng.close()
        out.update({
            "ms_per_step_no_output": bare * 1e3 / steps,
            "ms_per_step_with_output": with_out * 1e3 / steps,
            "output_overhead_ms_per_snapshot": (with_out - bare) * 1e3 / nout,
            "loop_issue_s": t_issue, "snapshots": nout, "snapshot_bytes": snap_bytes,
            "diagnostics_ms": diag_s * 1e3,
            "host_snapshot_write_ms": fmt_s * 1e3,
            "host_snapshot_MBps": snap_bytes / fmt_s / 1e6,
        })
        O = _oracle()
        if O.ref_available():
            diag = np.zeros(13)
            t0 = time.perf_counter()
            O.ref_write_run(root / "ref", dict(name="output_bench", solver="direct", steps=steps,
                                               dt=dt, snapshot_every=every,
                                               gravitational_constant=1.0, softening=eps,
                                               tree_theta=0.6, tree_leaf_capacity=16,
                                               fmm_expansion_order=4),
                            0, 0.0, p, v, a, mass, groups, diag, snapshots=3)
            ref_s = (time.perf_counter() - t0) / 3
            same = (root / "ref" / "snapshot_000000.csv").read_bytes() == \
                (root / "host" / "snapshot_000000.csv").read_bytes()
            out["cpu_baseline"] = {"kind": "reference", "cores": 1,
                                   "sample": "SnapshotWriter::write_snapshot of the same "
                                             "N=65536 state, 3 snapshots",
                                   "ms_per_snapshot": ref_s * 1e3,
                                   "MBps": snap_bytes / ref_s / 1e6,
                                   "bytes_identical_to_ours": same}
    finally:
        shutil.rmtree(root, ignore_errors=True)
    return out


def workload_ic(args):
    """Initial conditions written in HBM by the device generator (one
    Philox4x32-10 stream per galaxy, one thread per particle) against host
    generation plus upload: the reference's disk model (two galaxies of
    the README collision) and the Plummer sphere at N = 2^20, and the
    two-disk N = 65,536 collision.  cpu_baseline: the reference's
    generate_galaxies compiled from its source (oracle/_ref) on the same
    two galaxies."""
    import gcsim

    reps, n = 5, 1 << 20
    gal = [dict(n_particles=n // 2, mass=1.0, radius=0.85, position=[-0.72, -0.10, 0.06],
                velocity=[0.34, 0.10, -0.015], orientation=0.25, group_id=0, thickness=0.045,
                inclination=0.62),
           dict(n_particles=n // 2, mass=1.0, radius=0.85, position=[0.72, 0.10, -0.06],
                velocity=[-0.34, -0.10, 0.015], orientation=3.42, group_id=1, thickness=0.045,
                inclination=-0.72)]
    out = {"workload": "ic: reference disks and Plummer N=2^20, two-disk N=65536; "
                       "device generation vs host generation + upload"}
    eng = gcsim.Engine("hip-direct", "mixed", softening=0.025)

    def timed(fn):
        fn()
        eng.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        eng.synchronize()
        return (time.perf_counter() - t0) / reps

    cases = {
        "reference_disks_2^20": (
            lambda: eng.generate_reference_galaxies(gal, 20260502),
            lambda: gcsim.generators.reference_galaxies(gal, 1.0, 0.025, 20260502)[:3], n),
        "plummer_2^20": (lambda: eng.generate_plummer(n, seed=1),
                         lambda: gcsim.generators.plummer(n, seed=1), n),
        "two_disk_65536": (lambda: eng.generate_two_disk(32768, seed=20260502),
                           lambda: gcsim.generators.config1(), 65536),
    }
    for name, (dev, host, count) in cases.items():
        dev_s = timed(dev)
        t0 = time.perf_counter()
        p, v, m = host()
        host_gen = time.perf_counter() - t0
        t0 = time.perf_counter()
        eng.upload(p, v, m)
        eng.synchronize()
        up = time.perf_counter() - t0
        out[name] = {"device_ms": dev_s * 1e3, "host_generate_ms": host_gen * 1e3,
                     "upload_ms": up * 1e3, "particles_per_s_device": count / dev_s,
                     "state_GBps_device": 80.0 * count / dev_s / 1e9,
                     "speedup_vs_host_plus_upload": (host_gen + up) / dev_s}
        print(f"ic {name} done", file=sys.stderr, flush=True)
    eng.close()
    O = _oracle()
    if O.ref_available():
        t0 = time.perf_counter()
        O.ref_generate(gal, 1.0, 0.025, 20260502)
        ref_s = time.perf_counter() - t0
        out["cpu_baseline"] = {"kind": "reference", "cores": 1,
                               "sample": "generate_galaxies, the same two galaxies (N=2^20)",
                               "ms": ref_s * 1e3, "particles_per_s": n / ref_s}
    return out


def workload_copy(args):
    """Calibration for the HBM counters: calibration_copy_kernel (plain
    float4 copy) over 256 MiB, so FETCH_SIZE / WRITE_SIZE per dispatch can
    be scaled to known bytes."""
    import ctypes as C

    import gcsim

    nbytes = 256 << 20
    ms = C.c_double()
    st = gcsim.load().gc_calibration_copy(0, nbytes, 10, C.byref(ms))
    if st != 0:
        raise gcsim.GcError(st, "calibration copy failed")
    return {"workload": "copy calibration", "kernel": "calibration_copy_kernel",
            "bytes_read_per_copy": nbytes, "bytes_written_per_copy": nbytes,
            "ms_per_copy": ms.value, "GBps_read_plus_write": 2 * nbytes / (ms.value * 1e-3) / 1e9}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=65536)
    ap.add_argument("--precision", default="mixed", choices=["mixed", "fp64"])
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 GPU comparison")
    ap.add_argument("--no-sub", action="store_true",
                    help="skip the C2 (FMM) and C4 (batch) sub-records of the default line")
    ap.add_argument("--pairs", default="auto", choices=["auto", "targets", "symmetric"],
                    help="mixed force form: per target, or each unordered pair once")
    ap.add_argument("--workload", default="direct",
                    choices=["direct", "fmm", "tree", "batch", "adjoint", "drift", "output",
                             "ic", "copy"],
                    help="direct = the headline line; the others print their own records")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: measures one GPU only (asked for {args.gpus}): the engine runs "
              "on a single device and launches no multi-GPU sharding.", file=sys.stderr)
        return 2

    import gcsim
    from gcsim import generators

    if not gcsim.solver_available():
        print("bench.py: no gfx950 device visible", file=sys.stderr)
        return 1
    if args.workload != "direct":
        fn = {"fmm": workload_fmm, "tree": workload_tree, "batch": workload_batch,
              "adjoint": workload_adjoint,
              "drift": workload_drift, "output": workload_output, "ic": workload_ic,
              "copy": workload_copy}[args.workload]
        print(json.dumps(fn(args)))
        return 0

    n, eps, dt = args.n, 0.02, 5e-4
    pos, vel, mass = generators.config1(n=n)
    pairs = float(n) * float(n - 1)

    eng = gcsim.Engine("hip-direct", args.precision, softening=eps)
    if args.precision == "mixed":
        eng.set_pairs(args.pairs)
    eng.upload(pos, vel, mass)
    eng.compute_accelerations()
    _, _, a_step0 = eng.download()

    # FP64 GPU path on the same state: accuracy of the mixed forces and the
    # FP64 kernel's rate.  Run before the timed steps (untimed), so the
    # warm-up steps start on a GPU that has left its idle clocks.
    fp64_leg = None
    if not args.no_fp64 and args.precision == "mixed":
        e64 = gcsim.Engine("hip-direct", "fp64", softening=eps)
        e64.upload(pos, vel, mass)
        e64.compute_accelerations()
        _, _, a64 = e64.download()
        f_ms, _ = e64.time_force_kernel(reps=3)
        e64.close()
        # the FP64 throughput form (rounding-level differences from the
        # bit-exact kernel above)
        ef = gcsim.Engine("hip-direct", "fp64-fast", softening=eps)
        ef.upload(pos, vel, mass)
        ef.compute_accelerations()
        _, _, af = ef.download()
        ff_ms, _ = ef.time_force_kernel(reps=10)
        ef.close()
        fp64_leg = (a64, f_ms, af, ff_ms)

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
    sources = eng.sources() if args.precision == "mixed" else None
    pair_form = eng.pairs() if args.precision == "mixed" else "targets"
    diag = eng.diagnostics()
    # host-buffer use (the reference's compute_accelerations on a host
    # particle vector): upload, one force evaluation, download, over PCIe
    eng.upload(pos, vel, mass)
    eng.compute_accelerations()
    eng.download()
    reps = 5
    t0 = time.perf_counter()
    for _ in range(reps):
        eng.upload(pos, vel, mass)
        eng.compute_accelerations()
        eng.download()
    host_ms = (time.perf_counter() - t0) * 1e3 / reps
    eng.close()

    peak = FP64_PEAK_TFLOPS if args.precision == "fp64" else FP32_PEAK_TFLOPS
    if args.precision == "fp64":
        ipt, jsplit = 1, 0
        kernel_name = "direct_fp64_kernel"
    elif pair_form == "symmetric":
        kernel_name = "direct_sym_kernel"
    else:
        kernel_name = "direct_mixed_kernel"
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
            "sources": sources, "pairs": pair_form,
            "parallelism": "single GPU",
        },
        "roofline": {
            "bound": "valu",
            "kernel": kernel_name,
            "achieved": round(achieved, 3),
            "peak": peak,
            "unit": "TFLOP/s",
            "frac": round(achieved / peak, 4),
            "traffic": _traffic(kernel_name),
            "kernel_ms": round(k_ms, 4),
            "flop_per_pair": FLOP_PER_PAIR,
        },
        "energy_after_steps": diag["total_energy"],
        "host_buffers": {
            "ms_per_evaluation": round(host_ms, 4),
            "Gpair_per_s": round(pairs / (host_ms * 1e-3) / 1e9, 3),
            "sample": f"upload of host arrays, one force evaluation and download of "
                      f"positions, velocities and accelerations, mean of {reps}",
        },
    }

    if fp64_leg is not None:
        a64, f_ms, af, ff_ms = fp64_leg
        err = np.linalg.norm(a_step0 - a64, axis=1) / np.maximum(np.linalg.norm(a64, axis=1), 1e-12)
        exact_rate = pairs / (f_ms * 1e-3) / 1e9
        fast_rate = pairs / (ff_ms * 1e-3) / 1e9
        fp64_rate = max(exact_rate, fast_rate)
        out["fp64_gpu"] = {
            "kernel_ms": round(min(f_ms, ff_ms), 4), "Gpair_per_s": round(fp64_rate, 3),
            "kernel": "direct_fp64_fast_kernel" if ff_ms < f_ms else "direct_fp64_kernel",
            "frac_fp64_peak": round(FLOP_PER_PAIR * fp64_rate / 1e3 / FP64_PEAK_TFLOPS, 4),
            "mixed_over_fp64": round((k_pairs / (k_ms * 1e-3) / 1e9) / fp64_rate, 3),
            "bit_exact_kernel": {"kernel_ms": round(f_ms, 3), "Gpair_per_s": round(exact_rate, 3),
                                 "frac_fp64_peak": round(FLOP_PER_PAIR * exact_rate / 1e3
                                                         / FP64_PEAK_TFLOPS, 4)},
            "fast_vs_bit_exact_max_rel": float(_rel_err(af, a64).max()),
        }
        out["accuracy_vs_fp64"] = {"max_rel_force_err": float(err.max()),
                                   "median_rel_force_err": float(np.median(err))}

    if not args.no_cpu and int(os.environ.get("RANK", "0")) == 0:
        sys.path.insert(0, str(ROOT / "oracle"))
        import oracle  # CPU baseline leg only

        threads = oracle.direct_targets_omp_threads(0)
        t0 = time.perf_counter()
        a_cpu = oracle.direct_targets(pos, mass, 1.0, eps, threads=threads)
        cpu_s = time.perf_counter() - t0
        # the reference's own mode: one core, on the first 2048 targets
        sample1 = min(n, 2048)
        t0 = time.perf_counter()
        oracle.direct_targets(pos, mass, 1.0, eps, 0, sample1, threads=1)
        one_s = time.perf_counter() - t0
        pairs1 = sample1 * (n - 1)
        out["cpu_baseline"] = {
            "value": round(pairs / cpu_s / 1e9, 4),
            "unit": "Gpair-interactions/s",
            "cores": threads,
            "kind": "port",
            "sample": f"one full FP64 force evaluation at N={n} (targets form, "
                      f"{pairs:.3e} pairs) on {threads} OpenMP threads, {cpu_s:.2f} s",
            "one_core": {"value": round(pairs1 / one_s / 1e9, 4), "cores": 1,
                         "sample": f"targets [0, {sample1}) against all {n} sources, "
                                   f"{pairs1:.3e} pairs, {one_s:.2f} s"},
        }
        if "accuracy_vs_fp64" in out:
            err = (np.linalg.norm(a_step0 - a_cpu, axis=1)
                   / np.maximum(np.linalg.norm(a_cpu, axis=1), 1e-12))
            out["accuracy_vs_cpu_oracle"] = {"max_rel_force_err": float(err.max()),
                                             "median_rel_force_err": float(np.median(err))}

    if not args.no_sub and args.precision == "mixed":
        # the other measured workloads, after the headline (untimed by it)
        fmm = workload_fmm(args)
        out["c2_fmm"] = {k: fmm[k] for k in (
            "workload", "ms_per_step", "build_ms", "eval_ms", "eval_kernel_ms", "p2p_pairs",
            "m2p_evaluations", "Gp2p_per_s_eval", "Gm2p_per_s_eval", "tflops_eval_20_440",
            "stats", "range_eighth", "accuracy_vs_fp64_direct_4096",
            "fp32_vs_oracle_fmm_4096_max", "cpu_baseline") if k in fmm}
        batch = workload_batch(args, steps=100)
        out["c4_batch"] = {k: batch[k] for k in (
            "workload", "runs", "particles", "pairs_per_step", "schedule", "ms_per_step",
            "batched_Gpair_per_s", "force_kernel_ms", "force_kernel_frac_fp32_peak",
            "one_at_a_time_Gpair_per_s", "speedup")}

    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
