"""Headline benchmark: FP64 TVD-RK3 explicit Euler step on one MI355X.

Workload (plain GPU scale): unit cube, 128^3 hexes split into 6 tetrahedra
each (12,582,912 cells), interior nodes jittered uniformly by +-15% of the
spacing (splitmix64, seed 20260101), FP64, freestream Mach 0.5 (rho=1,
p=1/gamma) plus a Gaussian density pulse (amplitude 0.1, sigma 0.1) at the
centre, all six sides farfield, CFL 0.5, Rusanov + Green-Gauss + Barth.

A step is one TVD-RK3 step (three stages) over all cells, dt computed on the
device; the timed region replays captured HIP graphs of the step loop with
all inputs resident in HBM.  Metric: cell-stage-updates/s = N * 3 * K / t.

Prints one JSON line.  --gpus N>1 is refused: this build runs on one GPU.
"""
import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
SEED = 20260101


def bytes_model(N, F):
    """Algorithmic bytes of one stage, FP64: the compulsory traffic of a
    two-sweep stage that streams face records -- sweep 1 reads U (40N),
    volumes (8N), face records (80F) and writes limited slopes (120N); sweep 2
    reads U, slopes, face records, volumes, U_old (stages 2-3) and writes U'.
    These are the bytes `roofline.achieved` is priced on; the kernels here
    recompute geometry from the nodes and move fewer (`roofline.traffic`)."""
    sweep1 = 168 * N + 80 * F
    sweep2_s1 = 208 * N + 80 * F        # stage 1: no U_old read
    sweep2_s23 = 248 * N + 80 * F
    step = 1232 * N + 480 * F
    return sweep1, sweep2_s1, sweep2_s23, step


def bytes_model_mixed(N, F):
    """The same compulsory-traffic model with FP32 work arrays and geometry
    (FP64 state): a face record is 44 B, limited slopes 60 B per cell, volumes
    stay 8 B.  Sweep 1 = 108 N + 44 F, sweep 2 = 148 N + 44 F (stage 1) or
    188 N + 44 F (with U_old); per RK3 step 860 N + 264 F."""
    sweep1 = 108 * N + 44 * F
    sweep2_s1 = 148 * N + 44 * F
    sweep2_s23 = 188 * N + 44 * F
    step = 860 * N + 264 * F
    return sweep1, sweep2_s1, sweep2_s23, step


def profiled_traffic(kernel, n, fname="sweep_traffic_128cubed.md", profiled_n=128):
    """HBM bytes per launch of `kernel` from the committed PMC summary
    (FETCH_SIZE + WRITE_SIZE, copy-calibrated) of the workload of size
    profiled_n^3 (None at other sizes)."""
    path = os.path.join(ROOT, "profiles", fname)
    if n != profiled_n or not os.path.exists(path):
        return None
    for line in open(path):
        f = line.split()
        if len(f) == 3 and f[0] == "traffic_bytes_per_launch" and f[1] == kernel:
            return int(f[2])
    return None


def copy_rate_gbs(mib=1024, reps=10):
    """Device-to-device copy rate on this GPU (read + write bytes / time), the
    practical ceiling the roofline fraction is compared with (torch's copy)."""
    import torch
    n = mib * (1 << 20) // 8
    src = torch.ones(n, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        dst.copy_(src)
    e1.record()
    torch.cuda.synchronize()
    return 2 * n * 8 * reps / (e0.elapsed_time(e1) / 1e3) / 1e9


def measured_copy_gbs():
    """copy_rate_gbs in a child process: torch brings its own HIP runtime,
    which cannot share a process with the one the library is linked to."""
    r = subprocess.run([sys.executable, os.path.abspath(__file__), "--copy-rate"], capture_output=True,
                       text=True, timeout=300)
    if r.returncode != 0:
        return None
    return json.loads(r.stdout.strip().splitlines()[-1])["copy_GBs"]


def cpu_baseline(sample_n, steps):
    """Oracle (serial restatement of the reference solver) on one host core."""
    import netflux_amd as nf
    from tests.cases import pulse_config, pulse_state
    from tests.oracle_bind import Oracle

    mesh = nf.Mesh.box(sample_n, jitter=0.15, seed=SEED, patches=("FarField",) * 6)
    cfg = pulse_config()
    ora = Oracle.from_mesh(mesh, cfg)
    ora.set_state(pulse_state(mesh))
    # one core, as the reference runs (serial); the oracle is built with the
    # reference's release flags (g++ -O3 -flto -DNDEBUG, no -march)
    cpus = sorted(os.sched_getaffinity(0))
    os.sched_setaffinity(0, {cpus[0]})
    try:
        t0 = time.perf_counter()
        ora.run(steps)
        t = time.perf_counter() - t0
    finally:
        os.sched_setaffinity(0, set(cpus))
    model = "unknown CPU"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    N = mesh.n_cells
    return {"value": N * 3 * steps / t, "unit": "cell-stage-updates/s", "cores": 1, "kind": "port",
            "sample": f"oracle, {sample_n}^3x6 = {N} cells, {steps} TVD-RK3 steps, same workload family "
                      f"(jitter 0.15, pulse, farfield), one thread pinned to core {cpus[0]} of "
                      f"{len(cpus)} available, {model}, {t:.1f} s"}


def steady_block(args):
    """The metric's second half: time to a 1e-8 density-residual drop per
    solve, learned + batched against plain FP64, on bump channels of
    20k-100k cells (netflux_amd.steady.time_to_drop).  The network is the
    committed one, trained by the library's kernels on plain-solver steady
    solutions of other cases."""
    from netflux_amd import steady as SD
    if not os.path.exists(args.steady_net):
        return {"skipped": f"no network at {args.steady_net}"}
    net, meta = SD.load_network(args.steady_net)
    t0 = time.perf_counter()
    rep = SD.time_to_drop(net, n_cases=args.steady_cases)
    print(f"[bench] steady comparison {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    rep["metric"] = "time-to-1e-8 density-residual drop per solve"
    rep["unit"] = "s per instance"
    rep["higher_is_better"] = False
    rep["network"] = {"file": os.path.relpath(args.steady_net, ROOT), "trained_on": meta}
    rep["config"] = ("bump channel [0,3]x[0,1]x[0,0.5] (Gaussian floor bump, sigma 0.2), cells log-uniform "
                     "20k-100k, Mach U[0.3,0.7], h U[0.02,0.1], farfield x-planes, slip walls, unlimited "
                     "second order, CFL 1.0, TVD-RK3, U_old = U^0; drop measured against each case's "
                     "freestream residual")
    return rep


def poly_bytes_model(N, F, G):
    """Compulsory HBM bytes of one TVD-RK3 step of the polyhedral path
    (second order, Barth) in its per-cell slot layout, with
    slots = 2 (F - G) + G cell-face pairs: per stage prim (U in, V out: 80 N),
    cell ghosts (i, j, S, type, V in, V out: 113 G), sweep A (slot words
    24 N, V and volume 48 N; per slot S and r 48 B in, face value 40 B out)
    and sweep B (slot words and neighbour slots 30 N, U, U_old, volume 88 N,
    U out 40 N; per slot S 24 B and its face value 40 B)."""
    slots = 2 * (F - G) + G
    stage = 80 * N + 113 * G + (72 * N + 88 * slots) + (158 * N + 64 * slots)
    return 3 * stage


def bench_hex(args, nf, pulse_config, pulse_state):
    """Plain FP64 TVD-RK3 steps on an n^3 hexahedral box (mixed-element
    path): nf_run's device-resident Driver loop, wall clock over K steps."""
    n = args.n if args.n != 128 else 96
    t0 = time.perf_counter()
    mesh = nf.Mesh.hex_box(n, jitter=0.15, seed=SEED, patches=("FarField",) * 6)
    U0 = pulse_state(mesh)
    cfg = pulse_config(mach=0.5)
    solver = nf.EulerSolver(cfg, mesh)
    N, F, Fint = solver.N, solver.sizes.n_total_faces, solver.sizes.n_interior_faces
    print(f"[bench] hex mesh {n}^3 = {N} cells, {F} faces, setup {time.perf_counter() - t0:.1f} s",
          file=sys.stderr, flush=True)
    solver.set_state(U0)
    # at least one full graph of graph_steps steps, so the timed region only
    # replays the captured graph (the context keeps it)
    solver.run(max(args.warmup, args.graph_steps), check_every=args.graph_steps)
    # timed region: K steps of nf_run (calc_timestep + step kept on the device,
    # graph replays of graph_steps steps); it synchronises before returning
    t0 = time.perf_counter()
    st = solver.run(args.steps, check_every=args.graph_steps)
    wall = time.perf_counter() - t0
    assert st["steps_done"] == args.steps, st
    ms_per_step = wall * 1e3 / args.steps
    G = F - Fint
    step_bytes = poly_bytes_model(N, F, G)
    # per-kernel HIP-event durations on the solver's stream
    kt = solver.time_kernels(args.timing_steps)
    kern = {name: {"avg_ms": avg, "calls": calls} for name, (avg, calls) in kt.items()}
    dom = max(kern, key=lambda k: kern[k]["avg_ms"] * kern[k]["calls"])
    slots = 2 * Fint + G
    per_launch = {"poly_sweepA": 72 * N + 88 * slots, "poly_sweepB": 158 * N + 64 * slots}.get(dom)
    roofline = None
    if per_launch:
        achieved = per_launch / (kern[dom]["avg_ms"] / 1e3) / 1e9
        roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(achieved / HBM_PEAK_GBS, 4),
                    "traffic": profiled_traffic(dom, n, "poly_kernels_hex96.md", profiled_n=96),
                    "traffic_source": "profiles/poly_kernels_hex96.md (rocprofv3 FETCH_SIZE x2 + WRITE_SIZE x1; "
                                      "an upper bound for scattered reads)",
                    "kernel": dom, "kernel_avg_ms": round(kern[dom]["avg_ms"], 4),
                    "algorithmic_bytes_per_launch": int(per_launch)}
    out = {
        "metric": "cell-stage-updates/s",
        "value": N * 3 * args.steps / wall,
        "unit": "cell-stage-updates/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_per_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (generated jittered hexahedral mesh, Gaussian pulse)",
        "config": {"workload": f"plain FP64 TVD-RK3 step, mixed-element path, unit cube {n}^3 hexahedra "
                               f"({N} cells), jitter 0.15, Mach 0.5 + pulse, farfield, CFL 0.5",
                   "cells": N, "faces": F, "graph_steps": args.graph_steps,
                   "timing": "wall clock over nf_run (device-resident Driver loop, graph replays)"},
        "step_roofline_frac": round(step_bytes / (ms_per_step / 1e3) / 1e9 / HBM_PEAK_GBS, 4),
        "step_algorithmic_bytes": step_bytes,
        "kernels": kern,
        "roofline": roofline,
        "cpu_baseline": None,
    }
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--n", type=int, default=128, help="hexes per side")
    ap.add_argument("--graph-steps", type=int, default=10)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sample-n", type=int, default=64)
    ap.add_argument("--cpu-steps", type=int, default=10)
    ap.add_argument("--no-copy-reference", dest="copy_reference", action="store_false",
                    help="skip timing a plain device copy beside the roofline")
    ap.add_argument("--copy-rate", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--timing-steps", type=int, default=3)
    ap.add_argument("--order", default="morton", choices=["input", "rcm", "morton"])
    ap.add_argument("--precision", default="fp64", choices=["fp64", "mixed"],
                    help="mixed: FP32 face arithmetic with FP64 state (not the headline)")
    ap.add_argument("--no-graph", action="store_true", help="eager launches in the timed region (profiling)")
    ap.add_argument("--autotune", type=int, default=0, metavar="STEPS",
                    help="pick the sweep launch configuration with the autotuner first")
    ap.add_argument("--steady-cases", type=int, default=8,
                    help="bump channels of the time-to-1e-8 comparison (0: skip it)")
    ap.add_argument("--steady-net", default=os.path.join(ROOT, "netflux_amd", "data", "warm_start_bump.npz"),
                    help="warm-start network for the time-to-1e-8 comparison")
    ap.add_argument("--mesh", default="tet", choices=["tet", "hex"],
                    help="hex: an n^3 hexahedral mesh through the mixed-element (polyhedral) path")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="only run a device-to-device copy of this size (PMC byte-counter calibration)")
    args = ap.parse_args()
    if args.copy_rate:
        print(json.dumps({"copy_GBs": copy_rate_gbs()}))
        return
    if args.copy_calibration_mib:
        import torch
        n = args.copy_calibration_mib * (1 << 20) // 8
        src = torch.ones(n, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes": n * 8, "copies": 3}))
        return
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(multi-GPU sharding is design-only)", file=sys.stderr)
        sys.exit(2)

    import netflux_amd as nf
    from tests.cases import pulse_config, pulse_state

    if args.mesh == "hex":
        return bench_hex(args, nf, pulse_config, pulse_state)

    t0 = time.perf_counter()
    mesh = nf.Mesh.box(args.n, jitter=0.15, seed=SEED, patches=("FarField",) * 6)
    U0 = pulse_state(mesh)
    cfg = pulse_config(mach=0.5, cell_order=args.order, precision=args.precision)
    solver = nf.EulerSolver(cfg, mesh)
    N = solver.N
    F = solver.sizes.n_total_faces
    print(f"[bench] mesh {args.n}^3x6 = {N} cells, {F} faces, setup {time.perf_counter() - t0:.1f} s",
          file=sys.stderr, flush=True)

    solver.set_state(U0)
    tune = solver.autotune(args.autotune) if args.autotune else None
    st = solver.run(args.warmup, check_every=args.graph_steps, use_graph=not args.no_graph)
    # timed region: K steps, graph replays; nf_run synchronises before return
    solver.set_state(U0)
    solver.run(1, use_graph=False)  # first step has the zero-U_old quirk; excluded
    t0 = time.perf_counter()
    st = solver.run(args.steps, check_every=args.graph_steps, use_graph=not args.no_graph)
    wall = time.perf_counter() - t0
    ms_gpu = st["elapsed_ms"]
    assert st["steps_done"] == args.steps, st
    ms_per_step = ms_gpu / args.steps
    value = N * 3 * args.steps / (ms_gpu / 1e3)
    print(f"[bench] {args.steps} steps: {ms_gpu:.1f} ms (hipEvent), {wall * 1e3:.1f} ms wall",
          file=sys.stderr, flush=True)

    # per-kernel durations with HIP events on the solver's stream
    kt = solver.time_kernels(args.timing_steps)
    mixed = args.precision == "mixed"
    s1, s2a, s2b, step_bytes = (bytes_model_mixed if mixed else bytes_model)(N, F)
    kern = {}
    for name, (avg, calls) in kt.items():
        kern[name] = {"avg_ms": avg, "calls": calls}
    tot = {k: v["avg_ms"] * v["calls"] for k, v in kern.items()}
    dom = max(tot, key=tot.get)
    sweeps = ({"sweepA_barth_mx": s1, "sweepB_mx": (s2a + 2 * s2b) / 3.0} if mixed
              else {"sweepA_barth": s1, "sweepB": (s2a + 2 * s2b) / 3.0})
    per_launch = sweeps.get(dom, s1)
    achieved = per_launch / (kern[dom]["avg_ms"] / 1e3) / 1e9
    traffic_file = "sweep_traffic_mixed_128cubed.md" if mixed else "sweep_traffic_128cubed.md"
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": profiled_traffic(dom, args.n, traffic_file),
                "traffic_source": f"profiles/{traffic_file} (rocprofv3 FETCH_SIZE x2 + WRITE_SIZE x1)",
                "kernel": dom,
                "kernel_avg_ms": round(kern[dom]["avg_ms"], 4),
                "algorithmic_bytes_per_launch": int(per_launch),
                "other_sweep": {k: {"algorithmic_bytes_per_launch": int(v),
                                    "frac": round(v / (kern[k]["avg_ms"] / 1e3) / 1e9 / HBM_PEAK_GBS, 4)}
                                for k, v in sweeps.items() if k != dom and k in kern}}
    if args.copy_reference:
        cp = measured_copy_gbs()
        if cp:
            roofline["measured_copy_GBs"] = round(cp, 1)
            roofline["frac_of_measured_copy"] = round(achieved / cp, 4)
    step_frac = step_bytes / (ms_per_step / 1e3) / 1e9 / HBM_PEAK_GBS

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(args.cpu_sample_n, args.cpu_steps)
    steady = None
    if args.steady_cases and args.precision == "fp64":
        del solver  # frees the 128^3 context first
        steady = steady_block(args)
        if cpu and "plain_fp64" in steady:
            # the reference's serial method on one host core, extrapolated: the
            # plain GPU solve's step count of each case x 3 stages x its cells /
            # the oracle's measured cell-stage-updates/s (same scheme except the
            # limiter, so an estimate)
            est = [n * 3 * c / cpu["value"] for n, c in zip(steady["plain_fp64"]["steps"], steady["cells"])]
            steady["cpu_extrapolated_per_instance_s"] = sum(est) / len(est)

    out = {
        "metric": "cell-stage-updates/s",
        "value": value,
        "unit": "cell-stage-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32 faces / f64 state",
        "data": "synthetic (generated jittered Kuhn mesh, Gaussian pulse)",
        "config": {"workload": f"plain {'FP64' if args.precision == 'fp64' else 'mixed-precision'} TVD-RK3 step, unit cube {args.n}^3 hex->6 tets "
                               f"({N} cells), jitter 0.15, Mach 0.5 + pulse, farfield, CFL 0.5",
                   "cells": N, "faces": F, "graph_steps": args.graph_steps, "cell_order": args.order,
                   "launch": tune},
        "step_roofline_frac": None if step_frac is None else round(step_frac, 4),
        "kernels": kern,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "steady": steady,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
