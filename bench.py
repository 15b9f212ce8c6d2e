#!/usr/bin/env python3
"""Benchmark of the SEISMIC_CPML elastic time loop on one MI355X.

Default workload: the classic SEISMIC_CPML shape 101 x 641 x 640 (homogeneous,
cp=3300 m/s, 10-point C-PML, Ricker force source, 65 receivers, energy every 10
steps), FP32 wavefields and all 18 memory variables resident in HBM.

A "step" is one time step of that shot (stress pass + velocity pass + receivers +
source).  Timed region: K steps, bracketed by torch-free device syncs; each
kernel is bracketed by hipEvents on the shot's own stream, which gives the
dominant kernel's average duration for the roofline.  Prints ONE JSON line.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload classic]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Mcell-updates/s + % HBM roofline; mixed-shot shots/s & p99 latency vs plain loop"
PEAK_HBM = 8.0e12  # MI355X HBM3E, bytes/s


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_workload(name, nstep):
    from ela3d_amd import workloads
    if name == "classic":
        cfg, model = workloads.classic(nstep=nstep)
        desc = ("classic SEISMIC_CPML 101x641x640 homogeneous, cp=3300 m/s, dx=10 m, dt=1.6e-3 s, "
                "10-point C-PML, Ricker force f0=7 Hz, 65 receivers, energy every 10 steps")
    elif name == "layered":
        cfg, model = workloads.layered(nstep=nstep)
        desc = "seeded 6-layer anticline 401x401x301, explosive source f0=10 Hz, FP32"
    elif name == "layered_fp16":
        cfg, model = workloads.layered(nstep=nstep, fp16=True)
        desc = "seeded 6-layer anticline 401x401x301, FP16 storage with power-of-two scales"
    elif name == "homogeneous_cube":
        cfg, model = workloads.homogeneous_cube(nstep=nstep)
        desc = "homogeneous cube 101^3, Ricker force f0=7 Hz, 42 receivers, energy every step"
    else:
        raise SystemExit(f"unknown workload {name}")
    return cfg, model, desc


def cpu_baseline(cfg, model, budget_s=15.0):
    """Time the oracle's FP32 restatement of the reference loop on this host's cores."""
    import oracle
    threads = min(os.cpu_count() or 1, 16)
    sub = cfg.replace(nstep=1000000)
    st = oracle.State(sub, np.float32)
    t = time.perf_counter()
    oracle.run(sub, model, 1, st=st, dtype=np.float32, threads=threads)
    one = time.perf_counter() - t
    nsteps = int(max(2, min(400, budget_s / max(one, 1e-6))))
    t = time.perf_counter()
    oracle.run(sub, model, nsteps, st=st, it0=1, dtype=np.float32, threads=threads)
    wall = time.perf_counter() - t
    rate = cfg.cells * nsteps / wall / 1e6
    # single-thread point, the reference's own execution model, on a short sample
    t = time.perf_counter()
    oracle.run(sub, model, 1, st=st, it0=1 + nsteps, dtype=np.float32, threads=1)
    w1 = time.perf_counter() - t
    return {
        "value": round(rate, 2), "unit": "Mcell-updates/s", "cores": threads, "kind": "port",
        "sample": (f"oracle FP32 restatement of the reference loop, same grid, {nsteps} steps "
                   f"({wall:.1f} s) on {threads} threads; 1 thread: "
                   f"{cfg.cells / w1 / 1e6:.1f} Mcell-updates/s over 1 step"),
        "value_1thread": round(cfg.cells / w1 / 1e6, 2),
    }


PMC_FILES = {"classic": "pmc_classic_101x641x640_fp32.json"}


def load_traffic(workload, kernel):
    """HBM bytes per dispatch of `kernel` from the committed rocprofv3 --pmc summary
    (FETCH_SIZE and WRITE_SIZE scaled by a copy kernel of known size), with the
    launch shape it was measured at."""
    name = PMC_FILES.get(workload)
    path = os.path.join(ROOT, "profiles", name) if name else ""
    if not name or not os.path.exists(path):
        return None, None
    with open(path) as f:
        d = json.load(f)
    k = d.get("kernels", {}).get(kernel)
    return (None if k is None else k.get("hbm_bytes_per_dispatch")), d.get("launch_shape")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2500)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="classic")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--graph", action="store_true", help="also time the hipGraph replay path")
    ap.add_argument("--no-tune", action="store_true", help="skip the launch-shape autotuner")
    ap.add_argument("--calibrate", type=float, default=0.0,
                    help="only run the stream-copy kernel over this many GiB (HBM ceiling)")
    ap.add_argument("--copy-variant", type=int, default=-1,
                    help="copy kernel for --calibrate (0-3); default: all four")
    ap.add_argument("--shape", default="",
                    help="fixed launch shape block,zseg_interior,zseg_slab,vec (skips the tuner)")
    args = ap.parse_args()
    if args.calibrate > 0:
        import ela3d_amd
        nbytes = int(args.calibrate * (1 << 30))
        variants = [args.copy_variant] if args.copy_variant >= 0 else [0, 1, 2, 3]
        for v in variants:
            ms = ela3d_amd.copy_bench(nbytes, reps=20, variant=v)
            print(json.dumps({"copy_variant": v, "copy_bytes_read": nbytes,
                              "copy_bytes_written": nbytes, "ms_per_copy": ms,
                              "GBps": 2 * nbytes / (ms * 1e-3) / 1e9}), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); the time loop "
              "shards by z-slabs across GPUs, which this build does not run.", file=sys.stderr)
        sys.exit(2)

    import ela3d_amd
    from ela3d_amd import Shot
    if ela3d_amd.device_count() < 1:
        raise SystemExit("no GPU visible")
    K, W = args.steps, args.warmup
    cfg, model, desc = make_workload(args.workload, K + W)
    log(f"[bench] workload {args.workload}: {desc}; steps={K} warmup={W}")
    tuned = None
    if args.shape:
        b, zi, zs, pp = (int(x) for x in args.shape.split(","))
        ela3d_amd.set_tuning(cfg, block=b, zseg_interior=zi, zseg_slab=zs, vec=pp)
        log(f"[bench] fixed launch shape block={b} zseg={zi}/{zs} vec={pp}")
    elif not args.no_tune:
        t = time.perf_counter()
        tuned = ela3d_amd.autotune(cfg, model, steps=6)
        log(f"[bench] autotune: block={tuned['block']} zseg={tuned['zseg_interior']}/"
            f"{tuned['zseg_slab']} vec={tuned['vec']} ({tuned['ms_per_step']:.3f} ms/step over "
            f"{tuned['candidates']} shapes, {time.perf_counter() - t:.1f} s)")
    with Shot(cfg, *model) as shot:
        shot.step(W)
        shot.sync()
        info = shot.info()
        t0 = time.perf_counter()
        stats, wall_ms = shot.profile(K)
        shot.sync()
        host_s = time.perf_counter() - t0
        log(f"[bench] timed {K} steps: event wall {wall_ms:.1f} ms, host wall {host_s * 1e3:.1f} ms")
        graph_ms = None
        if args.graph:
            g_steps = min(K, 500)
            cfg_g = cfg.replace(nstep=g_steps + W)
            with Shot(cfg_g, *model) as sg:
                sg.step(W)
                sg.sync()
                t = time.perf_counter()
                sg.step(g_steps)
                sg.sync()
                graph_ms = (time.perf_counter() - t) * 1e3 / g_steps
            log(f"[bench] graph replay: {graph_ms:.4f} ms/step")
    ms_per_step = host_s * 1e3 / K
    mcells = cfg.cells * K / host_s / 1e6
    for s in stats:
        avg = s["total_ms"] / max(1, s["launches"])
        bw = s["alg_bytes_per_launch"] / (avg * 1e-3) if avg > 0 else 0.0
        log(f"[bench]   {s['name']:<20} avg {avg * 1e3:9.2f} us  alg {s['alg_bytes_per_launch'] / 1e6:9.2f} MB"
            f"  {bw / 1e9:8.1f} GB/s  ({100 * bw / PEAK_HBM:5.1f}% of 8 TB/s)")
    dom = max(stats, key=lambda s: s["total_ms"])
    avg_ms = dom["total_ms"] / dom["launches"]
    achieved = dom["alg_bytes_per_launch"] / (avg_ms * 1e-3)
    step_bw = info["alg_bytes_per_step"] / (ms_per_step * 1e-3)
    traffic, traffic_shape = load_traffic(args.workload, dom["name"])
    log(f"[bench] step: {ms_per_step:.4f} ms, {mcells:.1f} Mcell-updates/s, algorithmic "
        f"{info['alg_bytes_per_step'] / 1e6:.1f} MB/step -> {step_bw / 1e9:.1f} GB/s "
        f"({100 * step_bw / PEAK_HBM:.1f}% of peak)")
    out = {
        "metric": METRIC,
        "value": round(mcells, 2),
        "unit": "Mcell-updates/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": round(ms_per_step, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if cfg.storage == 0 else ("f16-storage/f32" if cfg.storage == 1 else "f64"),
        "data": "synthetic",
        "config": {
            "workload": desc, "nx": cfg.nx, "ny": cfg.ny, "nz": cfg.nz,
            "cells": cfg.cells, "npoints_pml": cfg.npoints_pml, "nrec": cfg.nrec,
            "energy_every": cfg.energy_every, "schedule": "two-pass (stress, velocity)",
            "alg_bytes_per_step": info["alg_bytes_per_step"],
            "step_hbm_frac": round(step_bw / PEAK_HBM, 4),
            "graph_ms_per_step": graph_ms,
            "launch_shape": {"block": info["block"], "zseg": info["zseg"],
                             "zseg_slab": info["zseg_slab"], "vec": info["vec"]},
        },
        "roofline": {
            "bound": "hbm", "kernel": dom["name"], "achieved": round(achieved / 1e9, 2),
            "peak": PEAK_HBM / 1e9, "unit": "GB/s", "frac": round(achieved / PEAK_HBM, 4),
            "traffic": traffic,
            "traffic_launch_shape": traffic_shape,
            "alg_bytes_per_launch": dom["alg_bytes_per_launch"],
            "avg_launch_ms": round(avg_ms, 5),
        },
        "cpu_baseline": None,
    }
    if not args.no_cpu:
        log("[bench] CPU baseline (oracle restatement) ...")
        out["cpu_baseline"] = cpu_baseline(cfg, model)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
