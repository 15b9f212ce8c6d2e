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

``--workload serving`` instead measures the shot server on the seeded stream of
256 mixed-size shots (edges 64..256, NSTEP 500..2000, Poisson arrivals at
4 shots/s): closed-loop shots/s of the batched server against one-at-a-time
execution (the plain per-step-synchronous loop and graph replay), and p50/p99
latency under the Poisson arrivals.
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


def make_workload(name, nstep=None):
    """(Config, model, description) of a named workload; nstep None keeps its own NSTEP."""
    from ela3d_amd import workloads
    kw = {} if nstep is None else {"nstep": nstep}
    if name == "classic":
        cfg, model = workloads.classic(**kw)
        desc = ("classic SEISMIC_CPML 101x641x640 homogeneous, cp=3300 m/s, dx=10 m, dt=1.6e-3 s, "
                "10-point C-PML, Ricker force f0=7 Hz, 65 receivers, energy every 10 steps")
    elif name == "layered":
        cfg, model = workloads.layered(**kw)
        desc = "seeded 6-layer anticline 401x401x301, explosive source f0=10 Hz, FP32"
    elif name == "layered_fp16":
        cfg, model = workloads.layered(fp16=True, **kw)
        desc = "seeded 6-layer anticline 401x401x301, FP16 storage with power-of-two scales"
    elif name == "homogeneous_cube":
        cfg, model = workloads.homogeneous_cube(**kw)
        desc = "homogeneous cube 101^3, Ricker force f0=7 Hz, 42 receivers, energy every step"
    else:
        raise SystemExit(f"unknown workload {name}")
    return cfg, model, desc


def cpu_baseline(cfg, model, budget_s=20.0):
    """Time the oracle's FP32 restatement of the reference loop on all of this host's
    cores (OpenMP), and on one core (the reference's own execution model)."""
    import oracle
    nproc = os.cpu_count() or 1
    avail = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else nproc
    omp = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    threads = min(avail, omp) if omp > 0 else avail  # the cores this job may use
    sub = cfg.replace(nstep=1000000)
    st = oracle.State(sub, np.float32)
    t = time.perf_counter()
    oracle.run(sub, model, 1, st=st, dtype=np.float32, threads=threads)
    one = time.perf_counter() - t
    nsteps = int(max(2, min(4000, budget_s / max(one, 1e-6))))
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
        "nproc": nproc, "cores_available": avail, "omp_num_threads": omp or None,
        "sample": (f"oracle FP32 restatement of the reference loop, same grid, {nsteps} steps "
                   f"({wall:.1f} s) on {threads} OpenMP threads (every core this job may use: "
                   f"{avail} in its affinity mask, OMP_NUM_THREADS {omp or 'unset'}; nproc {nproc}); 1 thread: "
                   f"{cfg.cells / w1 / 1e6:.1f} Mcell-updates/s over 1 step"),
        "value_1thread": round(cfg.cells / w1 / 1e6, 2),
    }


PMC_FILES = {"classic": "pmc_classic_101x641x640_fp32.json"}


# the launch-shape fields a kernel's traffic depends on
SHAPE_KEYS = {"fused": ("fused", "tile", "zseg", "strips"), "interior": ("block", "zseg", "vec", "merged"),
              "cpml": ("block_slab", "zseg_slab", "vec_slab", "merged")}


def load_traffic(workload, kernel, shape=None):
    """HBM bytes per dispatch of `kernel` from the committed rocprofv3 --pmc summary
    (FETCH_SIZE and WRITE_SIZE scaled by a copy kernel of known size), with the
    launch shape it was measured at: the measurement at `shape` where there is one
    (entries "<kernel>" and "<kernel>@<tile>x<zseg>"), else the kernel's main entry."""
    name = PMC_FILES.get(workload)
    path = os.path.join(ROOT, "profiles", name) if name else ""
    if not name or not os.path.exists(path):
        return None, None
    with open(path) as f:
        d = json.load(f)
    cands = [v for k, v in d.get("kernels", {}).items() if k == kernel or k.startswith(kernel + "@")]
    if not cands:
        return None, None

    keys = SHAPE_KEYS["fused" if kernel.startswith("fused") else "cpml" if "cpml" in kernel else "interior"]

    def same(v):
        ls = v.get("launch_shape", d.get("launch_shape")) or {}
        return shape is not None and all(ls.get(k, 0) == shape.get(k, 0) for k in keys)

    k = next((v for v in cands if same(v)), d["kernels"].get(kernel, cands[0]))
    return k.get("hbm_bytes_per_dispatch"), k.get("launch_shape", d.get("launch_shape"))


def accuracy_run(workload, nstep=None):
    """The workload (its full NSTEP, or the first `nstep` steps) in its own storage and
    in FP64 storage (the GPU FP64 path, itself pinned to the CPU FP64 oracle by the
    parity tests): seismogram relative L2 and the largest energy deviation relative to
    the FP64 energy peak."""
    from ela3d_amd import STORAGE_FP64, Shot
    cfg, model, _ = make_workload(workload, nstep)
    out = {}
    for name, c in (("run", cfg), ("fp64", cfg.replace(storage=STORAGE_FP64))):
        t = time.perf_counter()
        with Shot(c, *model) as s:
            s.run()
            out[name] = (s.seismograms(), s.energy())
        log(f"[bench] accuracy leg {name}: {c.nstep} steps in {time.perf_counter() - t:.1f} s")
    s32, e32 = out["run"]
    s64, e64 = out["fp64"]
    rel = float(np.linalg.norm((s32 - s64).ravel()) / max(np.linalg.norm(s64.ravel()), 1e-300))
    tot32, tot64 = e32.sum(1), e64.sum(1)
    edev = float(np.max(np.abs(tot32 - tot64)) / max(np.max(tot64), 1e-300)) if tot64.size else None
    log(f"[bench] accuracy: seismogram rel L2 vs FP64 {rel:.3e}, max energy deviation {edev}")
    return {"nstep": cfg.nstep, "seismogram_rel_l2_vs_fp64": rel, "max_energy_dev_vs_fp64": edev,
            "bar": 1e-4, "within_bar": bool(rel <= 1e-4),
            "reference": "same workload with FP64 storage on the GPU"}


def percentile(a, q):
    return float(np.percentile(np.asarray(a, np.float64), q)) if len(a) else None


def run_server(shots, arrivals=None, label="", **opts):
    """Submit every shot (at its arrival time if given) and wait for all.

    Returns (makespan_s, results) with per-shot status, submit/admit/done times
    relative to the start of the run."""
    import ela3d_amd
    res = [None] * len(shots)
    with ela3d_amd.Server(**opts) as srv:
        t_start = time.perf_counter()
        tickets = []
        for q, (cfg, model) in enumerate(shots):
            if arrivals is not None:
                dt = arrivals[q] - (time.perf_counter() - t_start)
                if dt > 0:
                    time.sleep(dt)
            tickets.append(srv.submit(cfg, *model))
        base = None
        for q, t in enumerate(tickets):
            info = srv.wait(t)
            srv.release(t)
            res[q] = info
            if (q + 1) % 32 == 0 or q + 1 == len(tickets):
                log(f"[serving] {label}: {q + 1}/{len(tickets)} shots done, "
                    f"{time.perf_counter() - t_start:.1f} s")
        makespan = time.perf_counter() - t_start
        stats = srv.stats()
    base = min(r["t_submit"] for r in res)
    for r in res:
        for k in ("t_submit", "t_admit", "t_done"):
            r[k] -= base
    return makespan, res, stats


def serving_summary(nshots=40, nstep_range=(100, 400), seed=11, max_active=8, steps_per_round=50,
                    budget_gb=200.0, rate=4.0, edges=None):
    """A shortened seeded serving stream with the full stream's edge mix: closed-loop
    shots/s of the batched server, of one-at-a-time graph replay and of the plain
    per-step-synchronous loop, and the batched server's p50/p99 latency under Poisson
    arrivals (open loop)."""
    import ela3d_amd
    from ela3d_amd import workloads
    edges = workloads.SERVING_EDGES if edges is None else tuple(edges)
    specs = workloads.serving_stream(nshots=nshots, seed=seed, nstep_range=nstep_range, rate=rate,
                                     edges=edges)
    shots = [sp.materialize() for sp in specs]
    for n in sorted({sp.edge for sp in specs}):
        cfg, model = [sh for sh, sp in zip(shots, specs) if sp.edge == n][0]
        ela3d_amd.autotune(cfg.replace(nstep=64), model, steps=4)
    cells = sum(sp.cell_updates for sp in specs)
    common = dict(hbm_budget_bytes=int(budget_gb * 1e9), steps_per_round=steps_per_round)
    out = {"workload": (f"{nshots} seeded shots, cubic edges {list(edges)}, "
                        f"NSTEP {nstep_range[0]}-{nstep_range[1]}, Poisson {rate} shots/s"),
           "cell_updates": cells, "max_active": max_active}
    runs = (("batched", dict(max_active=max_active)), ("graph_one_at_a_time", dict(max_active=1)),
            ("plain_loop", dict(plain_loop=True)))
    for name, opts in runs:
        mk, res, st = run_server(shots, label=name, **common, **opts)
        service = [r["t_done"] - r["t_admit"] for r in res]
        d = {"shots_per_s": round(len(res) / mk, 4), "makespan_s": round(mk, 3),
             "gcell_updates_per_s": round(cells / mk / 1e9, 3),
             "failed": sum(1 for r in res if r["status"] != 0)}
        if name != "batched":
            lat = workloads.fifo_latencies([sp.arrival for sp in specs], service)
            d["poisson_p50_s"] = round(percentile(lat, 50), 4)
            d["poisson_p99_s"] = round(percentile(lat, 99), 4)
        out[name] = d
    mk, res, st = run_server(shots, arrivals=[sp.arrival for sp in specs], label="batched Poisson",
                             max_active=max_active, **common)
    lat = [r["t_done"] - r["t_submit"] for r in res]
    out["batched"]["poisson_p50_s"] = round(percentile(lat, 50), 4)
    out["batched"]["poisson_p99_s"] = round(percentile(lat, 99), 4)
    out["latency"] = ("batched: measured open loop at the Poisson arrival times; one-at-a-time: "
                      "FIFO (Lindley) over each shot's measured service time")
    b = out["batched"]["shots_per_s"]
    out["speedup_vs_graph_one_at_a_time"] = round(b / out["graph_one_at_a_time"]["shots_per_s"], 3)
    out["speedup_vs_plain_loop"] = round(b / out["plain_loop"]["shots_per_s"], 3)
    log(f"[bench] serving: {json.dumps(out)}")
    return out


def slabs_main(args):
    """The workload as args.slabs equal z-slabs on one GPU: each step is the stress
    phase of every slab, the halo exchanges, the velocity phase, the exchanges."""
    import ela3d_amd
    from ela3d_amd import Shot
    K, W = args.steps, args.warmup
    cfg, model, desc = make_workload(args.workload, K + W)
    n = args.slabs
    cuts = [round(q * cfg.nz / n) for q in range(n + 1)]
    slabs = [Shot(cfg, *model, slab=(a, b)) for a, b in zip(cuts[:-1], cuts[1:])]

    def step():
        for phase in (0, 1):
            for s in slabs:
                s.step_phase(phase)
            for lower, upper in zip(slabs[:-1], slabs[1:]):
                lower.halo_exchange(upper, phase)

    try:
        for _ in range(W):
            step()
        for s in slabs:
            s.sync()
        t = time.perf_counter()
        for _ in range(K):
            step()
        for s in slabs:
            s.sync()
        wall = time.perf_counter() - t
    finally:
        for s in slabs:
            s.close()
    ms = wall * 1e3 / K
    plane = cfg.nx * cfg.ny
    out = {"metric": "z-slab step time on one GPU", "workload": desc, "slabs": n, "cuts": cuts,
           "steps": K, "ms_per_step": round(ms, 4),
           "mcell_updates_per_s": round(cfg.cells / (ms * 1e-3) / 1e6, 1),
           "halo_bytes_per_interface_per_step": 6 * plane * 4,
           "timing": "host wall over K steps; every phase's exchange waits for both slabs' streams"}
    log(f"[bench] slabs: {json.dumps(out)}")
    print(json.dumps(out), flush=True)


def serving_main(args):
    import ela3d_amd
    from ela3d_amd import workloads
    if ela3d_amd.device_count() < 1:
        raise SystemExit("no GPU visible")
    specs = workloads.serving_stream(nshots=args.shots, seed=args.seed)
    arrivals = [s.arrival for s in specs]
    t = time.perf_counter()
    shots = [s.materialize() for s in specs]
    log(f"[serving] {len(shots)} shots materialised in {time.perf_counter() - t:.1f} s; "
        f"{sum(s.cell_updates for s in specs) / 1e9:.1f} Gcell-updates in total")
    if not args.no_tune:  # one launch shape per grid size, chosen before the timed runs
        for n in sorted({s.edge for s in specs}):
            cfg, model = [sh for sh, sp in zip(shots, specs) if sp.edge == n][0]
            tt = ela3d_amd.autotune(cfg.replace(nstep=64), model, steps=4)
            log(f"[serving] tuned {n}^3: block={tt['block']} zseg={tt['zseg_interior']}/"
                f"{tt['zseg_slab']} vec={tt['vec']} {tt['ms_per_step']:.3f} ms/step")
    cells = sum(s.cell_updates for s in specs)
    alg = 0.0
    for cfg, _ in shots:
        alg += cfg.plan()["alg_bytes_per_step"] * cfg.nstep
    common = dict(hbm_budget_bytes=int(args.budget_gb * 1e9), steps_per_round=args.round)
    modes = {}
    # batched server, closed loop (everything submitted at once)
    mk, res, st = run_server(shots, label="batched", max_active=args.max_active, **common)
    modes["batched"] = (mk, res, st)
    # one shot at a time: graph replay, then the plain per-step-synchronous loop
    mk, res, st = run_server(shots, label="one-at-a-time graph", max_active=1, **common)
    modes["graph_one_at_a_time"] = (mk, res, st)
    if not args.no_plain:
        mk, res, st = run_server(shots, label="plain loop", plain_loop=True, **common)
        modes["plain_loop"] = (mk, res, st)
    # batched server under the Poisson arrivals (open loop)
    mk_open, res, st = run_server(shots, arrivals=arrivals, label="batched Poisson",
                                  max_active=args.max_active, **common)
    lat_open = [r["t_done"] - r["t_submit"] for r in res]
    failed_open = sum(1 for r in res if r["status"] != 0)
    report = {}
    for name, (mk, res, st) in modes.items():
        failed = sum(1 for r in res if r["status"] != 0)
        service = [r["t_done"] - r["t_admit"] for r in res]
        d = {"shots_per_s": round(len(res) / mk, 4), "makespan_s": round(mk, 3),
             "gcell_updates_per_s": round(cells / mk / 1e9, 3), "failed": failed,
             "peak_device_bytes": st["peak_bytes"]}
        if name != "batched":  # one-at-a-time: Poisson latency from the measured service times
            lat = workloads.fifo_latencies(arrivals, service)
            d["poisson_p50_s"] = round(percentile(lat, 50), 4)
            d["poisson_p99_s"] = round(percentile(lat, 99), 4)
            d["latency_model"] = "FIFO (Lindley) over measured per-shot service times"
        report[name] = d
    report["batched"]["poisson_p50_s"] = round(percentile(lat_open, 50), 4)
    report["batched"]["poisson_p99_s"] = round(percentile(lat_open, 99), 4)
    report["batched"]["poisson_makespan_s"] = round(mk_open, 3)
    report["batched"]["poisson_failed"] = failed_open
    report["batched"]["latency_model"] = "measured, open loop (submitted at the Poisson arrival times)"
    b = modes["batched"][0]
    ref = "plain_loop" if "plain_loop" in report else "graph_one_at_a_time"
    ratio = modes[ref][0] / b
    for k, v in report.items():
        log(f"[serving] {k:<22} {json.dumps(v)}")
    agg = alg / b
    out = {
        "metric": METRIC,
        "value": report["batched"]["shots_per_s"],
        "unit": "shots/s",
        "n_gpus": 1, "steps": len(shots), "warmup": 0,
        "ms_per_step": round(b * 1e3 / len(shots), 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": (f"serving stream: {len(shots)} seeded shots, cubic edges "
                                f"{list(workloads.SERVING_EDGES)}, NSTEP 500-2000, Poisson 4 shots/s"),
                   "max_active": args.max_active, "steps_per_round": args.round,
                   "hbm_budget_bytes": int(args.budget_gb * 1e9), "cell_updates": cells,
                   "speedup_vs_" + ref: round(ratio, 3), "modes": report},
        "roofline": {"bound": "hbm", "kernel": "all (aggregate over the stream)",
                     "achieved": round(agg / 1e9, 2), "peak": PEAK_HBM / 1e9, "unit": "GB/s",
                     "frac": round(agg / PEAK_HBM, 4), "traffic": None},
        "cpu_baseline": None,
    }
    if not args.no_cpu:
        log("[serving] CPU baseline (oracle restatement) on a sample shot ...")
        spec = [s for s in specs if s.edge == 128][0]
        cfg, model = shots[specs.index(spec)]
        cb = cpu_baseline(cfg, model, budget_s=10.0)
        mean_cells = cells / len(specs)
        cb["sample"] = (f"{cb['sample']}; a {spec.edge}^3 shot of the stream, shots/s = "
                        f"Mcell-updates/s / mean {mean_cells / 1e6:.0f} M cell-updates per shot")
        cb["value"] = round(cb["value"] * 1e6 / mean_cells, 5)
        cb["unit"] = "shots/s"
        out["cpu_baseline"] = cb
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2500)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="classic")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--fused-tile", type=int, default=0,
                    help="with --shape: run the fused interior step with this many cells per workgroup "
                         "(its z-segment is the shape's interior zseg)")
    ap.add_argument("--fused-strips", type=int, default=0,
                    help="with --shape and --fused-tile: xy strips of the fused step's workgroup order "
                         "(0: the default)")
    ap.add_argument("--profile-every", type=int, default=10,
                    help="within the K timed steps, every N-th step runs eagerly with hipEvents around each "
                         "launch (the per-kernel roofline); the others replay the step graphs (default 10; "
                         "1 instruments every step)")
    ap.add_argument("--no-graph", dest="graph", action="store_false",
                    help="skip timing the hipGraph replay path (the production path, reported in config)")
    ap.add_argument("--no-tune", action="store_true", help="skip the launch-shape autotuner")
    ap.add_argument("--calibrate", type=float, default=0.0,
                    help="only run the stream-copy kernel over this many GiB (HBM ceiling)")
    ap.add_argument("--copy-variant", type=int, default=-1,
                    help="copy kernel for --calibrate (0-8); default: all nine")
    ap.add_argument("--shape", default="",
                    help="fixed launch shape block,zseg,vec[,block_slab,zseg_slab,vec_slab"
                         "[,overlap,merged]] (interior, then absorbing layers; skips the tuner)")
    ap.add_argument("--slabs", type=int, default=0,
                    help="split the workload into this many z-slabs on the one GPU, step them by phases "
                         "with halo exchanges, and report the step time (one GPU's share of a z "
                         "decomposition at its real shape)")
    ap.add_argument("--shots", type=int, default=256, help="serving: shots in the stream")
    ap.add_argument("--seed", type=int, default=7, help="serving: stream seed")
    ap.add_argument("--max-active", type=int, default=8, help="serving: concurrent shots")
    ap.add_argument("--round", type=int, default=50, help="serving: steps per round")
    ap.add_argument("--budget-gb", type=float, default=200.0, help="serving: HBM budget (GB)")
    ap.add_argument("--no-plain", action="store_true", help="serving: skip the plain-loop leg")
    ap.add_argument("--accuracy", action="store_true",
                    help="run the accuracy leg over the whole NSTEP of the workload")
    ap.add_argument("--accuracy-steps", type=int, default=600,
                    help="steps of the accuracy leg (the workload in its storage and in FP64 "
                         "storage; seismogram relative L2 and largest energy deviation); 0: skip")
    ap.add_argument("--serving-shots", type=int, default=40,
                    help="shots of the shortened serving stream reported beside the main line "
                         "(NSTEP 100-400, the full stream's edge mix); 0: skip")
    args = ap.parse_args()
    if args.calibrate > 0:
        import ela3d_amd
        nbytes = int(args.calibrate * (1 << 30))
        variants = [args.copy_variant] if args.copy_variant >= 0 else list(range(9))
        arrays = {4: (12, 6), 5: (10, 3), 6: (12, 9), 7: (12, 6), 8: (12, 6)}
        for v in variants:
            nr, nw = arrays.get(v, (1, 1))
            size = nbytes if v < 4 else nbytes // 8  # multi-stream: every array this long
            ms = ela3d_amd.copy_bench(size, reps=20, variant=v)
            print(json.dumps({"copy_variant": v, "arrays_read": nr, "arrays_written": nw,
                              "bytes_per_array": size, "ms_per_copy": ms,
                              "GBps": (nr + nw) * size / (ms * 1e-3) / 1e9}), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); the time loop "
              "shards by z-slabs across GPUs, which this build does not run.", file=sys.stderr)
        sys.exit(2)

    if args.workload == "serving":
        return serving_main(args)
    if args.slabs > 0:
        return slabs_main(args)
    import ela3d_amd
    from ela3d_amd import Shot
    if ela3d_amd.device_count() < 1:
        raise SystemExit("no GPU visible")
    K, W = args.steps, args.warmup
    cfg, model, desc = make_workload(args.workload, K + W)
    log(f"[bench] workload {args.workload}: {desc}; steps={K} warmup={W}")
    tuned = None
    if args.shape:
        v = [int(x) for x in args.shape.split(",")]
        if len(v) == 3:
            v += [0, 4, 1]
        if len(v) == 6:
            v += [0, 0]
        if len(v) != 8:
            raise SystemExit("--shape takes 3, 6 or 8 integers")
        b, zi, pp, bs, zs, ps, ov, mg = v
        ela3d_amd.set_tuning(cfg, block=b, zseg_interior=zi, vec=pp, block_slab=bs, zseg_slab=zs,
                             vec_slab=ps, overlap=bool(ov), merged=bool(mg), fused=args.fused_tile > 0,
                             tile=args.fused_tile if args.fused_tile > 0 else 1024, strips=args.fused_strips)
        log(f"[bench] fixed launch shape interior block={b} zseg={zi} vec={pp}; "
            f"layers block={bs or b} zseg={zs} vec={ps}")
    elif not args.no_tune:
        t = time.perf_counter()
        tuned = ela3d_amd.autotune(cfg, model, steps=6)
        log(f"[bench] autotune: interior block={tuned['block']} zseg={tuned['zseg_interior']} "
            f"vec={tuned['vec']}; layers block={tuned['block_slab']} zseg={tuned['zseg_slab']} "
            f"vec={tuned['vec_slab']} overlap={tuned['overlap']} merged={tuned['merged']} "
            f"fused={tuned.get('fused', 0)} tile={tuned.get('tile', 0)} strips={tuned.get('strips', 0)} "
            f"({tuned['ms_per_step']:.3f} ms/step over "
            f"{tuned['candidates']} shapes, {time.perf_counter() - t:.1f} s)")
    with Shot(cfg, *model) as shot:
        shot.step(W)
        shot.sync()
        info = shot.info()
        t0 = time.perf_counter()
        stats, wall_ms = shot.profile(K, every=args.profile_every)
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
    accuracy = None
    if args.accuracy or args.accuracy_steps > 0:
        accuracy = accuracy_run(args.workload, None if args.accuracy else args.accuracy_steps)
    serving = serving_summary(nshots=args.serving_shots) if args.serving_shots > 0 else None
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
    shape_now = {"block": info["block"], "zseg": info["zseg"], "vec": info["vec"],
                 "block_slab": info["block_slab"], "zseg_slab": info["zseg_slab"], "vec_slab": info["vec_slab"],
                 "overlap": info["overlap"], "merged": info["merged"], "fused": info["fused"], "tile": info["tile"],
                 "strips": info["strips"]}
    traffic, traffic_shape = load_traffic(args.workload, dom["name"], shape_now)
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
            "energy_every": cfg.energy_every,
            "schedule": "fused interior step" if info["fused"] else "two-pass (stress, velocity)",
            "timing": (f"{K} steps: every {args.profile_every}-th eager with hipEvents around each launch "
                       "(roofline), the others hipGraph replays"),
            "alg_bytes_per_step": info["alg_bytes_per_step"],
            "step_hbm_frac": round(step_bw / PEAK_HBM, 4),
            "graph_ms_per_step": None if graph_ms is None else round(graph_ms, 5),
            "graph_mcell_updates_per_s": None if graph_ms is None else round(cfg.cells / (graph_ms * 1e-3) / 1e6, 2),
            "accuracy": accuracy,
            "serving": serving,
            "launch_shape": shape_now,
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
