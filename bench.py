#!/usr/bin/env python3
"""Benchmark: seconds per grid Poisson solve to FP64 relative residual 1e-8
at 256^3 on one MI355X.

A step is one full device-resident solve call of the library (assembly,
multigrid setup, flexible-CG iterations and the velocity / field update; the
inputs are already in HBM). Steps cycle through the 256^3 workload's four
systems: the ghost-fluid pressure projection (sphere drop over a pool,
smoothed random velocities) and the FDM magnetic potential with chi = 0.5, 2
and 8 (cosine-perturbed interface, H_ext = (0,1,0), band 1.5 dx). Magnetic
solves start cold (no warm start) and every pressure step gets pristine
velocities, so each step is a full solve.

Prints one JSON line (see the driver contract); `value` is the mean seconds
per solve of plain MG-PCG (the library's default). `learned_fcg_s` is the
same solves with the learned preconditioner (the package's trained U-Net
weights and the fall-back threshold recorded with them) over the same step
count, and `learned_over_mg_time` their ratio; `unet_roofline` times the
network alone at the first coarse level (HIP events per convolution, FLOPs
against the BF16 MFMA peak). The `kernels` object times the hot fine-level
kernels with HIP events on the library's stream, over one more pass of each
magnetic system (all cells active; pressure solves skip wavefronts of air
cells) after the timed steps; `roofline` is the one of them with the most
time. The `cpu_baseline` object times the FP64 CPU oracle (smoothed-
aggregation AMG + SPAI-0, BiCGStab for pressure, CG for magnetic, hierarchy
rebuilt per solve) on one pressure and one magnetic solve at the same size on
the host cores (median of 5 after a warm-up; also at the magnetic solve's
as-shipped 1e-6), and `checks` holds, for every system and both
preconditioners, the GPU solution's relative residual on the oracle's own
system and its L-infinity distance to the oracle's solution at 1e-11.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "simferrofluid-cuda_amd"))

METRIC = "s per Poisson solve to rel. residual 1e-8 at 256³ (learned FCG vs plain MG-PCG)"
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=8)
    ap.add_argument("--warmup", type=int, default=4)
    ap.add_argument("--n", type=int, default=256, help="cells per axis")
    ap.add_argument("--tol", type=float, default=1e-8)
    ap.add_argument("--precond", choices=["mg", "jacobi"], default="mg")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-profile", action="store_true", help="skip per-kernel event timing")
    ap.add_argument("--no-learned", action="store_true",
                    help="skip the learned-preconditioner leg (same systems, same step count)")
    ap.add_argument("--systems", default="p,m0.5,m2,m8")
    ap.add_argument("--workload", choices=["default", "spike512"], default="default",
                    help="spike512: 20 alternating pressure / magnetic solves on the 512^3 spike "
                         "interface, magnetic warm-started (prints its own JSON line)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run a plain 1 GiB device copy 5 times (counter calibration)")
    ap.add_argument("--summarize-profiles", metavar="DIR",
                    help="turn the rocprofv3 CSVs under DIR into the profiles/ summaries")
    return ap.parse_args()


# Kernels timed with HIP events in the profiled pass (sfp_profile.h ids): the
# rocprofv3 symbol, algorithmic bytes per cell (each distinct array once per
# launch, as the library counts them) and what the kernel does.
HOT_KERNELS = {
    "smooth_pair": ("k_smooth_pair_rows<true", 26.0,
                    "fine-level pairs of Chebyshev steps in one pass over whole z rows, FP32 "
                    "7-point stencil, the magnetic diagonal rebuilt from the couplings: the first "
                    "two pre-smoothing steps (reads the stored first iterate, b, 3 couplings, "
                    "writes x: 24 B/cell) and the last two post-smoothing steps (reads x, b, d, 3 "
                    "couplings, writes x: 28 B/cell)"),
    "smooth_restrict": ("k_smooth_pair<true, true>", 28.5,
                        "fine-level last pre-smoothing step and the residual restriction in one "
                        "pass, FP32: reads x, b, d, diag, 3 couplings, writes x and the coarse b"),
    "smooth": ("k_smooth_flat", 28.0,
               "fine-level single Chebyshev steps over the flat cell index (the first two "
               "pre-smoothing steps from the stored iterate, the last one, the first "
               "post-smoothing step with the coarse correction), FP32 7-point stencil: reads x, "
               "b, 3 couplings (the magnetic diagonal rebuilt from them; and d), writes x (and "
               "d)"),
    "matvec64": ("k_fcg_matvec_flat", 52.0,
                 "flexible-CG w = A z with the FP64 operator and its 3 inner products "
                 "(two cells of a z row per lane): reads z, 3 FP64 couplings (the magnetic "
                 "diagonal rebuilt from them, no activity flags), r, q_prev, writes w"),
    "update64": ("k_fcg_update", 80.0,
                 "flexible-CG update of p, q, r (FP64) with r.r, the FP32 copy b of r and the next "
                 "cycle's first iterate c0 D^-1 b; x read and written every other iteration "
                 "(88 B then 72 B per cell)"),
    "restrict": ("k_restrict_flat", 21.5, "fine-level FP32 residual summed over 2x2x2 children, "
                 "with the coarse level's first Chebyshev iterate"),
}


def summarize_profiles(d, n=256):
    """Writes profiles/<kernel>_256cube_trace.md, profiles/traffic_<kernel>_256.json
    and profiles/traffic_<kernel>_256cube.md for the kernel bench.py's roofline
    names (the one with the most event time), from the CSVs of the measurement
    commands."""
    import csv
    import statistics

    def rows(path):
        with open(path) as f:
            return list(csv.DictReader(f))

    out_dir = os.path.join(ROOT, "profiles")
    os.makedirs(out_dir, exist_ok=True)
    bench = json.load(open(os.path.join(d, "bench_full.json")))
    name = bench["roofline"]["name"]
    sym, bpc, what = HOT_KERNELS[name]
    trace = rows(os.path.join(d, "prof_trace", "trace_kernel_trace.csv"))
    sm = [r for r in trace if sym in r["Kernel_Name"]]
    grid = max(int(r["Grid_Size_X"]) * int(r["Grid_Size_Y"]) for r in sm)
    fine_all = [(int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) * 1e-3 for r in sm
                if int(r["Grid_Size_X"]) * int(r["Grid_Size_Y"]) == grid]
    # batched FCG enqueues iterations past convergence whose kernels return at
    # once (a few us): they are not launches of the stencil and are left out
    med = statistics.median(fine_all)
    fine = [t for t in fine_all if t > 0.2 * med]
    gated = len(fine_all) - len(fine)
    stats = rows(os.path.join(d, "prof_trace", "trace_kernel_stats.csv"))
    cells = n ** 3
    alg = bpc * cells
    avg = statistics.mean(fine)
    lines = [f"# {sym} at {n}^3: rocprofv3 kernel trace", "", f"{what}.", "",
             "Commands (GPU box, repo root):", "",
             "    python3 bench.py > gpurun_out/bench_full.json",
             "    rocprofv3 --kernel-trace --stats -f csv -d gpurun_out/prof_trace -o trace -- \\",
             "        python3 bench.py --no-cpu-baseline --no-profile --no-learned --steps 4 --warmup 0 --systems m2,m8",
             "    python3 bench.py --summarize-profiles gpurun_out", "",
             f"(magnetic systems: every cell active, so every launch moves the full {bpc:g} B per cell)", "",
             f"Fine-level dispatches of `{sym}` (grid {grid} work-items): {len(fine)} "
             f"({gated} more returned at once: iterations enqueued past convergence, left out)",
             f"average duration {avg:.2f} us (min {min(fine):.2f}, max {max(fine):.2f});",
             f"algorithmic bytes per launch {bpc:g} B x {cells} cells = {alg / 1e6:.1f} MB "
             f"-> {alg / (avg * 1e-6) / 1e9:.0f} GB/s = {alg / (avg * 1e-6) / 8e12:.3f} of 8 TB/s.",
             f"bench.py's HIP-event average over its profiled pass: "
             f"{bench['roofline']['avg_launch_us']:.2f} us.", "",
             "Top kernels by total time (all levels; from trace_kernel_stats.csv):", "",
             "| kernel | calls | total ms | avg us | % |", "|---|---:|---:|---:|---:|"]
    for r in stats[:14]:
        kname = r["Name"].replace("sfp::(anonymous namespace)::", "").split("(")[0]
        lines.append(f"| `{kname}` | {r['Calls']} | {float(r['TotalDurationNs']) / 1e6:.3f} | "
                     f"{float(r['AverageNs']) / 1e3:.2f} | {float(r['Percentage']):.2f} |")
    open(os.path.join(out_dir, f"{name}_{n}cube_trace.md"), "w").write("\n".join(lines) + "\n")

    def counter(sub, name):
        rs = rows(os.path.join(d, sub, [f for f in os.listdir(os.path.join(d, sub))
                                        if f.endswith("counter_collection.csv")][0]))
        return rs

    def real(vals):  # drop the dispatches that returned at once (see above)
        m = statistics.median(vals)
        return [v for v in vals if v > 0.2 * m]

    fetch = real([float(r["Counter_Value"]) for r in counter("prof_fetch", "FETCH_SIZE")
                  if sym in r["Kernel_Name"] and int(r["Grid_Size"]) == grid * 1])
    write = real([float(r["Counter_Value"]) for r in counter("prof_write", "WRITE_SIZE")
                  if sym in r["Kernel_Name"] and int(r["Grid_Size"]) == grid * 1])
    cf = [float(r["Counter_Value"]) for r in counter("prof_copy_fetch", "FETCH_SIZE")
          if "copyBuffer" in r["Kernel_Name"]]
    cw = [float(r["Counter_Value"]) for r in counter("prof_copy_write", "WRITE_SIZE")
          if "copyBuffer" in r["Kernel_Name"]]
    copy_kib = (1 << 30) / 1024.0
    fs, ws = copy_kib / statistics.mean(cf), copy_kib / statistics.mean(cw)
    rd = statistics.mean(fetch) * 1024 * fs
    wr = statistics.mean(write) * 1024 * ws
    res = {"kernel": sym, "cells": cells, "grid_work_items": grid,
           "fetch_kib_avg": statistics.mean(fetch), "write_kib_avg": statistics.mean(write),
           "fetch_scale": fs, "write_scale": ws, "hbm_bytes_read": rd, "hbm_bytes_written": wr,
           "hbm_bytes_per_launch": rd + wr, "algorithmic_bytes_per_launch": alg,
           "traffic_over_algorithmic": (rd + wr) / alg}
    json.dump(res, open(os.path.join(out_dir, f"traffic_{name}_{n}.json"), "w"), indent=1)
    md = [f"# {sym} at {n}^3: HBM traffic from FETCH_SIZE / WRITE_SIZE", "",
          "Commands (GPU box, repo root), one counter per pass:", "",
          "    python3 bench.py > gpurun_out/bench_full.json",
          "    rocprofv3 --pmc FETCH_SIZE -f csv -d gpurun_out/prof_fetch -o fetch -- \\",
          "        python3 bench.py --no-cpu-baseline --no-profile --no-learned --steps 2 --warmup 0 --systems m2,m8",
          "    rocprofv3 --pmc WRITE_SIZE -f csv -d gpurun_out/prof_write -o write -- \\",
          "        python3 bench.py --no-cpu-baseline --no-profile --no-learned --steps 2 --warmup 0 --systems m2,m8",
          "    rocprofv3 --pmc FETCH_SIZE -f csv -d gpurun_out/prof_copy_fetch -o copyf -- python3 bench.py --copy-probe",
          "    rocprofv3 --pmc WRITE_SIZE -f csv -d gpurun_out/prof_copy_write -o copyw -- python3 bench.py --copy-probe",
          "    python3 bench.py --summarize-profiles gpurun_out", "",
          "Calibration: a plain 1 GiB device copy (`__amd_rocclr_copyBuffer`) reads "
          f"{statistics.mean(cf):.0f} KiB by FETCH_SIZE and writes {statistics.mean(cw):.0f} KiB by "
          f"WRITE_SIZE for {copy_kib:.0f} KiB each way: scale factors {fs:.4f} (fetch) and {ws:.4f} (write).", "",
          f"| quantity | per launch |", "|---|---:|",
          f"| FETCH_SIZE (KiB, mean of {len(fetch)} fine-level dispatches) | {statistics.mean(fetch):.0f} |",
          f"| WRITE_SIZE (KiB, mean of {len(write)}) | {statistics.mean(write):.0f} |",
          f"| bytes read (scaled) | {rd / 1e6:.1f} MB |",
          f"| bytes written (scaled) | {wr / 1e6:.1f} MB |",
          f"| algorithmic bytes ({bpc:g} B/cell) | {alg / 1e6:.1f} MB |",
          f"| traffic / algorithmic | {(rd + wr) / alg:.3f} |"]
    open(os.path.join(out_dir, f"traffic_{name}_{n}cube.md"), "w").write("\n".join(md) + "\n")

    # every hot fine-level kernel from the same trace and counter runs
    fetch_rows = counter("prof_fetch", "FETCH_SIZE")
    write_rows = counter("prof_write", "WRITE_SIZE")
    table, per = [], {}
    for key in ("update64", "matvec64", "smooth_pair", "smooth", "restrict"):
        ksym, kbpc, _ = HOT_KERNELS[key]
        ks = [r for r in trace if ksym in r["Kernel_Name"]]
        if not ks:
            continue
        kgrid = max(int(r["Grid_Size_X"]) * int(r["Grid_Size_Y"]) for r in ks)
        kt = real([(int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) * 1e-3 for r in ks
                   if int(r["Grid_Size_X"]) * int(r["Grid_Size_Y"]) == kgrid])
        kf = real([float(r["Counter_Value"]) for r in fetch_rows
                   if ksym in r["Kernel_Name"] and int(r["Grid_Size"]) == kgrid])
        kw = real([float(r["Counter_Value"]) for r in write_rows
                   if ksym in r["Kernel_Name"] and int(r["Grid_Size"]) == kgrid])
        kalg = kbpc * cells
        kavg = statistics.mean(kt)
        krd = statistics.mean(kf) * 1024 * fs
        kwr = statistics.mean(kw) * 1024 * ws
        per[key] = {"kernel": ksym, "launches": len(kt), "avg_us": kavg, "bytes_per_cell": kbpc,
                    "algorithmic_GBps": kalg / (kavg * 1e-6) / 1e9,
                    "frac_of_8TBps": kalg / (kavg * 1e-6) / 8e12, "hbm_bytes_read": krd,
                    "hbm_bytes_written": kwr, "traffic_over_algorithmic": (krd + kwr) / kalg,
                    "hbm_GBps": (krd + kwr) / (kavg * 1e-6) / 1e9}
        p = per[key]
        table.append(f"| `{ksym}` | {kbpc:g} | {len(kt)} | {kavg:.1f} | {p['algorithmic_GBps']:.0f} | "
                     f"{p['frac_of_8TBps']:.3f} | {krd / 1e6:.1f} / {kwr / 1e6:.1f} | "
                     f"{p['traffic_over_algorithmic']:.3f} | {p['hbm_GBps']:.0f} |")
    json.dump({"cells": cells, "fetch_scale": fs, "write_scale": ws, "kernels": per},
              open(os.path.join(out_dir, f"traffic_fine_kernels_{n}.json"), "w"), indent=1)
    fk = [f"# Fine-level kernels at {n}^3: duration and HBM traffic", "",
          "Durations from the kernel trace, traffic from the FETCH_SIZE / WRITE_SIZE runs scaled by",
          f"the copy probe ({fs:.4f} / {ws:.4f}); the commands are those of",
          f"`{name}_{n}cube_trace.md` and `traffic_{name}_{n}cube.md` (magnetic chi = 2 and 8 systems,",
          "every cell active; the launches that returned at once past convergence left out).", "",
          "| kernel | B/cell | launches | avg us | algorithmic GB/s | of 8 TB/s | read / written MB | "
          "traffic / algorithmic | HBM GB/s |",
          "|---|---:|---:|---:|---:|---:|---:|---:|---:|"] + table
    open(os.path.join(out_dir, f"fine_kernels_{n}cube.md"), "w").write("\n".join(fk) + "\n")
    ld = os.path.join(d, "prof_learned", "learned_kernel_stats.csv")
    if os.path.exists(ld):
        st = rows(ld)
        total = sum(float(r["TotalDurationNs"]) for r in st)
        unet = [r for r in st if "k_conv" in r["Name"] or "k_pool" in r["Name"] or "k_up2" in r["Name"]
                or "k_head" in r["Name"] or "k_pack" in r["Name"] or "k_corr" in r["Name"]]
        lm = [f"# Learned leg at {n}^3 (network on the {n // 2}^3 first coarse level): rocprofv3 kernel trace",
              "", "Command (GPU box, repo root):", "",
              "    rocprofv3 --kernel-trace --stats -f csv -d gpurun_out/prof_learned -o learned -- \\",
              "        python3 bench.py --no-cpu-baseline --no-profile --steps 2 --warmup 1 --systems p", "",
              "The run holds both legs (plain MG-PCG, then the learned preconditioner) of 3 pressure",
              "solves each. Network kernels: "
              f"{sum(float(r['TotalDurationNs']) for r in unet) / 1e6:.3f} ms of {total / 1e6:.3f} ms.", "",
              "| kernel | calls | total ms | avg us | % |", "|---|---:|---:|---:|---:|"]
        for r in st[:24]:
            kname = r["Name"].replace("sfp::(anonymous namespace)::", "").split("(")[0]
            lm.append(f"| `{kname}` | {r['Calls']} | {float(r['TotalDurationNs']) / 1e6:.3f} | "
                      f"{float(r['AverageNs']) / 1e3:.2f} | {float(r['Percentage']):.2f} |")
        open(os.path.join(out_dir, f"learned_{n}cube_trace.md"), "w").write("\n".join(lm) + "\n")
    print(json.dumps(res))


def copy_probe():
    """A plain device-to-device copy of 2^27 doubles (1 GiB read, 1 GiB
    written), run 5 times: the known-size reference for FETCH_SIZE/WRITE_SIZE."""
    import torch
    n = 1 << 27
    src = torch.ones(n, dtype=torch.float64, device="cuda:0")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(5):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes_read": n * 8, "copy_probe_bytes_written": n * 8,
                      "launches": 5}), flush=True)


def main():
    args = parse()
    if args.summarize_profiles:
        summarize_profiles(args.summarize_profiles)
        return
    if args.copy_probe:
        copy_probe()
        return
    if args.workload == "spike512":
        spike512(args)
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print(json.dumps({"error": "the solver runs on exactly one GPU; a multi-GPU "
                                   "(x-slab) decomposition is not built",
                          "n_gpus": args.gpus}), flush=True)
        sys.exit(2)

    import numpy as np
    import torch

    import sfp
    from sfp import scenes

    if not torch.cuda.is_available():
        raise SystemExit("bench.py needs a GPU")
    dev = torch.device("cuda:0")
    n = args.n
    t0 = time.time()
    psc = scenes.pressure_scene(n, seed=0)
    systems = []
    for tag in args.systems.split(","):
        if tag == "p":
            systems.append(("pressure", None))
        else:
            systems.append((f"magnetic_chi{tag[1:]}", float(tag[1:])))
    msc = scenes.magnetic_scene(n, seed=0, chi=2.0)
    log(f"scenes generated in {time.time() - t0:.1f}s")

    def to_dev(a):
        return torch.from_numpy(np.ascontiguousarray(a)).to(dev)

    phi_p = to_dev(psc["phi"])
    frac = [to_dev(a) for a in psc["fraction"]]
    us = [to_dev(a) for a in psc["solid_velocity"]]
    vel0 = [to_dev(a) for a in psc["velocity"]]
    phi_m = to_dev(msc["phi"])
    hext = [to_dev(a) for a in msc["hext"]]
    pressure_out = torch.zeros(psc["phi"].shape, dtype=torch.float64, device=dev)
    potential_out = torch.zeros(msc["phi"].shape, dtype=torch.float64, device=dev)
    total = args.warmup + args.steps
    n_p = sum(1 for i in range(total) if systems[i % len(systems)][1] is None)
    vel_ring = [[v.clone() for v in vel0] for _ in range(max(n_p, 1))]
    torch.cuda.synchronize()

    geo = psc["geo"]
    ctx = sfp.Context(geo.n, geo.spacing, geo.origin, tol=args.tol, max_iters=200,
                      precond=sfp.PRECOND_JACOBI if args.precond == "jacobi" else sfp.PRECOND_MG)
    ptr = lambda t: t.data_ptr()
    ring_i = [0]

    def refill_ring():
        for ring in vel_ring:
            for a, b in zip(ring, vel0):
                a.copy_(b)
        ring_i[0] = 0
        torch.cuda.synchronize()

    def run_step(i, ctx=ctx):
        name, chi = systems[i % len(systems)]
        if chi is None:
            v = vel_ring[ring_i[0] % len(vel_ring)]
            ring_i[0] += 1
            st = ctx.pressure_project_device([ptr(a) for a in v], ptr(phi_p),
                                             [ptr(a) for a in frac], [ptr(a) for a in us], None,
                                             0.0, ptr(pressure_out))
        else:
            ctx.reset_warm_start()
            st = ctx.magnetic_solve_device(ptr(phi_m), [ptr(a) for a in hext], chi,
                                           1.5 * geo.spacing, ptr(potential_out))
        return name, st

    for i in range(args.warmup):
        name, st = run_step(i)
        log(f"warmup {name}: iters={st['iterations']} rel={st['rel_residual']:.2e} "
            f"setup={st['setup_ms']:.2f}ms solve={st['solve_ms']:.2f}ms")
    torch.cuda.synchronize()
    records = []
    t_start = time.perf_counter()
    for i in range(args.warmup, total):
        ts = time.perf_counter()
        name, st = run_step(i)
        st["wall_ms"] = (time.perf_counter() - ts) * 1e3
        st["system"] = name
        records.append(st)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t_start
    for st in records:
        log(f"{st['system']}: iters={st['iterations']} rel={st['rel_residual']:.2e} "
            f"wall={st['wall_ms']:.2f}ms setup={st['setup_ms']:.2f} solve={st['solve_ms']:.2f} "
            f"apply={st['apply_ms']:.2f}")
    sec_per_solve = elapsed / args.steps
    cells = geo.ncells

    roof = None
    kernels = {}
    if not args.no_profile:
        # per-kernel HIP events on the library's stream, over one more pass of
        # each system after the timed steps (event pairs around every launch
        # would perturb the timed steps and keep FCG from batching iterations)
        # Only systems whose every cell is active count: pressure solves skip
        # wavefronts of air cells, which the per-cell byte figures would
        # overcount.
        refill_ring()
        ctx.profile(True)
        prof = [i for i, (_, chi) in enumerate(systems) if chi is not None] or list(range(len(systems)))
        for i in prof:
            run_step(i)
        torch.cuda.synchronize()
        for kid, kname in sfp.KERNEL_NAMES.items():
            kp = ctx.kernel_profile(kid)
            if kp["launches"]:
                gbps = kp["bytes"] / (kp["total_ms"] * 1e-3) / 1e9
                kernels[kname] = {"symbol": HOT_KERNELS[kname][0], "launches": kp["launches"],
                                  "total_ms": round(kp["total_ms"], 3),
                                  "avg_us": round(1e3 * kp["total_ms"] / kp["launches"], 2),
                                  "bytes_per_cell": HOT_KERNELS[kname][1],
                                  "GBps": round(gbps, 1), "frac": round(gbps / HBM_PEAK_GBS, 4)}
        if kernels:
            # the dominant kernel: the most event time over the profiled pass
            name = max(kernels, key=lambda k: kernels[k]["total_ms"])
            kp = ctx.kernel_profile([k for k, v in sfp.KERNEL_NAMES.items() if v == name][0])
            sym, bpc, what = HOT_KERNELS[name]
            achieved = kp["bytes"] / (kp["total_ms"] * 1e-3) / 1e9
            traffic = None
            tf = os.path.join(ROOT, "profiles", f"traffic_{name}_{n}.json")
            if os.path.exists(tf):
                with open(tf) as f:
                    traffic = json.load(f).get("hbm_bytes_per_launch")
            roof = {"kernel": f"{sym} ({what}; {bpc:g} B/cell)", "name": name,
                    "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                    "traffic": traffic,
                    "avg_launch_us": round(1e3 * kp["total_ms"] / kp["launches"], 2),
                    "algorithmic_bytes_per_launch": kp["bytes"] / kp["launches"],
                    "share_of_profiled_kernel_time": round(
                        kp["total_ms"] / sum(v["total_ms"] for v in kernels.values()), 3)}
        ctx.profile(False)

    by_sys = {}
    for st in records:
        d = by_sys.setdefault(st["system"], {"wall_ms": [], "iterations": [], "rel_residual": []})
        d["wall_ms"].append(round(st["wall_ms"], 3))
        d["iterations"].append(st["iterations"])
        d["rel_residual"].append(st["rel_residual"])

    # the solution of every system, for the checks against the oracle
    solutions = {"mg": solve_each(ctx, systems, refill_ring, run_step, pressure_out, potential_out)}

    learned = None
    if not args.no_learned:
        learned = learned_leg(args, systems, total, run_step, refill_ring, geo, by_sys, log,
                              solutions, pressure_out, potential_out)

    cpu = None
    checks = {}
    if not args.no_cpu_baseline:
        cpu, checks = cpu_leg(psc, msc, systems, solutions, geo, args, log)

    out = {
        "metric": METRIC,
        "value": round(sec_per_solve, 6),
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * sec_per_solve, 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 V-cycle / f64 outer FCG",
        "data": "synthetic",
        "config": {"workload": f"{n}^3 pressure projection + FDM magnetic potential "
                               f"(chi 0.5/2/8) solves to FP64 rel. residual {args.tol:g}",
                   "grid": [n, n, n], "precond": args.precond, "systems": args.systems,
                   "learned_fcg": None if learned is None else learned["weights"],
                   "parallelism": "single GPU"},
        "mg_pcg_s": round(sec_per_solve, 6),
        "learned_fcg_s": None if learned is None else learned["s_per_solve"],
        "learned_over_mg_time": None if learned is None else round(learned["s_per_solve"] / sec_per_solve, 3),
        "learned": learned,
        "cells_per_s": round(cells / sec_per_solve, 1),
        "per_system": by_sys,
        "kernels": kernels,
        "roofline": roof,
        "unet_roofline": None if learned is None else learned["unet_roofline"],
        "cpu_baseline": cpu,
        "checks": checks,
    }
    ctx.close()
    print(json.dumps(out), flush=True)


def spike512(args):
    """20 consecutive solves alternating pressure projection and magnetic
    potential on the spike-pattern interface (y = 0.2 - 64 cosine modes of
    amplitude <= 0.005, chi = 4, H_ext = (0, 1.5, 0)) at 512^3. The interface
    moves between pairs of solves (every mode's phase advances by 0.1 rad), as
    in a running simulation; the magnetic solves are warm-started from the
    previous one (as the reference does) and every pressure solve starts from
    pristine smoothed random velocities."""
    import numpy as np
    import torch

    import sfp
    from sfp import scenes

    n = 512 if args.n == 256 else args.n
    dev = torch.device("cuda:0")
    t0 = time.time()
    msc = scenes.spike_scene(n, seed=0, chi=4.0)
    geo = msc["geo"]
    vel0 = []
    for a in range(3):
        shp = geo.face_shape(a)
        u = scenes.uniform(0, scenes.ID_VEL + a, int(np.prod(shp))).reshape(shp) * 2.0 - 1.0
        vel0.append(torch.from_numpy(scenes._box_smooth(u, 3)).to(dev))
        del u
        log(f"spike512: velocity {a} smoothed ({time.time() - t0:.1f}s)")
    frac = [torch.from_numpy(f).to(dev) for f in scenes.wall_fractions(geo)]
    us = [torch.zeros(geo.face_shape(a), dtype=torch.float64, device=dev) for a in range(3)]
    modes = scenes._modes(0, 64, 0.005, mmax=12)
    phi = torch.from_numpy(msc["phi"]).to(dev)
    hext = [torch.from_numpy(h).to(dev) for h in msc["hext"]]
    vel = [v.clone() for v in vel0]
    out = torch.zeros(geo.n, dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    log(f"spike512: scene {n}^3 ready in {time.time() - t0:.1f}s")
    ptr = lambda ts: [t.data_ptr() for t in ts]  # noqa: E731
    ctx = sfp.Context(geo.n, geo.spacing, geo.origin, tol=args.tol, max_iters=200)
    recs = []
    t_all = 0.0
    for step in range(20):
        if step % 2 == 0 and step > 0:
            a, mm, nn, th = modes
            phi.copy_(torch.from_numpy(scenes._wavy_phi(geo, 0.2, (a, mm, nn, th + 0.1 * (step // 2)))))
            torch.cuda.synchronize()
        if step % 2 == 0:
            for v, v0 in zip(vel, vel0):
                v.copy_(v0)
            torch.cuda.synchronize()
            ts = time.perf_counter()
            st = ctx.pressure_project_device(ptr(vel), phi.data_ptr(), ptr(frac), ptr(us), None, 0.0,
                                             out.data_ptr())
        else:
            ts = time.perf_counter()
            st = ctx.magnetic_solve_device(phi.data_ptr(), ptr(hext), 4.0, 1.5 * geo.spacing,
                                           out.data_ptr())
        dt = time.perf_counter() - ts
        t_all += dt
        kind = "pressure" if step % 2 == 0 else "magnetic"
        recs.append({"system": kind, "ms": round(dt * 1e3, 2), "iterations": st["iterations"],
                     "rel_residual": st["rel_residual"], "converged": bool(st["converged"])})
        log(f"spike512 step {step} {kind}: {dt * 1e3:.1f} ms, {st['iterations']} its, "
            f"rel {st['rel_residual']:.2e}")
    ctx.close()
    res = {"workload": f"{n}^3 spike interface moving between solve pairs, 20 alternating "
                       f"pressure/magnetic solves, magnetic warm-started, FP64 rel. residual {args.tol:g}",
           "s_per_solve": round(t_all / 20, 5), "cells": geo.ncells,
           "cells_per_s": round(geo.ncells / (t_all / 20), 1),
           "all_converged": all(r["converged"] for r in recs), "solves": recs}
    print(json.dumps(res), flush=True)


def solve_each(ctx, systems, refill_ring, run_step, pressure_out, potential_out):
    """One more solve of each system on `ctx` (pristine velocities, cold
    magnetic solves): {system: (host solution N^3, stats)}."""
    import torch
    out = {}
    for i, (name, chi) in enumerate(systems):
        refill_ring()
        _, st = run_step(i, ctx)
        torch.cuda.synchronize()
        out[name] = ((pressure_out if chi is None else potential_out).cpu().numpy().copy(), st)
    return out


def unet_roofline(params, n, log):
    """The network of the learned cycle alone on one n^3 volume (the first
    coarse level of the 2n^3 solve), timed with HIP events per convolution:
    FLOPs per convolution = 2 * 27 * Cin * Cout * voxels of its level."""
    import numpy as np

    import sfp
    cin = [8, 16, 16, 32, 32, 64, 64, 32, 32, 16]    # as the network stores them: the
    cout = [16, 16, 32, 32, 64, 64, 32, 32, 16, 16]  # input convolution has 8 (4 used)
    level = [0, 0, 1, 1, 2, 2, 1, 1, 0, 0]
    with sfp.UNet(1, n, n, n, training=False) as net:
        net.set_params(params)
        x = np.random.default_rng(0).standard_normal((1, n, n, n, 4)).astype(np.float32)
        net.forward(x)
        fwd_ms, conv_ms = net.profile_forward(10)
    flops = [2.0 * 27 * cin[k] * cout[k] * (n >> level[k]) ** 3 for k in range(10)]
    k = int(np.argmax(conv_ms))
    rate = flops[k] / (conv_ms[k] * 1e-3) / 1e12
    total = sum(flops) / (fwd_ms * 1e-3) / 1e12
    log(f"network {n}^3: {fwd_ms:.3f} ms per pass ({total:.0f} TFLOP/s over its convolutions' FLOPs); "
        f"dominant conv {k} ({cin[k]}->{cout[k]}): {conv_ms[k] * 1e3:.1f} us, {rate:.0f} TFLOP/s")
    return {"kernel": f"k_conv3r layer {k} ({cin[k]}->{cout[k]} channels, 3x3x3, BF16 MFMA, FP32 "
                      f"accumulation) at {n >> level[k]}^3", "bound": "mfma",
            "achieved": round(rate, 1), "peak": 2500.0, "unit": "TFLOP/s",
            "frac": round(rate / 2500.0, 4), "flops_per_launch": flops[k],
            "avg_launch_us": round(conv_ms[k] * 1e3, 2),
            "network_ms_per_pass": round(fwd_ms, 4),
            "network_tflops": round(total, 1),
            "conv_us": [round(v * 1e3, 2) for v in conv_ms]}


def learned_leg(args, systems, total, run_step, refill_ring, geo, mg_by_sys, log, solutions,
                pressure_out, potential_out):
    """The same systems and step count with the learned preconditioner (the
    package's trained weights) in a second context."""
    import torch

    import sfp
    from sfp import learn

    params, meta = learn.load_weights()
    lctx = sfp.Context(geo.n, geo.spacing, geo.origin, tol=args.tol, max_iters=200)
    lctx.set_network(params)
    lctx.set_params(**learn.learned_params(meta))
    refill_ring()
    for i in range(args.warmup):
        name, st = run_step(i, lctx)
        log(f"learned warmup {name}: iters={st['iterations']} fallback={st['fallback']} "
            f"rel={st['rel_residual']:.2e} solve={st['solve_ms']:.2f}ms")
    torch.cuda.synchronize()
    records = []
    t_start = time.perf_counter()
    for i in range(args.warmup, total):
        name, st = run_step(i, lctx)
        st["system"] = name
        records.append(st)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t_start
    solutions["learned"] = solve_each(lctx, systems, refill_ring, run_step, pressure_out, potential_out)
    lctx.close()
    roof = unet_roofline(params, geo.n[0] // 2, log)
    by_sys = {}
    for st in records:
        log(f"learned {st['system']}: iters={st['iterations']} fallback={st['fallback']} "
            f"rel={st['rel_residual']:.2e} solve={st['solve_ms']:.2f}ms")
        d = by_sys.setdefault(st["system"], {"solve_ms": [], "iterations": [], "fallback": [],
                                             "rel_residual": [], "mg_iterations": None})
        d["solve_ms"].append(round(st["solve_ms"], 3))
        d["iterations"].append(st["iterations"])
        d["fallback"].append(st["fallback"])
        d["rel_residual"].append(st["rel_residual"])
        d["mg_iterations"] = mg_by_sys.get(st["system"], {}).get("iterations")
    return {"s_per_solve": round(elapsed / args.steps, 6), "per_system": by_sys,
            "converged_all": all(st["converged"] for st in records), "unet_roofline": roof,
            "weights": {"sha256": meta["sha256"], "params": meta["count"],
                        "trained_at": meta.get("n"), "steps": meta.get("steps"),
                        "smoother": learn.smoother(meta),
                        "stall_ratio": learn.learned_params(meta)["stall_ratio"]}}


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_leg(psc, msc, systems, solutions, geo, args, log):
    """The CPU baseline (the oracle's FP64 assembly + SA-AMG/SPAI-0 setup +
    BiCGStab for pressure, CG for magnetic, all rebuilt per call as the
    reference does), and the FP64 checks of every GPU solution on the
    oracle's own systems: relative residual and L-infinity distance to the
    oracle's solution at rel. residual 1e-11."""
    import statistics

    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle

    def pressure_call(tol):
        t = time.time()
        ps = oracle.pressure_system(psc)
        ta = time.time() - t
        x, it, rel, ts, tv = oracle.solve(ps["A"], ps["rhs"], "bicgstab", "amg", tol=tol, maxit=300)
        return ps, x, it, ta, ts, tv

    def magnetic_call(chi, tol):
        t = time.time()
        ms = oracle.magnetic_system(dict(msc, chi=chi, band=1.5 * geo.spacing))
        ta = time.time() - t
        x, it, rel, ts, tv = oracle.solve(ms["A"], ms["rhs"], "cg", "amg", tol=tol, maxit=300)
        return ms, x, it, ta, ts, tv

    # timing: one warm-up, then the median of 5 calls of each
    runs = {"pressure_1e-8": lambda: pressure_call(args.tol),
            "magnetic_chi2_1e-8": lambda: magnetic_call(2.0, args.tol),
            "magnetic_chi2_1e-6": lambda: magnetic_call(2.0, 1e-6)}
    for f in runs.values():
        f()
    timing = {}
    for name, f in runs.items():
        rec = []
        for _ in range(5):
            _, _, it, ta, ts, tv = f()
            rec.append((ta + ts + tv, ta, ts, tv, it))
        rec.sort()
        tot, ta, ts, tv, it = rec[2]
        timing[name] = {"total_s": round(tot, 4), "assembly_s": round(ta, 4), "setup_s": round(ts, 4),
                        "solve_s": round(tv, 4), "iterations": it,
                        "all_totals_s": [round(r[0], 4) for r in rec]}
        log(f"cpu {name}: median {tot:.2f}s (assembly {ta:.2f} setup {ts:.2f} solve {tv:.2f}, {it} its)")

    # checks of every GPU solution against the oracle's system and solution
    checks = {}
    for name, chi in systems:
        if chi is None:
            sysd, xr, it, _, _, _ = pressure_call(1e-11)
            pick = lambda a: a.ravel()[sysd["mat2grid"]]  # noqa: E731
        else:
            sysd, xr, it, _, _, _ = magnetic_call(chi, 1e-11)
            pick = lambda a: a.ravel()  # noqa: E731
        scale = float(np.abs(xr).max())
        for leg, sol in solutions.items():
            x, st = sol[name]
            xg = pick(x)
            checks[f"{leg}_{name}"] = {
                "oracle_rel_residual": oracle.rel_residual(sysd["A"], sysd["rhs"], xg),
                "linf_over_linf_oracle": float(np.abs(xg - xr).max() / scale),
                "iterations": st["iterations"], "fallback": st["fallback"],
                "library_rel_residual": st["rel_residual"]}
            c = checks[f"{leg}_{name}"]
            log(f"check {leg} {name}: oracle rel {c['oracle_rel_residual']:.2e}, "
                f"Linf/Linf {c['linf_over_linf_oracle']:.2e}")
        del sysd
    per_solve = (timing["pressure_1e-8"]["total_s"] + timing["magnetic_chi2_1e-8"]["total_s"]) / 2
    cpu = {"value": round(per_solve, 4), "unit": "s", "cores": oracle.num_threads(), "kind": "port",
           "cpu_model": cpu_model(),
           "sample": f"one pressure (BiCGStab) and one magnetic chi=2 (CG) solve at {args.n}^3 to "
                     f"rel. residual {args.tol:g}, each the median of 5 calls after a warm-up: "
                     f"CSR assembly + SA-AMG/SPAI-0 setup + solve per call (the reference rebuilds "
                     f"its solver each call); OpenMP threads as cores",
           "timing": timing}
    return cpu, checks


if __name__ == "__main__":
    main()
