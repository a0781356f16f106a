#!/usr/bin/env python3
"""Throughput of the residual + LU-SGS iteration on one MI355X (Mcell-iter/s and HBM roofline).

A step is one NS_Time_advance_LU_SGS-equivalent iteration (residual, spectral radii and dt,
LU-SGS forward and backward sweeps, update, limit_flow, boundary conditions, halo, force_vt_kw)
over every block of the workload.  The grid and the state are resident in HBM before the timed
region starts; the timed region is exactly --steps iterations bracketed by device
synchronisation, and HIP events at the stage boundaries (on the library's own stream) give the
device time of each stage over the same iterations.

Workloads (the default is the largest single-GPU configuration, swept_wing_24block):
  naca0012_cgrid_257x97x33   single-block C-grid around NACA 0012, 257x97x33 nodes (786,432
                             cells), 2e-6 wall spacing, 20-chord far field; SA RANS, Re 6e6,
                             Mach 0.8, AoA 1.25 deg, WENO5 reconstruction, FP64
  channel_bump_97x33x9       inviscid channel with a 10% bump, 97x33x9 nodes, MUSCL3 + Roe
  naca4412_48block           577x193x129 C-grid around NACA 4412 cut 6x4x2 at seeded planes into 48
                             uneven blocks (14,155,776 cells); SST, Re 3e6, Mach 0.7, MUSCL3 + AUSMPW, CFL 0.1
  swept_wing_24block         1025x257x233 swept (30 deg), tapered (0.5) wing grid cut 4x3x2 into 24
                             blocks (60,817,408 cells); SA, Re 1.2e7, Mach 0.84, AoA 3, WENO5 + Roe

cpu_baseline: the compiled reference (oracle/_ref/ref_driver_omp, the reference's Fortran built
from its own sources with amdflang) on a bounded number of iterations of the same workload (for
the two large multi-block grids the same generator at half resolution per direction, with the
wall distance the device computed for it handed over as wall_dist.dat, which the reference reads
instead of its brute-force search), timed around its NS_Time_advance loop with every host core the
process may use; when that binary is not present, the oracle's C restatement is timed instead
(kind "port").

roofline: the largest kernel of the iteration that runs alone in its stage (so the stage's HIP
events time exactly that kernel), its algorithmic bytes per launch (the per-cell words of its pass
times the cells of the launch) over its event-timed duration.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "opencfd-ec_amd"))

from ocfd_amd import cases as C  # noqa: E402
import ocfd_amd as A  # noqa: E402

HBM_PEAK_GBS = 8000.0
# Algorithmic FP64 words per cell and iteration of the four passes (residual + dt, LU forward,
# LU backward, update): each per-cell array read once and written once, neighbours from cache.
PASS_WORDS = {5: (31, 32, 32, 15), 6: (73, 38, 38, 18), 7: (75, 41, 41, 21)}
PASS_STAGES = (("residual", ("pre", "inviscid", "viscous_dt")), ("lu_forward", ("lu_forward",)),
               ("lu_backward", ("lu_backward",)), ("update", ("update",)))
# HBM bytes per cell of each pass, measured with rocprofv3 FETCH_SIZE and WRITE_SIZE on the MI355X
# for the single-block C-grid in FP64 and scaled by a plain device copy of known size (the summary
# and its commands: profiles/iteration_kernels_naca0012_cgrid_257x97x33.md)
MEASURED_TRAFFIC = {("naca0012_cgrid_257x97x33", "fp64"): {"lu_forward": 457.8, "lu_backward": 491.8},
                    # profiles/iteration_kernels_swept_wing_24block.md (k_lusgs_team<6, true|false, 8>)
                    ("swept_wing_24block", "fp64"): {"lu_forward": 488.6, "lu_backward": 494.8}}
# stages that hold exactly one kernel launch per iteration: (kernel, the pass whose per-cell words
# are its algorithmic bytes)
SINGLE_KERNEL_STAGES = {"lu_forward": ("k_lusgs_team<{nv}, true, {tj}>", 1), "lu_backward": ("k_lusgs_team<{nv}, false, {tj}>", 2)}
FLUX_NAMES = {1: "Steger-Warming", 2: "HLL", 3: "HLLC", 4: "Roe", 5: "Van Leer", 6: "AUSMPW"}
SCHEME_NAMES = {5: "MUSCL3", 7: "WENO5"}


def build_workload(name: str, flux: int):
    if name == "naca0012_cgrid_257x97x33":
        return C.naca_cgrid(257, 97, 33, Iflag_Flux=flux)
    if name == "channel_bump_97x33x9":
        return C.channel_bump(97, 33, 9)
    if name == "naca4412_48block":
        return C.multiblock_airfoil()
    if name == "swept_wing_24block":
        return C.swept_wing()
    raise SystemExit(f"unknown workload {name}")


def cpu_sample_case(name: str, flux: int):
    """The CPU baseline's sample: the workload itself where the reference finishes setup and a few
    iterations within seconds, else the same generator at 1/8 of the cells."""
    if name == "naca4412_48block":
        return C.multiblock_airfoil(scale=0.5), "the same 48-block generator at half resolution per direction"
    if name == "swept_wing_24block":
        return C.swept_wing(scale=0.5), "the same 24-block generator at half resolution per direction"
    return build_workload(name, flux), "the benchmark workload"


def batch_member(name: str, flux: int, q: int):
    """Member q of a batched run: the workload's grid and schemes at its own freestream (AoA raised
    by 0.5 deg per member), as different requests on the same geometry would be."""
    case = build_workload(name, flux)
    if q:
        case = C.Case(case.name, case.blocks, dict(case.params, AoA=float(case.params["AoA"]) + 0.5 * q))
    return case


def variant_rate(name: str, flux: int, precision: int, batch: int, steps: int, warmup: int, graph: bool,
                 plain: bool = False):
    """Mcell-iter/s of `batch` copies of the workload stepped together (one launch per kernel over
    all their blocks), in the given precision mode; inputs resident before the timed region.  With
    `plain`, one copy in the per-block launch mode instead: every kernel launched once per block,
    eagerly (no graphs), the LU-SGS sweeps on 8x8 tiles (no autotuning)."""
    members = [A.solver_from_case(batch_member(name, flux, q), precision=precision) for q in range(batch)]
    cells = sum(ncells(batch_member(name, flux, 0)) for _ in range(batch))
    if batch == 1:
        run = members[0]
        run.set_graph(graph and not plain)
    else:
        run = A.Batch()
        run.set_graph(graph)
        for m in members:
            run.add(m)
    if plain:
        run.set_per_block(True)
        tj = 8
    else:
        tj, _ = run.autotune_lu(1)
    run.step(warmup)
    t0 = time.perf_counter()
    run.step(steps)               # returns after the device has finished
    wall = time.perf_counter() - t0
    rms = [float(m.norms()[1][0]) for m in members]
    if batch > 1:
        run.close()
    for m in members:
        m.close()
    return dict(precision="fp64" if precision == 0 else "mixed", batch=batch, cells=cells, lu_tile=f"{tj}x{64 // tj}",
                launches="per block, eager" if plain else "batched" + (", graphs" if graph else ""),
                value=round(cells * steps / wall / 1e6, 3), ms_per_step=round(wall * 1e3 / steps, 4),
                res_rms1_max=max(rms))


def ncells(case) -> int:
    return sum((b.dims[0] - 1) * (b.dims[1] - 1) * (b.dims[2] - 1) for b in case.blocks)


def host_cores() -> int:
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    lim = os.environ.get("OMP_NUM_THREADS")
    if lim and lim.isdigit():
        n = min(n, int(lim))
    return max(1, n)


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(case, what: str, iters: int):
    """Reference CPU throughput on `iters` iterations of the workload sample (rank 0, N=1 only)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    cores = host_cores()
    cells = ncells(case)
    import refrun  # noqa: E402
    if iters <= 0:
        # about 75 M cell-iterations (10-30 s of the reference on the host), 3 to 20 iterations
        iters = max(3, min(20, int(round(75e6 / cells))))
    sample = f"{iters} iterations of {what} ({cells} cells) from freestream"
    if refrun.available(omp=True):
        try:
            dw = None
            if C.nvar_of(case.params) > 5 and cells > 2_000_000:
                # the wall distance of the sample from the device, as wall_dist.dat: the reference
                # reads it instead of its O(cells x wall nodes) search
                g = A.solver_from_case(case)
                dw = g.wall_distance()
                g.close()
                sample += ", wall_dist.dat from the device's wall distance"
            r = refrun.run_case(case, nsteps=iters, kshow=10 ** 8, omp_threads=cores, dump=False, timeout=900, wall_dist=dw)
            return dict(value=cells * iters / r["loop_seconds"] / 1e6, unit="Mcell-iter/s", cores=cores,
                        kind="reference", cpu=cpu_model(),
                        sample=sample + " (reference Fortran, MPI 1 rank x OpenMP)")
        except Exception as e:  # the reference binary could not run on this host
            print(f"# reference CPU build failed to run: {e}", file=sys.stderr)
    import oracle  # noqa: E402
    os.environ.setdefault("OMP_NUM_THREADS", str(cores))
    o = oracle.Oracle(case)
    o.init_flow()
    t0 = time.perf_counter()
    o.step(iters)
    dt = time.perf_counter() - t0
    return dict(value=cells * iters / dt / 1e6, unit="Mcell-iter/s", cores=1, kind="port", cpu=cpu_model(),
                sample=sample + " (oracle C restatement, serial loops)")


def _plain_solver(case):
    """The plain GPU solver of one request: per-block eager launches, 8x8 LU-SGS tiles."""
    s = A.solver_from_case(case)
    s.set_per_block(True)
    return s


def serve_main(args) -> int:
    """The serving sweep: requests arrive (Poisson), are admitted against the HBM budget and
    advance together through shared launches until their Res_rms(1) has dropped by --drop; the
    baselines run one request at a time to the same criterion, with the solver's batched launches
    over the request's blocks and graphs, and in the plain form (per-block eager launches).
    --serve-phase splits the served run and the baselines over separate processes."""
    from ocfd_amd import serve as SV
    dims = SV.DIMS if args.serve_dims == "small" else SV.DIMS_LARGE
    reqs = SV.sweep_requests(args.serve, lam=args.rate, seed=5, dims=dims)
    log = (lambda m: print("# " + m, file=sys.stderr, flush=True)) if args.verbose else None

    def server(make_solver=A.solver_from_case, graph=bool(args.graph)):
        return SV.Server(mem_budget=args.serve_budget * 1e9, round_iters=args.round_iters, drop=args.drop,
                         max_iters=args.max_iters, graph=graph, make_solver=make_solver)

    def rounded(d):
        return {k2: (round(v, 4) if isinstance(v, float) else v) for k2, v in d.items()}

    out = {
        "metric": "sweep solves/hour with p50/p99 latency",
        "value": None,
        "unit": "solves/h",
        "n_gpus": 1,
        "higher_is_better": True,
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic requests generated in-process (NACA 4-digit C-grids, random Mach/AoA/Re)",
        "config": {"workload": f"naca_sweep_{args.serve}", "requests": args.serve, "arrival_rate_per_s": args.rate,
                   "dims": [list(d) for d in dims], "drop": args.drop, "round_iters": args.round_iters,
                   "max_iters": args.max_iters, "mem_budget_GB": args.serve_budget, "phase": args.serve_phase},
    }
    res, served_ref = [], None
    if args.serve_compare:
        # the served run of an earlier process (--serve-phase served): its forces and solves/hour
        with open(args.serve_compare) as f:
            served_ref = json.load(f)
        out["served"] = served_ref["served"]
        out["value"] = served_ref["value"]
        out["served_forces"] = served_ref["served_forces"]
    if args.serve_phase in ("both", "served"):
        srv = server()
        t0 = time.perf_counter()
        res = srv.run(reqs, log=log)
        wall = time.perf_counter() - t0
        served = SV.summarize(res, wall)
        served.update(peak_active=srv.peak_active, peak_GB=round(srv.peak_used / 1e9, 2))
        print(f"# served: {json.dumps(served)}", file=sys.stderr, flush=True)
        out["served"] = rounded(served)
        out["value"] = round(served["solves_per_hour"], 2)
        out["served_forces"] = {r.rid: [r.CL, r.CD, r.iterations] for r in res}
    if args.serve_phase in ("both", "sequential"):
        k = min(args.seq_sample, len(reqs))
        # the baseline starts each request as soon as the previous one is done (no idle gaps): its
        # solves/hour is an upper bound for the sequential solver at this arrival rate
        sub = [SV.Request(r.rid, 0.0, r.case, r.cells) for r in sorted(reqs, key=lambda r: r.arrival)[:k]]
        kinds = {"sequential": (A.solver_from_case, bool(args.graph)), "sequential_plain": (_plain_solver, False)}
        for name in args.seq_kinds.split(","):
            make, graph = kinds[name]
            t0 = time.perf_counter()
            seq = server(make, graph).run_sequential(sub, log=log)
            swall = time.perf_counter() - t0
            sq = SV.summarize(seq, swall)
            print(f"# {name}: {json.dumps(sq)}", file=sys.stderr, flush=True)
            out[name] = dict(rounded(sq), sample=f"the first {k} requests by arrival, back to back",
                             forces={r.rid: [r.CL, r.CD, r.iterations] for r in seq})
            if "served_forces" in out:
                same = {int(rid): v for rid, v in out["served_forces"].items()}
                dev = [max(abs(q.CL - same[q.rid][0]) / max(abs(q.CL), 1e-12), abs(q.CD - same[q.rid][1]) / max(abs(q.CD), 1e-12))
                       for q in seq if q.rid in same]
                out[name]["iterations_equal_served"] = all(q.iterations == same[q.rid][2] for q in seq if q.rid in same)
                out[name]["max_rel_force_diff_to_served"] = max(dev) if dev else None
                out[name]["speedup_served"] = round(out["served"]["solves_per_hour"] / max(sq["solves_per_hour"], 1e-9), 3)
    print(json.dumps(out), flush=True)
    return 0


def precision_check(args) -> int:
    """CL, CD and CM of the workload after N iterations from freestream in FP64 and in mixed
    precision (FP32 reconstruction, fluxes and SA terms), and their relative differences."""
    case = build_workload(args.workload, args.flux)
    out = {}
    for prec in (0, 1):
        s = A.solver_from_case(case, precision=prec)
        s.set_graph(bool(args.graph))
        s.step(1)
        r0 = float(s.norms()[1][0])
        done = 1
        while done < args.precision_check:
            n = min(50, args.precision_check - done)
            s.step(n)
            done += n
            print(f"# precision {prec}: {done} iterations", file=sys.stderr, flush=True)
        f = s.forces()
        out["fp64" if prec == 0 else "mixed"] = dict(CL=float(f[9]), CD=float(f[10]), CM=float(f[8]),
                                                     res_rms1_drop=float(s.norms()[1][0]) / r0)
        s.close()
    rel = {k: abs(out["mixed"][k] - out["fp64"][k]) / max(abs(out["fp64"][k]), 1e-30) for k in ("CL", "CD", "CM")}
    print(json.dumps({"workload": args.workload, "iterations": args.precision_check, **out, "relative_difference": rel}), flush=True)
    return 0


def copy_probe(mib: int, reps: int) -> int:
    """A plain device copy of known size (the calibration kernel of the HBM byte counters): run
    under rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE, each dispatch must read and write `mib` MiB."""
    import torch
    n = mib * 2 ** 20 // 8
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_MiB": mib, "reps": reps, "bytes_read_per_dispatch": n * 8,
                      "bytes_written_per_dispatch": n * 8}), flush=True)
    return 0


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=40)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="swept_wing_24block")
    ap.add_argument("--flux", type=int, default=4,
                    help="Iflag_Flux (default 4 Roe; 3 = the reference's HLLC, which diverges on "
                         "the C-grid within a few iterations)")
    ap.add_argument("--graph", type=int, default=1, help="replay each stage as a captured hipGraph")
    ap.add_argument("--lu", choices=("pipelined", "hyperplane", "fenced"), default="pipelined", help="LU-SGS sweep schedule")
    ap.add_argument("--autotune", type=int, default=1,
                    help="iterations per LU-SGS tile shape of the autotuner before the warm-up (0: no tuning)")
    ap.add_argument("--lu-tile", type=int, choices=(4, 8, 16), default=None,
                    help="fixed LU-SGS tile width instead of autotuning (profiling runs)")
    ap.add_argument("--cpu-iters", type=int, default=0,
                    help="iterations of the CPU baseline sample (0: about 75 M cell-iterations, 3 to 20)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--precision", choices=("fp64", "mixed"), default="fp64",
                    help="precision of the headline run (fp64: the drop-in; mixed: FP32 reconstruction, "
                         "fluxes and SA terms)")
    ap.add_argument("--variants", default=None,
                    help="extra runs reported beside the headline, precision:batch pairs, or precision:plain for the "
                         "per-block eager launches ('' for none; default mixed:1,fp64:4,mixed:4 on the single-block "
                         "C-grid, mixed:1,fp64:plain on the swept wing, fp64:plain,mixed:1 on the 48-block grid, none elsewhere)")
    ap.add_argument("--serve", type=int, default=0, help="run the serving sweep with this many requests instead")
    ap.add_argument("--rate", type=float, default=2.0, help="serving: request arrivals per second")
    ap.add_argument("--drop", type=float, default=5e-3, help="serving: Res_rms(1) drop that ends a solve")
    ap.add_argument("--round-iters", type=int, default=20, help="serving: iterations between admission/convergence checks")
    ap.add_argument("--max-iters", type=int, default=6000, help="serving: iteration cap per request")
    ap.add_argument("--serve-budget", type=float, default=200.0, help="serving: HBM budget of admitted requests, GB")
    ap.add_argument("--seq-sample", type=int, default=32, help="serving: requests the sequential baselines run")
    ap.add_argument("--seq-kinds", default="sequential,sequential_plain",
                    help="serving: the sequential baselines to run (sequential: batched over the request's blocks "
                         "with graphs; sequential_plain: per-block eager launches)")
    ap.add_argument("--serve-compare", default=None,
                    help="serving: JSON line of an earlier --serve-phase served run to compare the baselines with")
    ap.add_argument("--serve-phase", choices=("both", "served", "sequential"), default="both",
                    help="serving: run the served sweep, the sequential baselines, or both")
    ap.add_argument("--serve-dims", choices=("small", "large"), default="small",
                    help="serving: request grids 97x33x9/129x49x13/193x65x17 (small) or 129x49x17/193x65x25/257x97x33 (large)")
    ap.add_argument("--verbose", action="store_true")
    ap.add_argument("--precision-check", type=int, default=0, metavar="N",
                    help="only run N iterations of the workload in FP64 and in mixed precision from freestream "
                         "and print the force coefficients of both")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (byte-counter calibration)")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU: the hot path runs on a single MI355X; "
              "sharding by block over several GPUs is designed, not implemented.", file=sys.stderr)
        return 2

    if args.copy_probe:
        return copy_probe(args.copy_probe, 5)
    if args.precision_check:
        return precision_check(args)
    if args.serve:
        return serve_main(args)
    case = build_workload(args.workload, args.flux)
    cells = ncells(case)
    precision = 0 if args.precision == "fp64" else 1
    print(f"# {args.workload}: {cells} cells in {len(case.blocks)} blocks; setting up on the device", file=sys.stderr, flush=True)
    s = A.solver_from_case(case, precision=precision)   # geometry, wall distance, freestream: on the device
    nv = s.nv
    s.set_graph(bool(args.graph))
    s.set_lu_schedule({"pipelined": A.LU_PIPELINED, "hyperplane": A.LU_HYPERPLANE, "fenced": A.LU_PIPELINED_FENCED}[args.lu])
    tune = None
    if args.lu_tile:
        s.set_lu_tile(args.lu_tile)
    elif args.lu == "pipelined" and args.autotune > 0:
        tj, tms = s.autotune_lu(args.autotune)          # ordinary iterations, untimed
        tune = dict(kept=f"{tj}x{64 // tj}", ms_per_sweep_pair={f"{t}x{64 // t}": round(v, 4) for t, v in tms.items()})
        print(f"# LU tile autotune: {tune}", file=sys.stderr, flush=True)
    for _ in range(args.warmup):
        s.step(1)
    s.sync()

    t0 = time.perf_counter()
    stage_ms = s.step_timed(args.steps)   # returns after the device has finished
    s.sync()
    wall = time.perf_counter() - t0
    ms_per_step = wall * 1e3 / args.steps
    value = cells * args.steps / wall / 1e6

    s_tile = s.lu_tile() if args.lu == "pipelined" else 8
    mx, rms = s.norms()
    overlimit = max(s.status(no)[0] for no in range(1, len(case.blocks) + 1))

    # per-pass device time and algorithmic bytes
    words = PASS_WORDS[nv]
    passes = {}
    for (pname, stages), w in zip(PASS_STAGES, words):
        t = sum(stage_ms[x] for x in stages) / args.steps
        alg = w * 8.0 * cells
        passes[pname] = dict(ms=t, alg_bytes=alg, gbs=alg / (t * 1e-3) / 1e9)
    # the roofline's kernel: the largest kernel that runs alone in its stage, so the stage's
    # HIP events (on the library's stream) time exactly that kernel over the timed launches
    singles = {k: v for k, v in SINGLE_KERNEL_STAGES.items() if not (k.startswith("lu_") and args.lu != "pipelined")}
    stage_t = {k: stage_ms[k] / args.steps for k in singles}
    if stage_t:
        dom = max(stage_t, key=stage_t.get)
        kname, pidx = singles[dom]
        kernel = kname.format(nv=nv, tj=s_tile)
        t = stage_t[dom]
        alg = words[pidx] * 8.0 * cells
    else:
        dom = max(passes, key=lambda k: passes[k]["ms"])
        kernel, t, alg = dom, passes[dom]["ms"], passes[dom]["alg_bytes"]
    gbs = alg / (t * 1e-3) / 1e9
    measured = MEASURED_TRAFFIC.get((args.workload, args.precision), {}).get(dom)
    roofline = dict(bound="hbm", achieved=round(gbs, 2), peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=round(gbs / HBM_PEAK_GBS, 5),
                    traffic=round(measured * cells) if measured else None,
                    kernel=kernel, stage=dom, ms_per_launch=round(t, 4), alg_bytes_per_launch=alg)
    iter_bytes = sum(words) * 8.0 * cells
    out = {
        "metric": "Mcell-iter/s/GPU + %HBM roofline",
        "value": round(value, 3),
        "unit": "Mcell-iter/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if precision == 0 else "f32+f64",
        "data": "synthetic grid and freestream start (generated in-process)",
        "config": {
            "workload": args.workload,
            "cells": cells,
            "blocks": len(case.blocks),
            "nvar": nv,
            "scheme": SCHEME_NAMES.get(int(case.params["Iflag_Scheme"]), str(case.params["Iflag_Scheme"])),
            "flux": FLUX_NAMES[int(case.params["Iflag_Flux"])],
            "graph": bool(args.graph),
            "lu_schedule": args.lu,
            "lu_tile": (tune["kept"] if tune else f"{s_tile}x{64 // s_tile}") if args.lu != "hyperplane" else None,
            "precision": args.precision,
        },
        "lu_autotune": tune,
        "roofline": roofline,
        "iteration_hbm": {"alg_bytes_per_cell": sum(words) * 8, "achieved": round(iter_bytes / (ms_per_step * 1e-3) / 1e9, 2),
                          "frac": round(iter_bytes / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBS, 5)},
        "stage_ms_per_step": {k: round(v / args.steps, 4) for k, v in stage_ms.items()},
        "passes": {k: {"ms": round(v["ms"], 4), "GB/s": round(v["gbs"], 1)} for k, v in passes.items()},
        "state": {"res_rms1": float(rms[0]), "overlimit": int(overlimit)},
    }
    s.close()
    if args.variants is None:
        args.variants = {"naca0012_cgrid_257x97x33": "mixed:1,fp64:4,mixed:4", "swept_wing_24block": "mixed:1,fp64:plain",
                         "naca4412_48block": "fp64:plain,mixed:1"}.get(args.workload, "")
    if args.variants:
        out["variants"] = []
        for v in args.variants.split(","):
            pr, bt = v.split(":")
            plain = bt == "plain"
            r = variant_rate(args.workload, args.flux, 0 if pr == "fp64" else 1, 1 if plain else int(bt), args.steps,
                             args.warmup, bool(args.graph), plain=plain)
            r["vs_headline"] = round(r["value"] / value, 3)
            out["variants"].append(r)
    if not args.no_cpu:
        ccase, what = cpu_sample_case(args.workload, args.flux)
        out["cpu_baseline"] = cpu_baseline(ccase, what, args.cpu_iters)
        out["cpu_baseline"]["value"] = round(out["cpu_baseline"]["value"], 4)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
