#!/usr/bin/env python3
"""Throughput of the residual + LU-SGS iteration on one MI355X (Mcell-iter/s and HBM roofline).

A step is one NS_Time_advance_LU_SGS-equivalent iteration (residual, spectral radii and dt,
LU-SGS forward and backward sweeps, update, limit_flow, boundary conditions, halo, force_vt_kw)
over every block of the workload.  The grid and the state are resident in HBM before the timed
region starts; the timed region is exactly --steps iterations bracketed by device
synchronisation, and HIP events at the stage boundaries (on the library's own stream) give the
device time of each stage over the same iterations.

Workloads:
  naca0012_cgrid_257x97x33   single-block C-grid around NACA 0012, 257x97x33 nodes (786,432
                             cells), 2e-6 wall spacing, 20-chord far field; SA RANS, Re 6e6,
                             Mach 0.8, AoA 1.25 deg, WENO5 reconstruction, FP64
  channel_bump_97x33x9       inviscid channel with a 10% bump, 97x33x9 nodes, MUSCL3 + Roe
  naca4412_48block           577x193x129 C-grid around NACA 4412 cut 6x4x2 at seeded planes into 48
                             uneven blocks (14,155,776 cells); SST, Re 3e6, Mach 0.7, MUSCL3 + AUSMPW
  swept_wing_24block         1025x257x233 swept (30 deg), tapered (0.5) wing grid cut 4x3x2 into 24
                             blocks (60,817,408 cells); SA, Re 1.2e7, Mach 0.84, AoA 3, WENO5 + Roe

cpu_baseline: the compiled reference (oracle/_ref/ref_driver_omp, the reference's Fortran built
from its own sources with amdflang) on a bounded number of iterations of the same workload,
timed around its NS_Time_advance loop with every host core the process may use; when that binary
is not present, the oracle's C restatement is timed instead (kind "port").
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


def cpu_baseline(case, what: str, iters: int):
    """Reference CPU throughput on `iters` iterations of the workload sample (rank 0, N=1 only)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    cores = host_cores()
    cells = ncells(case)
    import refrun  # noqa: E402
    sample = f"{iters} iterations of {what} ({cells} cells) from freestream"
    if refrun.available(omp=True):
        try:
            r = refrun.run_case(case, nsteps=iters, kshow=10 ** 8, omp_threads=cores, dump=False, timeout=900)
            return dict(value=cells * iters / r["loop_seconds"] / 1e6, unit="Mcell-iter/s", cores=cores,
                        kind="reference", sample=sample + " (reference Fortran, MPI 1 rank x OpenMP)")
        except Exception as e:  # the reference binary could not run on this host
            print(f"# reference CPU build failed to run: {e}", file=sys.stderr)
    import oracle  # noqa: E402
    os.environ.setdefault("OMP_NUM_THREADS", str(cores))
    o = oracle.Oracle(case)
    o.init_flow()
    t0 = time.perf_counter()
    o.step(iters)
    dt = time.perf_counter() - t0
    return dict(value=cells * iters / dt / 1e6, unit="Mcell-iter/s", cores=1, kind="port",
                sample=sample + " (oracle C restatement, serial loops)")


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=40)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="naca0012_cgrid_257x97x33")
    ap.add_argument("--flux", type=int, default=4,
                    help="Iflag_Flux (default 4 Roe; 3 = the reference's HLLC, which diverges on "
                         "the C-grid within a few iterations)")
    ap.add_argument("--graph", type=int, default=1, help="replay each stage as a captured hipGraph")
    ap.add_argument("--lu", choices=("pipelined", "hyperplane", "fenced"), default="pipelined", help="LU-SGS sweep schedule")
    ap.add_argument("--cpu-iters", type=int, default=20, help="iterations of the CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU: the hot path runs on a single MI355X in this build; "
              "multi-GPU sharding by block is a design note, not an implemented path.", file=sys.stderr)
        return 2

    case = build_workload(args.workload, args.flux)
    cells = ncells(case)
    s = A.solver_from_case(case)          # geometry, wall distance, freestream: all on the device
    nv = s.nv
    s.set_graph(bool(args.graph))
    s.set_lu_schedule({"pipelined": A.LU_PIPELINED, "hyperplane": A.LU_HYPERPLANE, "fenced": A.LU_PIPELINED_FENCED}[args.lu])
    for _ in range(args.warmup):
        s.step(1)
    s.sync()

    t0 = time.perf_counter()
    stage_ms = s.step_timed(args.steps)   # returns after the device has finished
    s.sync()
    wall = time.perf_counter() - t0
    ms_per_step = wall * 1e3 / args.steps
    value = cells * args.steps / wall / 1e6

    mx, rms = s.norms()
    overlimit = max(s.status(no)[0] for no in range(1, len(case.blocks) + 1))

    # per-pass device time and algorithmic bytes
    words = PASS_WORDS[nv]
    passes = {}
    for (pname, stages), w in zip(PASS_STAGES, words):
        t = sum(stage_ms[x] for x in stages) / args.steps
        alg = w * 8.0 * cells
        passes[pname] = dict(ms=t, alg_bytes=alg, gbs=alg / (t * 1e-3) / 1e9)
    dom = max(passes, key=lambda k: passes[k]["ms"])
    d = passes[dom]
    roofline = dict(bound="hbm", achieved=round(d["gbs"], 2), peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=round(d["gbs"] / HBM_PEAK_GBS, 5), traffic=None,
                    kernel=dom, ms_per_launch=round(d["ms"], 4), alg_bytes_per_launch=d["alg_bytes"])
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
        "dtype": "f64",
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
        },
        "roofline": roofline,
        "iteration_hbm": {"alg_bytes_per_cell": sum(words) * 8, "achieved": round(iter_bytes / (ms_per_step * 1e-3) / 1e9, 2),
                          "frac": round(iter_bytes / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBS, 5)},
        "stage_ms_per_step": {k: round(v / args.steps, 4) for k, v in stage_ms.items()},
        "passes": {k: {"ms": round(v["ms"], 4), "GB/s": round(v["gbs"], 1)} for k, v in passes.items()},
        "state": {"res_rms1": float(rms[0]), "overlimit": int(overlimit)},
    }
    if not args.no_cpu:
        ccase, what = cpu_sample_case(args.workload, args.flux)
        out["cpu_baseline"] = cpu_baseline(ccase, what, args.cpu_iters)
        out["cpu_baseline"]["value"] = round(out["cpu_baseline"]["value"], 4)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
