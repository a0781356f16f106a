#!/usr/bin/env python3
"""Writes the golden fixtures under tests/golden/ from the compiled reference.

Each fixture is one small case run by the reference's own Fortran (oracle/_ref/ref_driver, built
from the reference sources by oracle/ref.mk): the case definition (node coordinates, subface
table, control.ec values, seeded initial interior state) and what the reference produced from
it: the per-cell Res, DU, dt, mu_t and U after the first iteration, U after the last one, the
residual norms of every iteration (norms.txt) and force.log.  The tests rebuild the case from
the fixture alone, so fixtures stay valid when the case generators change.

    python tests/golden/make_golden.py            (needs oracle/_ref/ref_driver)
"""
from __future__ import annotations

import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(ROOT, "opencfd-ec_amd"), os.path.join(ROOT, "oracle")]

from ocfd_amd import cases as C  # noqa: E402
import refrun  # noqa: E402

NSTEPS = 3
CG = dict(wall_spacing=0.05)

# name -> (case, seed, amplitude)
FIXTURES = {
    "channel_roe_muscl3": (lambda: C.channel_bump(17, 9, 5), 11, 1e-3),
    "channel_cut_hllc_weno5": (lambda: C.cut_blocks(C.channel_bump(21, 11, 6, Iflag_Flux=3, Iflag_Scheme=7),
                                                    [[[11], [6], []]]), 12, 1e-3),
    "cgrid_sa_weno5_roe": (lambda: C.naca_cgrid(33, 13, 5, Iflag_Flux=4, **CG), 13, 1e-5),
    "cgrid_sa_limiter": (lambda: C.naca_cgrid(33, 13, 5, Iflag_Flux=4, Ldmax=1.001, LSAmax=5.004, **CG), 14, 1e-3),
    "cgrid_sst_muscl3_ausm": (lambda: C.naca_cgrid(33, 13, 5, Iflag_turbulence_model=3, Iflag_Scheme=5, Iflag_Flux=6,
                                                   Re=3e6, Ma=0.7, **CG), 15, 1e-5),
    "cgrid_laminar_iso_vl": (lambda: C.naca_cgrid(33, 13, 5, Iflag_turbulence_model=0, Re=1e4, Twall=400.0,
                                                  Iflag_Scheme=5, Iflag_Flux=5, **CG), 16, 1e-5),
}


def _interior(arr, nx, ny, nz, nv=0):
    if nv:
        return np.ascontiguousarray(arr.sl((1, nv), (1, nx - 1), (1, ny - 1), (1, nz - 1)))
    return np.ascontiguousarray(arr.sl((1, nx - 1), (1, ny - 1), (1, nz - 1)))


def make(name: str) -> str:
    builder, seed, amp = FIXTURES[name]
    case = builder()
    U0 = C.perturbed_state(case, seed=seed, amp=amp)
    r = refrun.run_case(case, nsteps=NSTEPS, kshow=1, U_init=U0)
    if "stateN" not in r:
        raise RuntimeError(f"{name}: the reference stopped before iteration {NSTEPS}")
    nv = C.nvar_of(case.params)
    out = {"params_json": np.array(json.dumps(case.params, default=list)), "nsteps": np.array(NSTEPS),
           "nblocks": np.array(len(case.blocks)), "bc3d_inc": np.array(r["bc3d_inc"])}
    for n, b in enumerate(case.blocks, start=1):
        nx, ny, nz = b.dims
        out[f"b{n}_x"], out[f"b{n}_y"], out[f"b{n}_z"] = b.x, b.y, b.z
        out[f"b{n}_sub"] = np.array([list(s.rng) + [s.bc] + list(s.link or (0,) * 7) for s in b.subfaces], dtype=np.int32)
        out[f"b{n}_U0"] = np.asarray(U0[n - 1])
        s1, sN = r["state1"][n - 1], r["stateN"][n - 1]
        out[f"b{n}_Res1"] = _interior(s1["Res"], nx, ny, nz, nv)
        out[f"b{n}_DU1"] = _interior(s1["DU"], nx, ny, nz, nv)
        out[f"b{n}_dt1"] = _interior(s1["dt"], nx, ny, nz)
        if "mu_t" in s1:
            out[f"b{n}_mut1"] = _interior(s1["mu_t"], nx, ny, nz)
        out[f"b{n}_U1"] = _interior(s1["U"], nx, ny, nz, nv)
        out[f"b{n}_UN"] = _interior(sN["U"], nx, ny, nz, nv)
        out[f"b{n}_overlimit"] = np.array(sN["overlimit"])
    ks, mx, rms = r["norms"]
    out["norms_kstep"], out["norms_max"], out["norms_rms"] = ks, mx, rms
    out["force_log"] = r["force"]
    path = os.path.join(HERE, name + ".npz")
    np.savez_compressed(path, **out)
    return path


if __name__ == "__main__":
    if not refrun.available():
        sys.exit("oracle/_ref/ref_driver is not built (make -f oracle/ref.mk serial)")
    for nm in (sys.argv[1:] or FIXTURES):
        p = make(nm)
        print(f"{p}: {os.path.getsize(p) / 1024:.0f} KiB")
