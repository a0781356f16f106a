#!/usr/bin/env python3
"""Writes tests/golden/converged_sa_cgrid_65x25x9.json: the converged force coefficients of an SA
C-grid from the compiled reference and from the oracle.

Case: the 65x25x9 NACA 0012 C-grid (wall spacing 0.002 chord, 20-chord far field, span 1), SA,
Re 1e5, Mach 0.5, AoA 2 deg, MUSCL3 (van Albada) + Van Leer, LU-SGS, CFL 10, from freestream
for 30000 iterations; by then Res_rms(1) has dropped by about 2e-11 and CL, CD, CM move by less
than 1e-13 relative per 1000 iterations.

The compiled reference (oracle/_ref/ref_driver, the reference's own Fortran) prints CL, CD, CS at
full precision (comput_force, sub_Post.f90:227-228) and writes force.log (CL, CD, CS and the
moments Ft0(7:9) to 9 digits, :240-241); the oracle gives all of them at full precision.  The
fixture keeps both; the GPU test holds the device to the reference's CL and CD and to the
oracle's moments, each within 1e-8 relative.

    python tests/golden/make_converged.py        (about 10 minutes; needs oracle/_ref/ref_driver)
"""
from __future__ import annotations

import json
import os
import re
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(ROOT, "opencfd-ec_amd"), os.path.join(ROOT, "oracle")]

from ocfd_amd import cases as C  # noqa: E402
import oracle  # noqa: E402
import refrun  # noqa: E402

NSTEPS = 30000
CASE_ARGS = dict(wall_spacing=0.002, Re=1e5, Ma=0.5, AoA=2.0, Iflag_Scheme=5, Iflag_Flux=5, CFL=10.0)
OUT = os.path.join(HERE, "converged_sa_cgrid_65x25x9.json")


def build_case():
    return C.naca_cgrid(65, 25, 9, **CASE_ARGS)


def run_reference():
    r = refrun.run_case(build_case(), nsteps=NSTEPS, kshow=1000, dump=False, timeout=3600)
    # the last "Total Force Coefficient (CL, CD, CS) =" block of the reference's stdout
    txt = r["stdout"]
    pos = txt.rfind("Total Force Coefficient")
    nums = re.findall(r"[-+]?\d*\.\d+(?:[EeDd][-+]?\d+)?", txt[pos:].split("\n", 1)[1])[:3]
    cl, cd, cs = (float(v.replace("D", "E").replace("d", "e")) for v in nums)
    steps, _, rms = r["norms"]
    f = r["force"][-1]
    return dict(CL=cl, CD=cd, CS=cs, forcelog_last=[float(v) for v in f],
                res_rms1=[[int(s), float(v)] for s, v in zip(steps, rms[:, 0])])


def run_oracle():
    o = oracle.Oracle(build_case())
    o.init_flow()
    o.step(1)
    r0 = float(o.monitor()["res_rms"][0])
    o.step(NSTEPS - 1)
    st = o.monitor()
    return dict(CL=float(st["CL"]), CD=float(st["CD"]), CS=float(st["CS"]), Ft=[float(v) for v in st["Ft"]],
                res_rms1_first=r0, res_rms1_last=float(st["res_rms"][0]))


def main():
    with ThreadPoolExecutor(2) as ex:
        fr, fo = ex.submit(run_reference), ex.submit(run_oracle)
        ref, orc = fr.result(), fo.result()
    data = dict(case="naca_cgrid(65, 25, 9, **case_args)", case_args=CASE_ARGS, iterations=NSTEPS,
                reference=ref, oracle=orc)
    with open(OUT, "w") as f:
        json.dump(data, f, indent=1)
    print(json.dumps({k: (ref[k], orc[k]) for k in ("CL", "CD", "CS")}))


if __name__ == "__main__":
    main()
