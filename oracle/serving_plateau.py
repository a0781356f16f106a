#!/usr/bin/env python3
"""Residual history of one serving-sweep request in the oracle (TEST INFRASTRUCTURE ONLY).

The serving sweep's stop rule is a drop of Res_rms(1) relative to its first iteration.  This
script runs, in the oracle on the CPU, the first request of the sweep's stream (seed 5, the
129x49x17 / 193x65x25 / 257x97x33 grids) on the smallest grid, and prints the drop and
the force coefficients every 200 iterations, to show where such a solve's residual settles.

    python oracle/serving_plateau.py [ITERATIONS]      (12001 by default: about 25 minutes)
"""
from __future__ import annotations

import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [os.path.join(os.path.dirname(HERE), "opencfd-ec_amd"), HERE]

from ocfd_amd import serve as SV  # noqa: E402
import oracle  # noqa: E402


def main() -> None:
    n = int(sys.argv[1]) if len(sys.argv) > 1 else 12001
    reqs = SV.sweep_requests(16, lam=2.0, seed=5, dims=SV.DIMS_LARGE)
    r = min(reqs, key=lambda q: (q.cells, q.rid))
    p = r.case.params
    print(f"request {r.rid}: {r.cells} cells in {len(r.case.blocks)} blocks, Ma {p['Ma']:.4f} AoA {p['AoA']:.4f} "
          f"Re {p['Re']:.4e}", flush=True)
    o = oracle.Oracle(r.case)
    o.init_flow()
    o.step(1)
    r0 = o.monitor()["res_rms"][0]
    t0, done = time.time(), 1
    while done < n:
        o.step(200)
        done += 200
        st = o.monitor()
        print(f"{done} drop {st['res_rms'][0] / r0:.3e} CL {st['CL']:.10f} CD {st['CD']:.10f} {time.time() - t0:.0f}s", flush=True)


if __name__ == "__main__":
    main()
