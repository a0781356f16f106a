"""Run the compiled reference (oracle/_ref/ref_driver) on a generated case and read its dumps.

Test infrastructure only.  The driver executes the reference's own setup and
NS_Time_advance routines (see oracle/ref_driver.f90); this module writes the case files into a
scratch directory, runs the driver there and parses the named-array stream files it writes.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import tempfile
from typing import Dict, List, Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF_DRIVER = os.path.join(HERE, "_ref", "ref_driver")
REF_DRIVER_OMP = os.path.join(HERE, "_ref", "ref_driver_omp")


def available(omp: bool = False) -> bool:
    return os.access(REF_DRIVER_OMP if omp else REF_DRIVER, os.X_OK)


class Arr:
    """A dumped Fortran array with its lower bounds; index with reference (Fortran) indices."""

    def __init__(self, data: np.ndarray, lb):
        self.data = data
        self.lb = tuple(lb)

    def at(self, *idx):
        return self.data[tuple(i - l for i, l in zip(idx, self.lb))]

    def sl(self, *ranges):
        """Inclusive Fortran ranges (lo, hi) per dim -> numpy view."""
        return self.data[tuple(slice(lo - l, hi - l + 1) for (lo, hi), l in zip(ranges, self.lb))]


def read_dump(path: str) -> List[Dict[str, object]]:
    blocks: List[Dict[str, object]] = []
    with open(path, "rb") as f:
        buf = f.read()
    pos = 0
    cur: Optional[Dict[str, object]] = None
    while pos < len(buf):
        (nlen,) = np.frombuffer(buf, "<i4", 1, pos); pos += 4
        name = buf[pos:pos + nlen].decode(); pos += nlen
        kind, nd = np.frombuffer(buf, "<i4", 2, pos); pos += 8
        bnds = np.frombuffer(buf, "<i4", 2 * nd, pos).reshape(nd, 2); pos += 8 * nd
        lb, ub = bnds[:, 0], bnds[:, 1]
        if kind == 0:
            cur = dict(block_no=int(lb[0]), nx=int(ub[0]), ny=int(lb[1]), nz=int(ub[1]),
                       nvar=int(lb[2]), overlimit=int(ub[2]))
            blocks.append(cur)
            continue
        shape = tuple(int(u - l + 1) for l, u in zip(lb, ub))
        n = int(np.prod(shape))
        dt = "<f8" if kind == 1 else "<i4"
        a = np.frombuffer(buf, dt, n, pos).reshape(shape, order="F")
        pos += n * 8 if kind == 1 else n * 4
        assert cur is not None
        cur[name] = Arr(a, lb)
    return blocks


def read_norms(path: str, nvar: int):
    rows = []
    with open(path) as f:
        for line in f:
            v = line.split()
            if v:
                rows.append([float(x) for x in v])
    a = np.array(rows).reshape(-1, 1 + 2 * nvar)
    return a[:, 0].astype(int), a[:, 1:1 + nvar], a[:, 1 + nvar:]


def read_forcelog(path: str):
    rows = []
    if not os.path.exists(path):
        return np.zeros((0, 7))
    with open(path) as f:
        for line in f:
            v = line.split()
            if v:
                rows.append([float(x) for x in v])
    return np.array(rows)


def run_case(case, nsteps: int = 1, kshow: int = 1, U_init=None, omp_threads: int = 0,
             keep: Optional[str] = None, timeout: float = 1800.0, dump: bool = True,
             wall_dist=None) -> Dict[str, object]:
    """Run the reference driver on ``case`` (an ocfd_amd.cases.Case) and return its dumps.

    With ``dump=False`` the driver writes no arrays; the result then holds the wall-clock seconds
    of the iteration loop (``loop_seconds``) and the norms/force logs only.  ``wall_dist`` (dw per
    block, numpy [i-1, j-1, k-1]) is written as wall_dist.dat, which the reference then reads
    instead of computing the wall distance by brute force (Comput_dist_wall, sub_comput_dw.f90:12-41)."""
    import sys
    sys.path.insert(0, os.path.join(os.path.dirname(HERE), "opencfd-ec_amd"))
    from ocfd_amd import cases as C  # noqa: E402

    d = tempfile.mkdtemp(prefix="ocfd_ref_")
    try:
        params = dict(case.params)
        if omp_threads > 0:
            params["NUM_THREADS"] = omp_threads
        C.write_case(C.Case(case.name, case.blocks, params), d)
        if U_init is not None:
            C.write_initial_state(os.path.join(d, "U_init.bin"), U_init)
        if wall_dist is not None:
            from ocfd_amd import restart as R  # noqa: E402
            R.write_wall_dist(os.path.join(d, "wall_dist.dat"), wall_dist)
        exe = REF_DRIVER_OMP if omp_threads > 0 else REF_DRIVER
        env = dict(os.environ)
        if omp_threads > 0:
            env["OMP_NUM_THREADS"] = str(omp_threads)
        r = subprocess.run(["bash", "-c", f"ulimit -s unlimited; exec {exe} {nsteps} {kshow} {1 if dump else 0}"],
                           cwd=d, capture_output=True, text=True, timeout=timeout, env=env)
        if r.returncode != 0:
            raise RuntimeError(f"reference driver failed ({r.returncode}):\n{r.stdout[-3000:]}\n{r.stderr[-3000:]}")
        out: Dict[str, object] = dict(stdout=r.stdout)
        for name in ("geom", "state0", "state1", "stateN"):
            p = os.path.join(d, name + ".bin")
            if os.path.exists(p):
                out[name] = read_dump(p)
        with open(os.path.join(d, "timing.txt")) as f:
            out["loop_seconds"] = float(f.read().split()[1])
        out["norms"] = read_norms(os.path.join(d, "norms.txt"), C.nvar_of(case.params))
        out["force"] = read_forcelog(os.path.join(d, "force.log"))
        with open(os.path.join(d, "bc3d.inc")) as f:
            out["bc3d_inc"] = f.read()
        for name in ("flow3d.dat", "SA3d.dat", "SST3d.dat", "wall_dist.dat", "Step_mess.dat"):
            p = os.path.join(d, name)
            if os.path.exists(p):
                with open(p, "rb") as f:
                    out[name] = f.read()
        return out
    finally:
        if keep:
            shutil.copytree(d, keep, dirs_exist_ok=True)
        shutil.rmtree(d, ignore_errors=True)
