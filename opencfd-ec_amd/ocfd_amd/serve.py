"""Steady-solve serving: requests arrive, are admitted against an HBM budget, and advance
together through shared batched launches until their residual has dropped by the requested
factor; each result is then returned and its memory freed.

A request is a complete case (the client sends its grid, freestream and options).  Admission:
the request's device memory (block arrays plus its share of the LU-SGS tables) must fit in what
the budget has left; a request that does not fit waits.  Convergence: Res_rms(1) relative to
its value after the request's first iteration (the reference itself has no stop criterion
besides t_end, opencfd_ec3d_v1.16a.f90:163).

`sweep_requests` draws the request stream of the serving benchmark: Poisson arrivals, NACA
4-digit sections extruded to span 1 on one of three C-grid sizes cut along i into 1-8 blocks,
random Mach / AoA / Re, SA, MUSCL3 + Van Leer (the reference's default scheme and flux), CFL 5.
On these grids the SA solves settle into a small limit cycle with Res_rms(1) between 1e-3 and
3e-3 of its first-iteration value; the sweep's stop criterion is therefore a drop of 5e-3.
"""
from __future__ import annotations

import dataclasses
import math
import queue
import threading
import time
from typing import Callable, Dict, List, Optional

import numpy as np

from . import OCFD_ENAN, Batch, OcfdError, Solver, solver_from_case
from . import cases as C


@dataclasses.dataclass
class Request:
    rid: int
    arrival: float                 # seconds after the start of the sweep
    case: C.Case
    cells: int = 0


@dataclasses.dataclass
class Result:
    rid: int
    arrival: float
    admitted: float
    done: float
    iterations: int
    converged: bool
    drop: float                    # final Res_rms(1) / Res_rms(1) after the first iteration
    CL: float
    CD: float
    cells: int
    failed: bool = False           # the residual held a NaN (OCFD_ENAN): the solve diverged

    @property
    def latency(self) -> float:
        return self.done - self.arrival


DIMS = ((97, 33, 9), (129, 49, 13), (193, 65, 17))
DIMS_LARGE = ((129, 49, 17), (193, 65, 25), (257, 97, 33))


def sweep_requests(n: int = 256, lam: float = 2.0, seed: int = 5, dims=DIMS, max_blocks: int = 8,
                   **overrides) -> List[Request]:
    """The serving sweep's request stream (deterministic for a seed)."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / lam, n))
    out = []
    for q in range(n):
        m, pp, tk = rng.uniform(0.0, 0.04), rng.uniform(0.30, 0.50), rng.uniform(0.08, 0.15)
        ni, nj, nk = dims[int(rng.integers(0, len(dims)))]
        nb = int(rng.integers(1, max_blocks + 1))
        Ma, AoA, Re = rng.uniform(0.3, 0.85), rng.uniform(-2.0, 6.0), rng.uniform(1e6, 1e7)
        kw = dict(m=m, p=pp, t=tk, span=1.0, Ma=Ma, AoA=AoA, Re=Re, Iflag_turbulence_model=2,
                  Iflag_Scheme=5, Iflag_Flux=5, CFL=5.0)
        kw.update(overrides)
        case = C.naca_cgrid(ni, nj, nk, **kw)
        if nb > 1:
            cuts = sorted(int(v) for v in np.linspace(1, ni, nb + 1)[1:-1].round())
            case = C.cut_blocks(case, [[cuts, [], []]])
        out.append(Request(q, float(t[q]), case, case.ncell))
    return out


class Server:
    """Continuous batching of steady solves on one GPU.

    mem_budget: bytes of device memory the admitted and prepared requests may hold together.
    round_iters: iterations per batched step between admission/convergence checks.
    A preparer thread builds each arrived request (grid, device geometry and wall distance, its
    first iteration on its own stream) while the batch runs; the serving loop only admits the
    prepared requests into the batch between rounds.  make_solver / make_batch build the solver of
    a case and the batch (the library's by default).
    """

    def __init__(self, mem_budget: float = 200e9, round_iters: int = 20, drop: float = 5e-3,
                 max_iters: int = 6000, graph: bool = True, make_solver: Callable = solver_from_case,
                 make_batch: Callable = Batch):
        self.mem_budget = mem_budget
        self.round_iters = round_iters
        self.drop = drop
        self.max_iters = max_iters
        self.graph = graph
        self.make_solver = make_solver
        self.make_batch = make_batch

    @staticmethod
    def lu_bytes(case) -> int:
        """Device bytes of a request's share of the batch's tables: the LU-SGS sweep storage
        (4 NVAR + 21 doubles per lane and step of every 64-column (j,k) tile, a tile marching
        nx - 1 + tj + tk - 2 steps; the widest tile shape) and 16 MB for the launch tables."""
        nv = C.nvar_of(case.params)
        steps = 0
        for b in case.blocks:
            nx, ny, nz = b.dims
            tiles = -(-(ny - 1) // 4) * -(-(nz - 1) // 16)
            steps += max(tiles, -(-(ny - 1) // 16) * -(-(nz - 1) // 4)) * (nx - 1 + 18)
        return int(steps * 64 * (4 * nv + 21) * 8 + 16e6)

    @staticmethod
    def rms1(s) -> float:
        """Res_rms(1) of a solver's last iteration; NaN when its residual holds a NaN (OCFD_ENAN)."""
        try:
            return float(s.norms()[1][0])
        except OcfdError as e:
            if e.code == OCFD_ENAN:
                return math.nan
            raise

    def _result(self, a, tnow, ratio, conv, failed):
        if failed:
            cl = cd = math.nan
        else:
            f = a["s"].forces()
            cl, cd = float(f[9]), float(f[10])
        return Result(a["req"].rid, a["req"].arrival, a["admitted"], tnow, a["iters"], conv, ratio, cl, cd,
                      a["req"].cells, failed)

    def _preparer(self, pending, ready, st, t0, clock):
        """Thread body: for each request in arrival order, wait for its arrival, build its solver
        and run its first iteration once the memory budget has room, then hand it over.

        The device work of this thread holds st["dev"], which the serving loop also holds while
        the batch may capture its graphs (the first round after a change of membership): a graph
        capture in progress on one thread must not see device calls from another.  The first
        iteration here runs without graphs, so this thread never captures."""
        try:
            for r in pending:
                while clock() - t0 < r.arrival:
                    if st["stop"].is_set():
                        return
                    time.sleep(min(0.02, max(0.0, r.arrival - (clock() - t0))))
                with st["dev"]:
                    s = self.make_solver(r.case)
                    need = s.mem_bytes() + self.lu_bytes(r.case)
                while True:
                    with st["lock"]:
                        held = st["used"] + st["reserved"]
                        if held == 0 or held + need <= self.mem_budget:
                            st["reserved"] += need
                            break
                    if st["stop"].is_set():
                        with st["dev"]:
                            s.close()
                        return
                    time.sleep(0.005)
                with st["dev"]:
                    s.set_graph(False)
                    s.step(1)
                    rms0 = self.rms1(s)
                ready.put((r, s, need, rms0, clock() - t0))
        except BaseException as e:          # handed to the serving loop, which raises it
            ready.put(e)
        finally:
            ready.put(None)

    def run(self, requests: List[Request], clock: Callable[[], float] = time.monotonic,
            log: Optional[Callable[[str], None]] = None) -> List[Result]:
        pending = sorted(requests, key=lambda r: r.arrival)
        active: Dict[int, dict] = {}
        results: List[Result] = []
        ready: "queue.Queue" = queue.Queue()
        st = dict(lock=threading.Lock(), dev=threading.Lock(), used=0, reserved=0, stop=threading.Event())
        flags = dict(recapture=True)
        self.peak_active, self.peak_used = 0, 0
        t0 = clock()
        batch = self.make_batch()
        prep = threading.Thread(target=self._preparer, args=(pending, ready, st, t0, clock), daemon=True)
        waiting = []                         # prepared solvers not yet admitted (closed on exit)
        try:
            batch.set_graph(self.graph)
            prep.start()
            preparing = True

            def admit(item):
                r, s, need, rms0, tprep = item
                with st["lock"]:
                    st["reserved"] -= need
                    st["used"] += need
                    used = st["used"]
                a = dict(req=r, s=s, rms0=rms0, iters=1, admitted=clock() - t0, need=need)
                if not math.isfinite(rms0):          # diverged in its first iteration
                    results.append(self._result(a, clock() - t0, math.nan, False, True))
                    s.close()
                    with st["lock"]:
                        st["used"] -= need
                    return
                batch.add(s)
                flags["recapture"] = True
                active[r.rid] = a
                self.peak_active = max(self.peak_active, len(active))
                self.peak_used = max(self.peak_used, used)
                if log:
                    log(f"t={a['admitted']:7.2f}s admit {r.rid} cells={r.cells} active={len(active)} mem={used / 1e9:.1f}GB")

            while preparing or active:
                # admit every prepared request; with nothing active, wait for the next one
                block = not active
                while True:
                    try:
                        item = ready.get(timeout=0.05) if block else ready.get_nowait()
                    except queue.Empty:
                        break
                    if item is None:
                        preparing = False
                        break
                    if isinstance(item, BaseException):
                        raise item
                    waiting.append(item[1])
                    admit(item)
                    waiting.remove(item[1])
                    block = False
                if not active:
                    continue
                if flags["recapture"]:
                    # the batch captures its stage graphs in this round: no device calls from the
                    # preparing thread meanwhile
                    with st["dev"]:
                        batch.step(self.round_iters)
                    flags["recapture"] = False
                else:
                    batch.step(self.round_iters)
                tnow = clock() - t0
                for rid in list(active):
                    a = active[rid]
                    a["iters"] += self.round_iters
                    rms = self.rms1(a["s"])
                    failed = not math.isfinite(rms)
                    ratio = rms / a["rms0"] if a["rms0"] > 0 else 0.0
                    conv = (not failed) and ratio <= self.drop
                    if conv or failed or a["iters"] >= self.max_iters:
                        results.append(self._result(a, tnow, ratio, conv, failed))
                        batch.remove(a["s"])
                        flags["recapture"] = True
                        a["s"].close()
                        with st["lock"]:
                            st["used"] -= a["need"]
                        del active[rid]
                        if log:
                            log(f"t={tnow:7.2f}s done {rid} iters={a['iters']} drop={ratio:.2e} active={len(active)}"
                                + (" FAILED (NaN)" if failed else ""))
        finally:
            st["stop"].set()
            while prep.is_alive():
                try:
                    item = ready.get(timeout=0.05)
                    if isinstance(item, tuple):
                        item[1].close()
                except queue.Empty:
                    pass
            while not ready.empty():
                item = ready.get_nowait()
                if isinstance(item, tuple):
                    item[1].close()
            for s in waiting:
                s.close()
            for a in active.values():
                try:
                    batch.remove(a["s"])
                except OcfdError:
                    pass
                a["s"].close()
            batch.close()
        return results

    def run_sequential(self, requests: List[Request], clock: Callable[[], float] = time.monotonic,
                       log: Optional[Callable[[str], None]] = None) -> List[Result]:
        """The baseline: one request at a time on its own solver, in arrival order, same stop rule."""
        results: List[Result] = []
        t0 = clock()
        for r in sorted(requests, key=lambda q: q.arrival):
            now = clock() - t0
            if r.arrival > now:
                time.sleep(r.arrival - now)
            adm = clock() - t0
            s = self.make_solver(r.case)
            try:
                s.set_graph(self.graph)
                s.step(1)
                rms0 = self.rms1(s)
                iters, ratio, failed = 1, 1.0, not math.isfinite(rms0)
                while not failed:
                    s.step(self.round_iters)
                    iters += self.round_iters
                    rms = self.rms1(s)
                    failed = not math.isfinite(rms)
                    ratio = rms / rms0 if rms0 > 0 else 0.0
                    if ratio <= self.drop or iters >= self.max_iters:
                        break
                a = dict(req=r, s=s, iters=iters, admitted=adm)
                conv = (not failed) and ratio <= self.drop
                results.append(self._result(a, clock() - t0, math.nan if failed else ratio, conv, failed))
            finally:
                s.close()
            if log:
                log(f"t={clock() - t0:7.2f}s sequential done {r.rid} cells={r.cells} iters={iters} drop={ratio:.2e}")
        return results


def summarize(results: List[Result], wall: float) -> Dict[str, float]:
    lat = np.array([r.latency for r in results]) if results else np.zeros(1)
    return dict(requests=len(results), converged=int(sum(r.converged for r in results)),
                failed=int(sum(r.failed for r in results)),
                p50_s=float(np.percentile(lat, 50)), p99_s=float(np.percentile(lat, 99)),
                solves_per_hour=3600.0 * len(results) / wall if wall > 0 else 0.0,
                mean_iters=float(np.mean([r.iterations for r in results])) if results else 0.0,
                cell_iters=float(sum(r.cells * r.iterations for r in results)), wall_s=wall)
