"""Serving loop: requests admitted against a memory budget and advanced in one batch return the
same solve, iteration for iteration, as the plain solver running each request alone."""
import numpy as np
import pytest

import ocfd_amd as A
from ocfd_amd import cases as C
from ocfd_amd import serve as SV

SMALL = ((65, 25, 9), (97, 33, 9))


def test_sweep_requests_are_deterministic():
    a = SV.sweep_requests(6, lam=3.0, seed=11, dims=SMALL, max_blocks=3)
    b = SV.sweep_requests(6, lam=3.0, seed=11, dims=SMALL, max_blocks=3)
    assert [r.arrival for r in a] == [r.arrival for r in b]
    assert all(np.diff([r.arrival for r in a]) > 0)
    for r, q in zip(a, b):
        assert r.cells == q.cells == r.case.ncell
        assert r.case.params == q.case.params
        assert 1 <= len(r.case.blocks) <= 3
        assert sum(blk.ncell for blk in r.case.blocks) == r.cells


def test_summary_percentiles():
    res = [SV.Result(q, 0.0, 0.0, float(q + 1), 10, True, 1e-6, 0.1, 0.01, 100) for q in range(100)]
    s = SV.summarize(res, wall=100.0)
    assert s["requests"] == 100 and s["converged"] == 100
    assert s["p50_s"] == pytest.approx(50.5) and s["p99_s"] == pytest.approx(99.01)
    assert s["solves_per_hour"] == pytest.approx(3600.0)


@pytest.mark.gpu
def test_served_results_equal_sequential():
    reqs = SV.sweep_requests(5, lam=50.0, seed=7, dims=SMALL, max_blocks=3, Iflag_Flux=4)
    # a budget that holds about two requests at a time: admission has to hold some back
    big = max(reqs, key=lambda r: r.cells)
    s = A.solver_from_case(big.case)
    one = s.mem_bytes() + SV.Server.lu_bytes(big.case)
    s.close()
    srv = SV.Server(mem_budget=2.2 * one, round_iters=10, drop=1e-2, max_iters=300, graph=True)
    served = {r.rid: r for r in srv.run(reqs)}
    assert srv.peak_active >= 2 and srv.peak_used <= srv.mem_budget
    seq = {r.rid: r for r in srv.run_sequential([SV.Request(r.rid, 0.0, r.case, r.cells) for r in reqs])}
    assert sorted(served) == sorted(seq) == [r.rid for r in reqs]
    for rid in seq:
        a, b = served[rid], seq[rid]
        assert a.iterations == b.iterations, rid
        assert a.converged == b.converged
        assert a.drop == b.drop and a.CL == b.CL and a.CD == b.CD, rid


class _FakeSolver:
    """Stand-in for a solver on the CPU: Res_rms(1) halves every 10 iterations; a request marked
    to diverge has a NaN residual (the library's OCFD_ENAN) from iteration `fail_at` on."""

    def __init__(self, case, fail_at=None):
        self.case, self.it, self.fail_at, self.closed = case, 0, fail_at, False

    def mem_bytes(self):
        return 1000

    def set_graph(self, on):
        pass

    def step(self, n=1):
        assert not self.closed
        self.it += n

    def norms(self):
        if self.fail_at is not None and self.it >= self.fail_at:
            raise A.OcfdError(A.OCFD_ENAN, "NaN in Residual is found !")
        return np.zeros(6), np.full(6, 0.5 ** (self.it / 10.0))

    def forces(self):
        return np.arange(12, dtype=float)

    def close(self):
        self.closed = True


class _FakeBatch:
    def __init__(self):
        self.members, self.closed = [], False

    def set_graph(self, on):
        pass

    def add(self, s):
        self.members.append(s)

    def remove(self, s):
        self.members.remove(s)

    def step(self, n=1):
        for s in self.members:
            s.step(n)

    def close(self):
        self.closed = True


@pytest.mark.parametrize("fail_at", [1, 21, 45])
def test_diverging_request_is_retired_not_raised(fail_at):
    """A request whose residual turns NaN (in its first iteration or later) is retired as failed;
    the other requests still complete, and every solver and the batch are closed."""
    reqs = [SV.Request(q, 0.001 * q, C.channel_bump(9, 5, 3), 128) for q in range(4)]
    made, batches = [], []

    def make_solver(case):
        s = _FakeSolver(case, fail_at=fail_at if case is reqs[1].case else None)
        made.append(s)
        return s

    def make_batch():
        b = _FakeBatch()
        batches.append(b)
        return b

    srv = SV.Server(mem_budget=1e12, round_iters=20, drop=1e-2, max_iters=400, graph=False,
                    make_solver=make_solver, make_batch=make_batch)
    res = {r.rid: r for r in srv.run(reqs)}
    assert sorted(res) == [0, 1, 2, 3]
    assert res[1].failed and not res[1].converged and np.isnan(res[1].CL)
    for q in (0, 2, 3):
        assert res[q].converged and not res[q].failed and res[q].CL == 9.0
    assert all(s.closed for s in made) and len(batches) == 1 and batches[0].closed
    seq = {r.rid: r for r in srv.run_sequential(reqs)}
    assert seq[1].failed and all(seq[q].converged for q in (0, 2, 3))
    assert all(s.closed for s in made)
    summ = SV.summarize(list(res.values()), 1.0)
    assert summ["failed"] == 1 and summ["converged"] == 3


def test_preparer_error_reaches_the_serving_loop():
    """An error while a request is prepared (here: its solver cannot be built) is raised by the
    serving loop after the batch and the solvers already built are closed."""
    reqs = [SV.Request(q, 0.0, C.channel_bump(9, 5, 3), 128) for q in range(3)]
    made, batches = [], []

    def make_solver(case):
        if case is reqs[2].case:
            raise A.OcfdError(A.OCFD_EHIP, "out of memory")
        s = _FakeSolver(case)
        made.append(s)
        return s

    def make_batch():
        batches.append(_FakeBatch())
        return batches[-1]

    srv = SV.Server(mem_budget=1e12, round_iters=20, drop=1e-2, max_iters=400, graph=False,
                    make_solver=make_solver, make_batch=make_batch)
    with pytest.raises(A.OcfdError):
        srv.run(reqs)
    assert all(s.closed for s in made) and batches[0].closed


@pytest.mark.gpu
def test_served_results_equal_sequential_large_grids():
    """Requests drawn like the serving sweep's at its grid sizes (129x49x17, 193x65x25, 257x97x33
    C-grids cut along i into 1-8 blocks; NACA 4-digit sections; random Mach, AoA, Re; SA, MUSCL3 +
    Van Leer): served through the batch, each request's iterations, residual drop and forces equal
    its sequential solve exactly."""
    reqs = SV.sweep_requests(6, lam=20.0, seed=5, dims=SV.DIMS_LARGE)
    assert {tuple(r.case.blocks[0].dims[1:]) for r in reqs} <= {(49, 17), (65, 25), (97, 33)}
    assert any(len(r.case.blocks) > 1 for r in reqs)
    srv = SV.Server(mem_budget=200e9, round_iters=20, drop=3e-2, max_iters=200, graph=True)
    served = {r.rid: r for r in srv.run(reqs)}
    assert srv.peak_active >= 2
    seq = {r.rid: r for r in srv.run_sequential([SV.Request(r.rid, 0.0, r.case, r.cells) for r in reqs])}
    assert sorted(served) == sorted(seq) == [r.rid for r in reqs]
    for rid in seq:
        a, b = served[rid], seq[rid]
        print(f"  request {rid}: cells {a.cells} iters {a.iterations} drop {a.drop:.3e} CL {a.CL:.6f}", flush=True)
        assert not a.failed and not b.failed
        assert a.iterations == b.iterations, rid
        assert a.converged == b.converged
        assert a.drop == b.drop and a.CL == b.CL and a.CD == b.CD, rid
