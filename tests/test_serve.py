"""Serving loop: requests admitted against a memory budget and advanced in one batch return the
same solve, iteration for iteration, as the plain solver running each request alone."""
import numpy as np
import pytest

import ocfd_amd as A
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
