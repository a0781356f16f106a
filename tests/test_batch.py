"""Batches: several solves advancing through shared launches give each solve's result alone.

Members differ in grid, block count and freestream (Mach, AoA, Re); they join and leave the
batch between steps, as requests do in the serving loop.  Every member's state must equal,
bit for bit, the same case stepped alone for the same number of iterations.
"""
import numpy as np
import pytest

import ocfd_amd as A
from ocfd_amd import cases as C

CG = dict(wall_spacing=0.02, Iflag_Flux=4)


def _cases():
    a = C.naca_cgrid(65, 25, 9, **CG)
    b = C.naca_cgrid(65, 25, 9, Ma=0.6, AoA=3.0, Re=3e6, **CG)
    c = C.cut_blocks(C.naca_cgrid(65, 25, 9, m=0.02, p=0.4, Ma=0.7, AoA=-1.0, **CG), [[[17, 40], [12], [5]]])
    return [a, b, c]


def _alone(case, seed, n):
    s = A.solver_from_case(case, states=C.perturbed_state(case, seed=seed, amp=1e-4))
    s.step(n)
    return [s.download_U(no) for no in range(1, len(case.blocks) + 1)], s.norms()[1]


@pytest.mark.gpu
@pytest.mark.parametrize("graph", [False, True])
def test_batch_members_equal_alone(graph):
    cases = _cases()
    S = [A.solver_from_case(cs, states=C.perturbed_state(cs, seed=10 + q, amp=1e-4)) for q, cs in enumerate(cases)]
    bt = A.Batch()
    bt.set_graph(graph)
    bt.add(S[0]); bt.add(S[1])
    bt.step(2)                       # a: 2, b: 2
    bt.add(S[2])
    bt.step(2)                       # a: 4, b: 4, c: 2
    bt.remove(S[0])
    bt.step(1)                       # b: 5, c: 3
    iters = [4, 5, 3]
    for q, (cs, s) in enumerate(zip(cases, S)):
        ref, ref_rms = _alone(cs, 10 + q, iters[q])
        for no in range(1, len(cs.blocks) + 1):
            assert np.array_equal(s.download_U(no), ref[no - 1]), f"member {q} block {no}"
        assert np.array_equal(s.norms()[1], ref_rms)
        assert s.counters()[0] == iters[q]
    bt.close()
