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


def _alone(case, seed, n, du=False):
    s = A.solver_from_case(case, states=C.perturbed_state(case, seed=seed, amp=1e-4))
    s.step(n)
    U = [s.download_U(no) for no in range(1, len(case.blocks) + 1)]
    if du:
        return U, s.norms()[1], [s.field(no, A.F_DU) for no in range(1, len(case.blocks) + 1)]
    return U, s.norms()[1]


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
        ref, ref_rms, ref_du = _alone(cs, 10 + q, iters[q], du=True)
        for no in range(1, len(cs.blocks) + 1):
            assert np.array_equal(s.download_U(no), ref[no - 1]), f"member {q} block {no}"
            # the last update's DU: a member that left the batch took it along, the others read
            # it from the batch's sweep storage
            assert np.array_equal(s.field(no, A.F_DU), ref_du[no - 1]), f"member {q} block {no} DU"
        assert np.array_equal(s.norms()[1], ref_rms)
        assert s.counters()[0] == iters[q]
    bt.close()


@pytest.mark.gpu
def test_member_refused_alone_and_destroy_leaves_batch():
    """A batch member cannot be stepped alone (OCFD_EARG); destroying a member takes it out of
    the batch, whose other members then step exactly as alone; after the batch is destroyed its
    members can be stepped alone again."""
    cases = _cases()
    S = [A.solver_from_case(cs, states=C.perturbed_state(cs, seed=30 + q, amp=1e-4)) for q, cs in enumerate(cases)]
    bt = A.Batch()
    bt.add(S[0]); bt.add(S[1]); bt.add(S[2])
    for call in (lambda: S[0].step(1), lambda: S[0].residual(), lambda: S[0].update_limit(), lambda: S[0].bc(),
                 lambda: S[0].halo(), lambda: S[0].init_flow(), lambda: S[0].autotune_lu(1)):
        with pytest.raises(A.OcfdError) as e:
            call()
        assert e.value.code == A.OCFD_EARG
    with pytest.raises(A.OcfdError):
        bt.add(S[1])                                   # already in the batch
    other = A.Batch()
    with pytest.raises(A.OcfdError):
        other.add(S[1])                                # a context is in one batch at most
    other.close()
    bt.step(1)
    S[0].norms(); S[0].forces()                        # monitors work on members
    S[0].close()                                       # destroyed while in the batch
    assert S[0] not in bt.members and len(bt.members) == 2
    bt.step(2)
    for q in (1, 2):
        ref, ref_rms = _alone(cases[q], 30 + q, 3)
        for no in range(1, len(cases[q].blocks) + 1):
            assert np.array_equal(S[q].download_U(no), ref[no - 1]), f"member {q} block {no}"
    bt.close()
    S[1].step(1)                                       # released by the batch's destruction
    ref, _ = _alone(cases[1], 31, 4)
    assert np.array_equal(S[1].download_U(1), ref[0])


@pytest.mark.gpu
def test_membership_churn_reuses_sweep_storage():
    """Members join and leave between steps as serving requests do, a new member taking the
    sweep storage a departed one released, and the tile width changes in between (which lays the
    storage out again); every member's state still equals its case stepped alone."""
    cases = _cases() + [C.naca_cgrid(65, 25, 9, Ma=0.5, AoA=1.0, **CG),
                        C.cut_blocks(C.naca_cgrid(49, 17, 5, Ma=0.65, AoA=2.0, **CG), [[[25], [], []]])]
    seeds = [50 + q for q in range(len(cases))]
    S = [A.solver_from_case(cs, states=C.perturbed_state(cs, seed=sd, amp=1e-4)) for cs, sd in zip(cases, seeds)]
    done = [0] * len(S)
    bt = A.Batch()
    bt.set_graph(True)

    def step(n):
        bt.step(n)
        for q, s in enumerate(S):
            if s in bt.members:
                done[q] += n

    bt.add(S[0]); bt.add(S[1]); bt.add(S[2])
    step(2)
    bt.remove(S[1])
    bt.add(S[3])                     # the first member laid out after a release
    step(2)
    bt.remove(S[0])
    bt.add(S[4])
    step(1)
    bt.set_lu_tile(4)                # every region laid out again for 4x16 tiles
    step(2)
    bt.add(S[1])                     # back in, after stepping nowhere meanwhile
    step(1)
    for q, (cs, s) in enumerate(zip(cases, S)):
        ref, ref_rms = _alone(cs, seeds[q], done[q])
        for no in range(1, len(cs.blocks) + 1):
            assert np.array_equal(s.download_U(no), ref[no - 1]), f"member {q} block {no} after {done[q]} iterations"
    assert done == [4, 3, 8, 6, 4]
    bt.close()
