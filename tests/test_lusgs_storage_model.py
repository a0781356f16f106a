"""CPU model of the pipelined LU-SGS sweep storage (ocfd_lusgs.hip, ocfd_dev.h lane_pos/pos_lane).

The team sweep of a tile reads its upstream tile's DU across the tile edge and waits while the
element holds the sentinel; the pack puts the sentinel back before each pair of sweeps, for a full
tile only into the leading lines of each step (where lane_pos puts the edge lanes), for a ragged
tile into every position.  The update reads DU by position, one 128-byte line per step and pass.
This model restates the task layout (lusgs_tasks) and the kernels' index arithmetic in Python and
checks, on ragged multi-block shapes and every tile width:
  - lane_pos is a permutation of the 64 lanes and pos_lane its inverse;
  - the upstream element a remote lane reads holds the DU of the cell it needs (same cell, the
    producer's storage step of that cell);
  - every element a remote lane waits on receives the sentinel from the pack;
  - the update's passes read every live cell's DU exactly once, at that cell's element.
No GPU and no oracle: the kernels' own parity against the hyperplane order is checked on the GPU
(test_lu_schedules_agree)."""
import random

import pytest

NT = 64
TILES = (4, 8, 16)


def lane_pos(TJ, lane):
    TK = NT // TJ
    jl, kl = lane % TJ, lane // TJ
    if jl == TJ - 1:
        return kl
    if jl == 0:
        return TK + kl
    if kl == TK - 1:
        return 2 * TK + (jl - 1)
    if kl == 0:
        return 2 * TK + (TJ - 2) + (jl - 1)
    return 2 * TK + 2 * (TJ - 2) + (kl - 1) * (TJ - 2) + (jl - 1)


def pos_lane(TJ, pos):
    TK = NT // TJ
    if pos < TK:
        jl, kl = TJ - 1, pos
    elif pos < 2 * TK:
        jl, kl = 0, pos - TK
    elif pos < 2 * TK + (TJ - 2):
        kl, jl = TK - 1, pos - 2 * TK + 1
    elif pos < 2 * TK + 2 * (TJ - 2):
        kl, jl = 0, pos - 2 * TK - (TJ - 2) + 1
    else:
        r = pos - 2 * TK - 2 * (TJ - 2)
        kl, jl = r // (TJ - 2) + 1, r % (TJ - 2) + 1
    return jl + TJ * kl


def tasks(blocks, TJ, forward):
    """lusgs_tasks: tiles of each block, storage in forward tile order, offsets in steps."""
    TK = NT // TJ
    out, pbase = [], 0
    for bi, (nx, ny, nz) in enumerate(blocks):
        nj, nk = ny - 1, nz - 1
        NJt, NKt = -(-nj // TJ), -(-nk // TK)
        sj = lambda J: min(TJ, nj - J * TJ)
        sk = lambda K: min(TK, nk - K * TK)
        pofs = {}
        for K in range(NKt):
            for J in range(NJt):
                pofs[(J, K)] = pbase
                pbase += (nx - 1) + sj(J) + sk(K) - 2
        for K in range(NKt):
            for J in range(NJt):
                t = dict(blk=bi, j0=1 + J * TJ, k0=1 + K * TK, tj=sj(J), tk=sk(K), pofs=pofs[(J, K)])
                Ju, Ku = (J - 1, K - 1) if forward else (J + 1, K + 1)
                if 0 <= Ju < NJt:
                    t.update(pofs_j=pofs[(Ju, K)], tj_up=sj(Ju))
                if 0 <= Ku < NKt:
                    t.update(pofs_k=pofs[(J, Ku)], tk_up=sk(Ku))
                out.append(t)
    return out


def sentinel_elements(blocks, TJ):
    """(tile storage offset, step, position) the pack writes the sentinel into (either DU array)."""
    TK = NT // TJ
    nsent = (2 * TK + 2 * (TJ - 2) + 15) // 16 * 16
    out = set()
    for U in tasks(blocks, TJ, True):
        nxm = blocks[U['blk']][0] - 1
        nsteps = nxm + U['tj'] + U['tk'] - 2
        npos = nsent if (U['tj'] == TJ and U['tk'] == TK) else NT
        for pos in range(npos):
            lane = pos_lane(TJ, pos)
            jl, kl = lane % TJ, lane // TJ
            for q in range(nsteps):
                if 1 <= q - jl - kl + 1 <= nxm:
                    out.add((U['pofs'], q, pos))
    return out


def producer_element(blocks, TJ, cell):
    """Storage element (tile offset, forward step, position) holding the DU of cell (blk, i, j, k)."""
    b, i, j, k = cell
    for U in tasks(blocks, TJ, True):
        if U['blk'] == b and U['j0'] <= j < U['j0'] + U['tj'] and U['k0'] <= k < U['k0'] + U['tk']:
            jl, kl = j - U['j0'], k - U['k0']
            return U['pofs'], i - 1 + jl + kl, lane_pos(TJ, jl + TJ * kl)
    raise AssertionError(cell)


def remote_reads(blocks, TJ, forward):
    """(element read, cell whose DU it must hold) for every wait of a remote lane."""
    for T in tasks(blocks, TJ, forward):
        nx, ny, nz = blocks[T['blk']]
        nxm = nx - 1
        nsteps = nxm + T['tj'] + T['tk'] - 2
        for D in (1, 2):
            for lane in range(NT):
                jl, kl = lane % TJ, lane // TJ
                live = jl < T['tj'] and kl < T['tk']
                jo = jl if forward else T['tj'] - 1 - jl
                ko = kl if forward else T['tk'] - 1 - kl
                j, k = T['j0'] + jl, T['k0'] + kl
                if D == 1:
                    has_d, inside = (j != 1 if forward else j != ny - 1), jo > 0
                else:
                    has_d, inside = (k != 1 if forward else k != nz - 1), ko > 0
                if not (live and has_d and not inside):
                    continue
                if D == 1:
                    pr, lane_r = T['pofs_j'], (T['tj_up'] - 1 if forward else 0) + TJ * kl
                    s0 = T['tj_up'] - 1 if forward else nxm + T['tk'] - 2
                else:
                    pr, lane_r = T['pofs_k'], jl + TJ * (T['tk_up'] - 1 if forward else 0)
                    s0 = T['tk_up'] - 1 if forward else nxm + T['tj'] - 2
                i0 = 1 - jo - ko if forward else nxm + jo + ko
                d = 1 if forward else -1
                for q in range(nsteps):
                    i = i0 + d * q
                    if 1 <= i <= nxm:
                        up = (j - d, k) if D == 1 else (j, k - d)
                        yield (pr, s0 + d * q, lane_pos(TJ, lane_r)), (T['blk'], i, up[0], up[1])


SHAPES = [[(17, 17, 9)], [(33, 20, 13), (9, 5, 18)], [(65, 25, 9)], [(12, 40, 7), (20, 3, 3)], [(9, 65, 17)]]
random.seed(7)
SHAPES += [[(random.randint(3, 30), random.randint(3, 40), random.randint(3, 40)) for _ in range(random.randint(1, 3))]
           for _ in range(3)]


@pytest.mark.parametrize("TJ", TILES)
def test_lane_positions_are_a_permutation(TJ):
    P = [lane_pos(TJ, l) for l in range(NT)]
    assert sorted(P) == list(range(NT))
    assert all(pos_lane(TJ, P[l]) == l for l in range(NT))


def test_edge_lanes_share_few_lines():
    # 8x8: the j-edge column across which a downstream tile reads fits one 128-byte line of a
    # step's field, the k-edge row two (row order: four and one)
    TJ = 8
    col = {lane_pos(TJ, TJ - 1 + TJ * k) // 16 for k in range(8)}
    row = {lane_pos(TJ, j + TJ * 7) // 16 for j in range(8)}
    assert len(col) == 1 and len(row) == 2


@pytest.mark.parametrize("TJ", TILES)
@pytest.mark.parametrize("shape", range(len(SHAPES)))
def test_remote_reads_find_the_upstream_cell_and_its_sentinel(TJ, shape):
    blocks = SHAPES[shape]
    sent = sentinel_elements(blocks, TJ)
    n = 0
    for forward in (True, False):
        for elem, cell in remote_reads(blocks, TJ, forward):
            assert elem == producer_element(blocks, TJ, cell), (forward, cell)
            assert elem in sent, (forward, cell)
            n += 1
    assert n > 0


@pytest.mark.parametrize("TJ", TILES)
@pytest.mark.parametrize("nxm,tj,tk", [(40, None, None), (33, -1, 0), (7, 2, 1), (64, 3, 2)])
def test_update_passes_read_every_cell_once(TJ, nxm, tj, tk):
    TK = NT // TJ
    tj = TJ if tj is None else (TJ + tj if tj < 0 else tj)
    tk = TK if tk is None else (TK + tk if tk <= 0 else tk)
    PK_W, UP_ROWS = 32, 16
    nsteps = nxm + tj + tk - 2
    seen = set()
    for i0 in range(1, nxm + 1, PK_W):
        W = min(PK_W, nxm - i0 + 1)
        for pos0 in range(0, NT, UP_ROWS):
            for slot in range((PK_W + TJ + TK - 2) * UP_ROWS):
                l, qo = slot % UP_ROWS, slot // UP_ROWS
                lane = pos_lane(TJ, pos0 + l)
                jl, kl = lane % TJ, lane // TJ
                ci = qo - (jl + kl)
                if jl < tj and kl < tk and 0 <= ci < W:
                    q, cell = i0 - 1 + qo, (i0 + ci, jl, kl)
                    assert 0 <= q < nsteps and q == cell[0] - 1 + jl + kl
                    assert cell not in seen
                    seen.add(cell)
    assert len(seen) == nxm * tj * tk
