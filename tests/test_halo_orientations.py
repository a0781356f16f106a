"""Block-interface halo for every orientation of a face link.

update_buffer_onemesh copies LAP = 4 layers of a block's interior cells into the ghost cells of
its neighbour, with the axis permutation L1..L3 and the directions that Convert_bc derives from
the Gridgen ranges (Umessage_send_mpi, sub_update_buffer_mpi.f90:42-141; Convert_bc,
sub_convert_inp.f90:131-247).  The check here does not use those formulas: two blocks split a
box at x = 0.5, the second block's node arrays are the box's physical axes permuted and reversed
by one of the 24 rotations of the cube, and every cell holds a code of its physical position
(cell index along x counted through both blocks, y and z cell indices, component).  After the
halo, every ghost cell within LAP layers of the interface must hold the code of the cell of the
other block that occupies its physical position: an explicit index map from the geometry itself.
The oracle is held to it on the CPU, the device path on the GPU.
"""
import itertools

import numpy as np
import pytest

import ocfd_amd as A
from ocfd_amd import cases as C
import oracle

NA, NB, NY, NZ = 6, 7, 8, 9          # nodes: block A along x, block B along x, shared y and z
LAP = 4


def rotations():
    """The 24 orientation-preserving signed permutations: (perm, flips), array axis a of the
    second block holds physical axis perm[a], reversed when flips[a]."""
    out = []
    for perm in itertools.permutations(range(3)):
        inv = sum(1 for a in range(3) for b in range(a + 1, 3) if perm[a] > perm[b])
        for flips in itertools.product((False, True), repeat=3):
            if (-1) ** inv * (-1) ** sum(flips) == 1:
                out.append((perm, flips))
    assert len(out) == 24
    return out


def _orient(phys, perm, flips):
    a = np.transpose(phys, perm)
    for ax in range(3):
        if flips[ax]:
            a = np.flip(a, axis=ax)
    return np.ascontiguousarray(a)


def two_blocks(perm, flips):
    xa, xb = np.linspace(0.0, 0.5, NA), np.linspace(0.5, 1.0, NB)
    y, z = np.linspace(0.0, 1.0, NY) ** 1.1, np.linspace(0.0, 1.0, NZ) ** 1.2
    A_xyz = np.meshgrid(xa, y, z, indexing="ij")
    B_xyz = [_orient(q, perm, flips) for q in np.meshgrid(xb, y, z, indexing="ij")]
    dA = (NA, NY, NZ)
    dB = B_xyz[0].shape
    n_phys = (NB, NY, NZ)
    a0 = perm.index(0)
    faceB = a0 + (4 if flips[a0] else 1)

    def img(a, node):                 # array node index of physical node `node` along array axis a
        return n_phys[perm[a]] - node + 1 if flips[a] else node

    # A face 4 -> B: A's tangential y is positive, z negative (full_face)
    sgnA = {1: 1, 2: -1}
    tgt = [0] * 6
    for a in range(3):
        if a == a0:
            v = img(a, 1)
            tgt[2 * a] = tgt[2 * a + 1] = v
        else:
            ph = perm[a]
            tgt[2 * a], tgt[2 * a + 1] = sgnA[ph] * img(a, 1), sgnA[ph] * img(a, n_phys[ph])
    subsA = [C.Subface(C.full_face(dA, 4), C.BC_INNER, tuple(tgt) + (2,))]
    subsA += [C.Subface(C.full_face(dA, f), C.BC_SYMMETRY) for f in (1, 2, 3, 5, 6)]
    # B's interface face -> A face 4
    srcB = C.full_face(dB, faceB)
    tgtA = [0] * 6
    tgtA[0] = tgtA[1] = NA
    for t in range(3):
        if t == a0:
            continue
        sg = 1 if srcB[2 * t] > 0 else -1
        ph = perm[t]

        def phys_node(idx, t=t, ph=ph):
            return n_phys[ph] - idx + 1 if flips[t] else idx
        tgtA[2 * ph], tgtA[2 * ph + 1] = sg * phys_node(1), sg * phys_node(dB[t])
    subsB = [C.Subface(srcB, C.BC_INNER, tuple(tgtA) + (1,))]
    subsB += [C.Subface(C.full_face(dB, f), C.BC_SYMMETRY) for f in range(1, 7) if f != faceB]
    params = C.make_params(Ma=0.5, If_viscous=0, Iflag_Scheme=5, Iflag_Flux=4, CFL=1.0)
    blocks = [C.Block(*A_xyz, subsA), C.Block(*B_xyz, subsB)]
    return C.Case("halo_pair", blocks, params), dB, a0


def code(X, q, r, m):
    return 1.0e6 * m + X + 100.0 * q + 10000.0 * r


def coded_states(perm, flips, dB):
    """U over the LAP box [m, i+3, j+3, k+3] of both blocks: interior cells coded, the rest -1."""
    nv = 5
    UA = np.full((nv, NA + 7, NY + 7, NZ + 7), -1.0)
    for m in range(nv):
        i, j, k = np.meshgrid(np.arange(1, NA), np.arange(1, NY), np.arange(1, NZ), indexing="ij")
        UA[m, 4:NA + 3, 4:NY + 3, 4:NZ + 3] = code(i, j, k, m)
    UB = np.full((nv,) + tuple(n + 7 for n in dB), -1.0)
    c = np.meshgrid(*[np.arange(1, n) for n in dB], indexing="ij")
    phys = [None] * 3
    n_phys = (NB, NY, NZ)
    for a in range(3):
        ph = perm[a]
        phys[ph] = n_phys[ph] - c[a] if flips[a] else c[a]
    for m in range(nv):
        UB[m, 4:dB[0] + 3, 4:dB[1] + 3, 4:dB[2] + 3] = code(NA - 1 + phys[0], phys[1], phys[2], m)
    return UA, UB


def check_ghosts(UA, UB, perm, flips, dB, a0):
    """Every ghost cell within LAP layers of the interface (tangential cells 1..n-1) against the
    code of the physical cell it overlaps."""
    bad = 0
    # A: cells i = NA .. NA+3 continue into B
    i, j, k = np.meshgrid(np.arange(NA, NA + LAP), np.arange(1, NY), np.arange(1, NZ), indexing="ij")
    for m in range(5):
        bad += int(np.count_nonzero(UA[m][i + 3, j + 3, k + 3] != code(i, j, k, m)))
    # B: LAP layers beyond its interface face, array cells along a0
    rng = [np.arange(1, n) for n in dB]
    rng[a0] = np.arange(dB[a0], dB[a0] + LAP) if flips[a0] else np.arange(1 - LAP, 1)
    c = np.meshgrid(*rng, indexing="ij")
    n_phys = (NB, NY, NZ)
    phys = [None] * 3
    for a in range(3):
        ph = perm[a]
        phys[ph] = n_phys[ph] - c[a] if flips[a] else c[a]
    for m in range(5):
        got = UB[m][c[0] + 3, c[1] + 3, c[2] + 3]
        bad += int(np.count_nonzero(got != code(NA - 1 + phys[0], phys[1], phys[2], m)))
    return bad


def _as_oracle(U):
    return np.ascontiguousarray(U.transpose(0, 3, 2, 1))       # [m, k, j, i]


@pytest.mark.parametrize("rot", range(24))
def test_oracle_halo_every_orientation(rot):
    perm, flips = rotations()[rot]
    case, dB, a0 = two_blocks(perm, flips)
    UA, UB = coded_states(perm, flips, dB)
    o = oracle.Oracle(case)
    o.init_flow()
    for blk, U in ((1, UA), (2, UB)):
        for m in range(5):
            o.set(blk, f"U{m + 1}", _as_oracle(U)[m])
    o.halo()
    gA = np.stack([o.get(1, f"U{m + 1}").transpose(2, 1, 0) for m in range(5)])
    gB = np.stack([o.get(2, f"U{m + 1}").transpose(2, 1, 0) for m in range(5)])
    inc = A.inc_tables(case)
    print(f"\n[perm {perm} flips {flips}] L = {inc[0][0][18:21].tolist()} / {inc[1][0][18:21].tolist()}")
    assert check_ghosts(gA, gB, perm, flips, dB, a0) == 0


@pytest.mark.gpu
def test_device_halo_every_orientation():
    for rot, (perm, flips) in enumerate(rotations()):
        print(f"\n[orientation {rot}: perm {perm} flips {flips}]", flush=True)
        case, dB, a0 = two_blocks(perm, flips)
        UA, UB = coded_states(perm, flips, dB)
        s = A.solver_from_case(case)
        s.upload_U(1, UA)
        s.upload_U(2, UB)
        s.halo()
        assert check_ghosts(s.download_U(1), s.download_U(2), perm, flips, dB, a0) == 0, rot
        s.close()
