"""Synthetic OpenCFD-EC cases: structured grids, Gridgen-style boundary tables and control.ec values.

A case is what the reference reads from disk: node coordinates of every block (``Mesh3d.dat``,
PLOT3D layout read by ``read_main_mesh``, sub_IO.f90:4-93), the Gridgen ``bc3d.inp`` subface table
(converted by ``convert_inp_inc``, sub_convert_inp.f90:25-247) and the ``&control_ec`` namelist
(``read_parameter_ec``, sub_read_parameter.f90:109-134).  Grids are generated here because the
reference ships no meshes.

Index convention: block arrays are numpy arrays indexed ``[i-1, j-1, k-1]`` over node indices
``1..nx, 1..ny, 1..nz``; they are written to disk in Fortran order, like the reference reads them.
"""
from __future__ import annotations

import dataclasses
import math
import os
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

# ---------------------------------------------------------------------------------------------
# constants mirroring sub_modules.f90:19-47 (ids used in control.ec and bc3d.inp)
BC_WALL, BC_SYMMETRY, BC_FARFIELD, BC_INFLOW, BC_OUTFLOW = 2, 3, 4, 5, 6
BC_EXTRAPOLATE = 401
BC_INNER, BC_PERIODIC_L, BC_PERIODIC_R = -1, -2, -3

SCHEME = dict(UD1=0, NND2=1, UD3=2, MUSCL2U=3, MUSCL2C=4, MUSCL3=5, OMUSCL2=6, WENO5=7, UD5=8,
              WENO7=9, CD2=20)
FLUX = dict(SW=1, HLL=2, HLLC=3, ROE=4, VANLEER=5, AUSM=6)


@dataclasses.dataclass
class Subface:
    """One Gridgen ``bc3d.inp`` subface record.

    ``rng`` is the signed node range (ib, ie, jb, je, kb, ke) exactly as written in the file;
    ``link`` is the signed linked range plus the linked block number (1-based) for inner faces.
    """
    rng: Tuple[int, int, int, int, int, int]
    bc: int
    link: Optional[Tuple[int, int, int, int, int, int, int]] = None


@dataclasses.dataclass
class Block:
    x: np.ndarray
    y: np.ndarray
    z: np.ndarray
    subfaces: List[Subface]

    @property
    def dims(self) -> Tuple[int, int, int]:
        return tuple(int(n) for n in self.x.shape)

    @property
    def ncell(self) -> int:
        nx, ny, nz = self.dims
        return (nx - 1) * (ny - 1) * (nz - 1)


@dataclasses.dataclass
class Case:
    name: str
    blocks: List[Block]
    params: Dict[str, object]

    @property
    def ncell(self) -> int:
        return sum(b.ncell for b in self.blocks)


# defaults of set_default_parameter (sub_read_parameter.f90:30-106), restricted to what the
# hot path reads.  Values are the *namelist* values (before set_const_para scaling).
DEFAULT_PARAMS: Dict[str, object] = dict(
    Ma=1.0, Re=1000.0, gamma=1.4, AoA=0.0, AoS=0.0, P_outlet=-1.0, t_end=100.0,
    Kstep_save=1000, Iflag_turbulence_model=0, Iflag_init=0, If_viscous=1, Iflag_local_dt=1,
    dt_global=0.01, CFL=1.0, dtmax=10.0, dtmin=1e-9, Time_Method=0, If_Residual_smoothing=0,
    w_LU=1.0, If_dtime_mesh=1, Iflag_Scheme=5, Iflag_Flux=5, IFlag_Reconstruction=0,
    Mesh_File_Format=0, Kstep_show=1, Num_Mesh=1, T_inf=288.15, Twall=-1.0, Kt_inf=1e-5,
    Wt_inf=0.01, NUM_THREADS=1, MUT_MAX=-1.0, Bound_Scheme=4, Ref_S=1.0, Ref_L=1.0,
    Centroid=(0.0, 0.0, 0.0), Cood_Y_UP=1, IFLAG_LIMIT_FLOW=1, PrL=0.7, PrT=0.9, Ldmin=1e-6,
    Ldmax=1000.0, Lpmin=1e-6, Lpmax=1000.0, Lumax=1000.0, LSAmax=1000.0, Periodic_dX=0.0,
    Periodic_dY=0.0, Periodic_dZ=0.0, IF_Scheme_Positivity=1,
)

# namelist variables of &control_ec that this module writes (subset of sub_read_parameter.f90:114-129)
_NML_KEYS = [
    "Ma", "Re", "AoA", "AoS", "P_outlet", "t_end", "gamma", "PrL", "PrT", "Kstep_save",
    "Iflag_turbulence_model", "Iflag_init", "If_viscous", "Iflag_local_dt", "dt_global", "CFL",
    "dtmax", "dtmin", "Time_Method", "If_Residual_smoothing", "w_LU", "If_dtime_mesh",
    "Iflag_Scheme", "Iflag_Flux", "IFlag_Reconstruction", "Mesh_File_Format", "Kstep_show",
    "Num_Mesh", "T_inf", "Twall", "Kt_inf", "Wt_inf", "NUM_THREADS", "MUT_MAX", "Bound_Scheme",
    "Ref_S", "Ref_L", "Centroid", "Cood_Y_UP", "IFLAG_LIMIT_FLOW", "Ldmin", "Ldmax", "Lpmin",
    "Lpmax", "Lumax", "LSAmax", "Periodic_dX", "Periodic_dY", "Periodic_dZ",
    "IF_Scheme_Positivity",
]


def make_params(**kw) -> Dict[str, object]:
    p = dict(DEFAULT_PARAMS)
    for k, v in kw.items():
        if k not in p:
            raise KeyError(f"unknown control parameter {k}")
        p[k] = v
    return p


# ---------------------------------------------------------------------------------------------
# writers (formats read by sub_IO.f90:4-93, sub_convert_inp.f90:25-122, sub_read_parameter.f90:132)

def _fortran_real(v: float) -> str:
    return repr(float(v)).replace("e", "d") if "e" in repr(float(v)) else repr(float(v)) + "d0"


def _fortran_value(v) -> str:
    if isinstance(v, (tuple, list)):
        return ", ".join(_fortran_value(x) for x in v)
    if isinstance(v, (bool, np.bool_)):
        return ".true." if v else ".false."
    if isinstance(v, (int, np.integer)):
        return str(int(v))
    return _fortran_real(float(v))


def control_ec_text(params: Dict[str, object]) -> str:
    lines = ["&control_ec"]
    for k in _NML_KEYS:
        lines.append(f"  {k}={_fortran_value(params[k])}")
    lines.append("/")
    return "\n".join(lines) + "\n"


def _fortran_record(f, payload: bytes) -> None:
    n = np.array([len(payload)], dtype="<i4").tobytes()
    f.write(n)
    f.write(payload)
    f.write(n)


def write_mesh3d(path: str, blocks: Sequence[Block]) -> None:
    """Unformatted sequential PLOT3D multi-block file (Mesh_File_Format=0)."""
    with open(path, "wb") as f:
        _fortran_record(f, np.array([len(blocks)], dtype="<i4").tobytes())
        dims = np.array([b.dims for b in blocks], dtype="<i4").reshape(-1)
        _fortran_record(f, dims.tobytes())
        for b in blocks:
            payload = b"".join(np.asarray(a, dtype="<f8").tobytes(order="F") for a in (b.x, b.y, b.z))
            _fortran_record(f, payload)


def bc3d_inp_text(blocks: Sequence[Block]) -> str:
    out = [" 1", f" {len(blocks)}"]
    for m, b in enumerate(blocks, start=1):
        nx, ny, nz = b.dims
        out.append(f" {nx} {ny} {nz}")
        out.append(f" Block {m}")
        out.append(f" {len(b.subfaces)}")
        for s in b.subfaces:
            out.append(" ".join(f"{v:6d}" for v in (*s.rng, s.bc)))
            if s.bc < 0:
                assert s.link is not None
                out.append(" ".join(f"{v:6d}" for v in s.link))
    return "\n".join(out) + "\n"


def write_case(case: Case, directory: str) -> None:
    os.makedirs(directory, exist_ok=True)
    write_mesh3d(os.path.join(directory, "Mesh3d.dat"), case.blocks)
    with open(os.path.join(directory, "bc3d.inp"), "w") as f:
        f.write(bc3d_inp_text(case.blocks))
    p = dict(case.params)
    p["Mesh_File_Format"] = 0
    with open(os.path.join(directory, "control.ec"), "w") as f:
        f.write(control_ec_text(p))


# ---------------------------------------------------------------------------------------------
# face helpers

def full_face(dims: Tuple[int, int, int], face: int) -> Tuple[int, int, int, int, int, int]:
    """Signed node range of a whole block face (face 1..6 = i-, j-, k-, i+, j+, k+).

    The two tangential dims get opposite signs so the Gridgen dimension matching
    (sub_convert_inp.f90:226-230) is unambiguous: the first tangential dim is positive,
    the second negative.
    """
    nx, ny, nz = dims
    n = (nx, ny, nz)
    d = (face - 1) % 3
    lo = face <= 3
    rng = []
    tang = [a for a in range(3) if a != d]
    for a in range(3):
        if a == d:
            v = 1 if lo else n[a]
            rng += [v, v]
        elif a == tang[0]:
            rng += [1, n[a]]
        else:
            rng += [-1, -n[a]]
    return tuple(rng)


def sub_range(dims, face, lo_hi: Dict[int, Tuple[int, int]]):
    """Signed range of part of a face: ``lo_hi`` maps tangential dim -> (start, end) node index."""
    rng = list(full_face(dims, face))
    d = (face - 1) % 3
    tang = [a for a in range(3) if a != d]
    for a, (s, e) in lo_hi.items():
        sign = 1 if a == tang[0] else -1
        rng[2 * a] = sign * s
        rng[2 * a + 1] = sign * e
    return tuple(rng)


# ---------------------------------------------------------------------------------------------
# C1: channel with a circular-arc bump (97x33x9 nodes by default)

def channel_bump(ni: int = 97, nj: int = 33, nk: int = 9, Lx: float = 3.0, Lz: float = 0.3,
                 bump: float = 0.1, inlet_bc: int = BC_FARFIELD, outlet_bc: int = BC_FARFIELD, **params) -> Case:
    """Single-block 3D inviscid channel, 10%-chord circular-arc bump on x in [1,2] of the lower wall.

    Lower and upper walls are slip walls (bc=2 with If_viscous=0), z is periodic through a
    self-link (bc -2/-3) with Periodic_dZ=Lz, inflow/outflow are far-field (bc=4) faces driven by
    P_outlet = 1/(gamma Ma^2); ``inlet_bc``/``outlet_bc`` select another condition for the i=1 and
    i=ni faces (5 inflow, 6 outflow, 401 extrapolation).
    """
    R = (0.25 + bump * bump) / (2 * bump)          # circle through (1,0),(2,0),(1.5,bump)
    x1 = np.linspace(0.0, Lx, ni)
    yb = np.zeros_like(x1)
    m = (x1 >= 1.0) & (x1 <= 2.0)
    yb[m] = np.sqrt(np.maximum(R * R - (x1[m] - 1.5) ** 2, 0.0)) - (R - bump)
    yb = np.maximum(yb, 0.0)
    eta = np.linspace(0.0, 1.0, nj)
    zz = np.linspace(0.0, Lz, nk)
    x = np.empty((ni, nj, nk)); y = np.empty_like(x); z = np.empty_like(x)
    x[:] = x1[:, None, None]
    y[:] = (yb[:, None] + (1.0 - yb[:, None]) * eta[None, :])[:, :, None]
    z[:] = zz[None, None, :]
    dims = (ni, nj, nk)
    subs = [
        Subface(full_face(dims, 1), inlet_bc),
        Subface(full_face(dims, 4), outlet_bc),
        Subface(full_face(dims, 2), BC_WALL),
        Subface(full_face(dims, 5), BC_WALL),
        Subface(full_face(dims, 3), BC_PERIODIC_L, full_face(dims, 6) + (1,)),
        Subface(full_face(dims, 6), BC_PERIODIC_R, full_face(dims, 3) + (1,)),
    ]
    Ma = params.pop("Ma", 0.5)
    gamma = params.get("gamma", 1.4)
    defaults = dict(Ma=Ma, If_viscous=0, Iflag_Scheme=5, Iflag_Flux=4, Time_Method=0, CFL=5.0,
                    Iflag_local_dt=1, dt_global=1.0, t_end=2000.0, Kstep_show=100000,
                    Kstep_save=100000, P_outlet=1.0 / (gamma * Ma * Ma), Periodic_dZ=Lz)
    defaults.update(params)
    return Case("channel_bump", [Block(x, y, z, subs)], make_params(**defaults))


# ---------------------------------------------------------------------------------------------
# NACA 4-digit C-grid (C2 and the serving requests)

def naca4(m: float, p: float, t: float, n: int):
    """Upper and lower surfaces (closed trailing edge) at n cosine-spaced chord stations."""
    beta = np.linspace(0.0, math.pi, n)
    xc = 0.5 * (1.0 - np.cos(beta))
    yt = 5.0 * t * (0.2969 * np.sqrt(xc) - 0.1260 * xc - 0.3516 * xc ** 2 + 0.2843 * xc ** 3
                    - 0.1036 * xc ** 4)
    if m > 0.0 and p > 0.0:
        yc = np.where(xc < p, m / p ** 2 * (2 * p * xc - xc ** 2),
                      m / (1 - p) ** 2 * ((1 - 2 * p) + 2 * p * xc - xc ** 2))
        dyc = np.where(xc < p, 2 * m / p ** 2 * (p - xc), 2 * m / (1 - p) ** 2 * (p - xc))
    else:
        yc = np.zeros_like(xc); dyc = np.zeros_like(xc)
    th = np.arctan(dyc)
    xu, yu = xc - yt * np.sin(th), yc + yt * np.cos(th)
    xl, yl = xc + yt * np.sin(th), yc - yt * np.cos(th)
    yu[-1] = yl[-1] = yc[-1]
    xu[-1] = xl[-1] = 1.0
    return (xu, yu), (xl, yl)


def solve_stretch(h1: float, total: float, ncell: int) -> float:
    """Geometric ratio r with h1 (r^n - 1)/(r - 1) = total."""
    lo, hi = 1.0 + 1e-12, 2.0
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        s = h1 * (mid ** ncell - 1.0) / (mid - 1.0)
        if s > total:
            hi = mid
        else:
            lo = mid
    return 0.5 * (lo + hi)


def naca_cgrid(ni: int = 257, nj: int = 97, nk: int = 33, m: float = 0.0, p: float = 0.0,
               t: float = 0.12, span: float = 1.0, wall_spacing: float = 2e-6,
               farfield: float = 20.0, sweep_deg: float = 0.0, taper: float = 1.0, **params) -> Case:
    """Single-block C-grid around an extruded NACA 4-digit section.

    i runs wake(lower) -> lower surface -> LE -> upper surface -> wake(upper); j runs from the wall
    and wake cut to the far field; k is the span.  The j=1 wake cut is a self-link with reversed i
    (bc=-1), the airfoil is a no-slip wall (bc=2), the outer and outflow faces are far field
    (bc=4), the span ends are symmetry planes (bc=3).  The wall-normal distribution is geometric
    from ``wall_spacing`` with the ratio solved so the outer boundary sits at ``farfield`` chords.
    With ``sweep_deg``/``taper`` every spanwise station z is the section scaled by the local chord
    c(z) = 1 - (1 - taper) z / span and shifted by z tan(sweep) (a swept, tapered extruded wing).
    """
    nw = int(round((ni - 1) * 0.15625)) + 1          # wake nodes per side incl. the TE node
    na = (ni - 2 * (nw - 1) - 1) // 2 + 1             # nodes per surface incl. LE and TE
    assert 2 * (nw - 1) + 2 * (na - 1) + 1 == ni, "ni incompatible with the C-grid split"
    (xu, yu), (xl, yl) = naca4(m, p, t, na)
    # wake x stations: geometric from the TE spacing to x = 1 + farfield
    h_te = math.hypot(xu[-1] - xu[-2], yu[-1] - yu[-2])
    rw = solve_stretch(h_te, farfield, nw - 1)
    xw = 1.0 + h_te * (rw ** np.arange(nw) - 1.0) / (rw - 1.0)
    xw[-1] = 1.0 + farfield
    # inner curve (airfoil + wake line), i = 1..ni
    xi = np.concatenate([xw[::-1], xl[::-1][1:], xu[1:], xw[1:]])
    yi = np.concatenate([np.zeros(nw), yl[::-1][1:], yu[1:], np.zeros(nw - 1)])
    yi[nw - 1] = 0.0; yi[ni - nw] = 0.0
    # outer curve: straight wake lines y=-/+R and a half circle about the TE
    R = farfield
    xo = np.empty(ni); yo = np.empty(ni)
    xo[:nw] = xw[::-1]; yo[:nw] = -R
    xo[ni - nw:] = xw; yo[ni - nw:] = R
    # airfoil stations map by arc-length fraction onto angles -pi/2 -> -pi (lower) and pi -> pi/2
    def arcfrac(xs, ys):
        s = np.concatenate([[0.0], np.cumsum(np.hypot(np.diff(xs), np.diff(ys)))])
        return s / s[-1]
    fl = arcfrac(xl[::-1], yl[::-1])                  # TE -> LE
    fu = arcfrac(xu, yu)                              # LE -> TE
    th_l = -0.5 * math.pi - 0.5 * math.pi * fl
    th_u = math.pi - 0.5 * math.pi * fu
    xo[nw - 1:nw - 1 + na] = 1.0 + R * np.cos(th_l); yo[nw - 1:nw - 1 + na] = R * np.sin(th_l)
    xo[nw - 2 + na:ni - nw + 1] = 1.0 + R * np.cos(th_u); yo[nw - 2 + na:ni - nw + 1] = R * np.sin(th_u)
    # wall-normal distribution: s_j in [0,1], first step wall_spacing/farfield
    r = solve_stretch(wall_spacing / farfield, 1.0, nj - 1)
    s = (r ** np.arange(nj) - 1.0) / (r ** (nj - 1) - 1.0)
    X2 = xi[:, None] + s[None, :] * (xo - xi)[:, None]
    Y2 = yi[:, None] + s[None, :] * (yo - yi)[:, None]
    zz = np.linspace(0.0, span, nk)
    if sweep_deg == 0.0 and taper == 1.0:
        x = np.repeat(X2[:, :, None], nk, axis=2)
        y = np.repeat(Y2[:, :, None], nk, axis=2)
    else:
        c = 1.0 - (1.0 - taper) * zz / span
        x = X2[:, :, None] * c[None, None, :] + (zz * math.tan(math.radians(sweep_deg)))[None, None, :]
        y = Y2[:, :, None] * c[None, None, :]
    z = np.broadcast_to(zz[None, None, :], (ni, nj, nk)).copy()
    dims = (ni, nj, nk)
    iw0, iw1 = nw, ni - nw + 1                          # airfoil node range on j=1
    subs = [
        Subface(sub_range(dims, 2, {0: (1, iw0)}), BC_INNER,
                sub_range(dims, 2, {0: (ni, iw1)}) + (1,)),
        Subface(sub_range(dims, 2, {0: (iw0, iw1)}), BC_WALL),
        Subface(sub_range(dims, 2, {0: (iw1, ni)}), BC_INNER,
                sub_range(dims, 2, {0: (iw0, 1)}) + (1,)),
        Subface(full_face(dims, 5), BC_FARFIELD),
        Subface(full_face(dims, 1), BC_FARFIELD),
        Subface(full_face(dims, 4), BC_FARFIELD),
        Subface(full_face(dims, 3), BC_SYMMETRY),
        Subface(full_face(dims, 6), BC_SYMMETRY),
    ]
    defaults = dict(Ma=0.8, Re=6e6, AoA=1.25, If_viscous=1, Iflag_turbulence_model=2,
                    Iflag_Scheme=7, Iflag_Flux=3, Time_Method=0, CFL=5.0, Iflag_local_dt=1,
                    dt_global=1.0, t_end=1e9, Kstep_show=100000, Kstep_save=100000)
    defaults.update(params)
    return Case("naca_cgrid", [Block(x, y, z, subs)], make_params(**defaults))


# ---------------------------------------------------------------------------------------------
# Block cutting into one-to-one multi-block grids (semantics of Block_cut_1.1.f90:58-70)

def _face_of(rng):
    a = [abs(v) for v in rng]
    for d in range(3):
        if a[2 * d] == a[2 * d + 1]:
            return d, a[2 * d]
    raise ValueError("subface range is not a face")


def _orientation(rng, link):
    """Map from source dim to (target dim, sign) following Convert_bc (sub_convert_inp.f90:145-246)."""
    s = [0 if rng[2 * k] == rng[2 * k + 1] else (1 if rng[2 * k] > 0 else -1) for k in range(3)]
    p = [0 if link[2 * k] == link[2 * k + 1] else (1 if link[2 * k] > 0 else -1) for k in range(3)]
    m = {}
    for k in range(3):
        for j in range(3):
            if s[k] == p[j]:
                m[k] = j
    return m


def cut_blocks(case: Case, cuts: Sequence[Sequence[Sequence[int]]]) -> Case:
    """Cut every block at interior node planes and rebuild a one-to-one Gridgen table.

    ``cuts[b][d]`` lists the node indices (1-based, strictly inside) where block ``b`` is cut
    along dim ``d``.  Physical subfaces are clipped to the children; inner links (including
    self-links such as a C-grid wake cut) are split so that each child subface maps onto exactly
    one child of the target block.
    """
    # children of every parent block
    kids = []          # per parent: list of (child_id, lo(3), hi(3)) in parent node indices
    newblocks: List[Block] = []
    for b, blk in enumerate(case.blocks):
        dims = blk.dims
        edges = [[1] + sorted(cuts[b][d]) + [dims[d]] for d in range(3)]
        lst = []
        for kk in range(len(edges[2]) - 1):
            for jj in range(len(edges[1]) - 1):
                for ii in range(len(edges[0]) - 1):
                    lo = (edges[0][ii], edges[1][jj], edges[2][kk])
                    hi = (edges[0][ii + 1], edges[1][jj + 1], edges[2][kk + 1])
                    sl = tuple(slice(lo[d] - 1, hi[d]) for d in range(3))
                    nb = Block(blk.x[sl].copy(), blk.y[sl].copy(), blk.z[sl].copy(), [])
                    newblocks.append(nb)
                    lst.append((len(newblocks), lo, hi))
        kids.append(lst)

    def child_at(parent, node, face_dim, face_side_hi):
        """Child of `parent` containing the face cell whose lower node corner is `node`."""
        for cid, lo, hi in kids[parent]:
            ok = True
            for d in range(3):
                if d == face_dim:
                    if (hi[d] if face_side_hi else lo[d]) != node[d]:
                        ok = False
                else:
                    if not (lo[d] <= node[d] < hi[d]):
                        ok = False
            if ok:
                return cid, lo, hi
        raise RuntimeError("no child owns this face cell")

    def local(rng_nodes, lo):
        return [v - lo[d] + 1 for d, v in enumerate(rng_nodes)]

    for b, blk in enumerate(case.blocks):
        for cid, lo, hi in kids[b]:
            child = newblocks[cid - 1]
            cdims = child.dims
            for face in range(1, 7):
                d = (face - 1) % 3
                at = lo[d] if face <= 3 else hi[d]
                tang = [a for a in range(3) if a != d]
                on_parent_face = (face <= 3 and at == 1) or (face > 3 and at == blk.dims[d])
                if not on_parent_face:
                    # internal cut face: one-to-one link to the neighbouring child, same orientation
                    nbr_node = list(lo)
                    nbr_node[d] = at
                    ncid, nlo, nhi = child_at(b, nbr_node, d, face > 3 and False or face <= 3)
                    nface = face + 3 if face <= 3 else face - 3
                    child.subfaces.append(Subface(full_face(cdims, face), BC_INNER,
                                                  full_face(newblocks[ncid - 1].dims, nface) + (ncid,)))
                    continue
                # clip the parent's subfaces on this face to the child's tangential extent
                for sf in blk.subfaces:
                    fd, fat = _face_of(sf.rng)
                    if fd != d or fat != at:
                        continue
                    a = [abs(v) for v in sf.rng]
                    lo_t = {t: max(min(a[2 * t], a[2 * t + 1]), lo[t]) for t in tang}
                    hi_t = {t: min(max(a[2 * t], a[2 * t + 1]), hi[t]) for t in tang}
                    if any(hi_t[t] <= lo_t[t] for t in tang):
                        continue
                    if sf.bc >= 0:
                        rng = sub_range(cdims, face, {t: (lo_t[t] - lo[t] + 1, hi_t[t] - lo[t] + 1) for t in tang})
                        child.subfaces.append(Subface(rng, sf.bc))
                        continue
                    # inner link: map clipped source cells to the target parent and split by target child
                    tb = sf.link[6] - 1
                    om = _orientation(sf.rng, sf.link)
                    ta = [abs(v) for v in sf.link[:6]]
                    tdim_fixed = [k for k in range(3) if ta[2 * k] == ta[2 * k + 1]][0]
                    tface_hi = ta[2 * tdim_fixed] != 1
                    # node map source -> target for tangential dims
                    def tmap(t, v):
                        k = om[t]
                        s0, s1 = sf.rng[2 * t], sf.rng[2 * t + 1]
                        t0, t1 = sf.link[2 * k], sf.link[2 * k + 1]
                        off = abs(v) - abs(s0)
                        step = 1 if abs(s1) >= abs(s0) else -1
                        tstep = 1 if abs(t1) >= abs(t0) else -1
                        return abs(t0) + (off * step) * tstep
                    # split the clipped source range at preimages of target child boundaries
                    brk = {t: {lo_t[t], hi_t[t]} for t in tang}
                    for tcid, tlo, thi in kids[tb]:
                        for t in tang:
                            k = om[t]
                            for edge in (tlo[k], thi[k]):
                                for v in range(lo_t[t], hi_t[t] + 1):
                                    if tmap(t, v) == edge:
                                        brk[t].add(v)
                    bt = {t: sorted(brk[t]) for t in tang}
                    for i0 in range(len(bt[tang[0]]) - 1):
                        for i1 in range(len(bt[tang[1]]) - 1):
                            seg = {tang[0]: (bt[tang[0]][i0], bt[tang[0]][i0 + 1]),
                                   tang[1]: (bt[tang[1]][i1], bt[tang[1]][i1 + 1])}
                            # target child owning the first cell of this piece
                            tnode = [0, 0, 0]
                            tnode[tdim_fixed] = ta[2 * tdim_fixed]
                            for t in tang:
                                a0, a1 = tmap(t, seg[t][0]), tmap(t, seg[t][1])
                                tnode[om[t]] = min(a0, a1)
                            tcid, tlo, thi = child_at(tb, tnode, tdim_fixed, tface_hi)
                            tdims = newblocks[tcid - 1].dims
                            src = sub_range(cdims, face, {t: (seg[t][0] - lo[t] + 1, seg[t][1] - lo[t] + 1)
                                                          for t in tang})
                            tgt = [0] * 6
                            tfix_local = ta[2 * tdim_fixed] - tlo[tdim_fixed] + 1
                            tgt[2 * tdim_fixed] = tgt[2 * tdim_fixed + 1] = tfix_local
                            for t in tang:
                                k = om[t]
                                sign = 1 if src[2 * t] > 0 else -1
                                tgt[2 * k] = sign * (tmap(t, seg[t][0]) - tlo[k] + 1)
                                tgt[2 * k + 1] = sign * (tmap(t, seg[t][1]) - tlo[k] + 1)
                            # the fixed dim's sign convention in the target follows Gridgen (positive)
                            child.subfaces.append(Subface(src, sf.bc, tuple(tgt) + (tcid,)))
    return Case(case.name + "_cut", newblocks, dict(case.params))


def seeded_cuts(n_nodes: int, parts: int, rng: np.random.Generator, lo: int = 17, hi: int = 161) -> List[int]:
    """Interior cut node indices splitting 1..n_nodes into `parts` pieces whose node counts lie in
    [lo, hi]: piece sizes are the even split perturbed by seeded noise, then clipped to the range."""
    cells = n_nodes - 1
    if not (parts * (lo - 1) <= cells <= parts * (hi - 1)):
        raise ValueError("no split of %d cells into %d pieces within [%d, %d] nodes" % (cells, parts, lo, hi))
    for _ in range(1000):
        w = rng.uniform(0.6, 1.4, parts)
        sz = np.floor(w / w.sum() * cells).astype(int)
        sz[-1] = cells - sz[:-1].sum()
        if np.all(sz >= lo - 1) and np.all(sz <= hi - 1):
            return [int(v) + 1 for v in np.cumsum(sz)[:-1]]
    raise RuntimeError("no admissible split found")


def multiblock_airfoil(seed: int = 3, scale: float = 1.0, **params) -> Case:
    """48-block uneven grid: the 577x193x129 C-grid around a NACA 4412 section (span 1) cut 6x4x2
    at seeded planes, every block dim within [17, 161] nodes (14,155,776 cells); SST, Re 3e6,
    Mach 0.7, MUSCL3 + AUSMPW (flux 6), CFL 0.1.  `scale` < 1 shrinks the parent for tests.

    CFL 0.1: from freestream the reference's SST diverges at CFL 1 to 5 on this grid (2e-6 wall
    spacing).  In a wall-adjacent cell the omega destruction term -beta rho omega^2 has a Jacobian
    about 60 times the LU-SGS diagonal there (the sweep treats the source explicitly), so omega
    overshoots below zero; the boundary condition is applied before force_vt_kw clamps it
    (NS_Time_advance, sub_time_advance.f90:6-23; wall_bound, sub_boundary.f90:405-454), so the
    ghost takes omega_g = 2 rho wt - U7 with the unclamped negative U7, which feeds back a larger
    omega, and |omega| squares every two iterations until it overflows (iteration 8 to 10).  At
    CFL 0.1 the k and omega residuals decay (the oracle over 100 iterations at 1/4 and 1/2 scale)."""
    def n(v):
        return v if scale == 1.0 else max(33, int(round((v - 1) * scale)) // 32 * 32 + 1)
    ni, nj, nk = n(577), n(193), n(129)
    args = dict(Iflag_turbulence_model=3, Re=3e6, Ma=0.7, Iflag_Scheme=5, Iflag_Flux=6, CFL=0.1)
    args.update(params)
    parent = naca_cgrid(ni, nj, nk, m=0.04, p=0.4, t=0.12, span=1.0, **args)
    rng = np.random.default_rng(seed)
    lo, hi = (17, 161) if scale == 1.0 else (5, 161)
    cuts = [[seeded_cuts(ni, 6, rng, lo, hi), seeded_cuts(nj, 4, rng, lo, hi), seeded_cuts(nk, 2, rng, lo, hi)]]
    case = cut_blocks(parent, cuts)
    case.name = "naca4412_48block"
    return case


def swept_wing(seed: int = 4, scale: float = 1.0) -> Case:
    """24-block swept, tapered wing: the 1025x257x233 C-grid around NACA 0012 extruded over a span of
    3 chords with 30 deg sweep and taper 0.5, wall spacing 1e-6, cut 4x3x2 at seeded planes
    (60,817,408 cells); SA, Mach 0.84, AoA 3 deg, Re 1.2e7, WENO5 + Roe.  The wall-normal ratio is
    solved for a 20-chord far field (a 1.1 ratio from 1e-6 over 256 cells would reach 4e5 chords)."""
    def n(v):
        return v if scale == 1.0 else max(33, int(round((v - 1) * scale)) // 32 * 32 + 1)
    ni, nj, nk = n(1025), n(257), n(233)
    parent = naca_cgrid(ni, nj, nk, span=3.0, wall_spacing=1e-6, sweep_deg=30.0, taper=0.5,
                        Iflag_turbulence_model=2, Re=1.2e7, Ma=0.84, AoA=3.0, Iflag_Scheme=7, Iflag_Flux=4)
    rng = np.random.default_rng(seed)
    cuts = [[seeded_cuts(ni, 4, rng, 5, 10 ** 6), seeded_cuts(nj, 3, rng, 5, 10 ** 6), seeded_cuts(nk, 2, rng, 5, 10 ** 6)]]
    case = cut_blocks(parent, cuts)
    case.name = "swept_wing_24block"
    return case


# ---------------------------------------------------------------------------------------------
# initial states

def freestream(params: Dict[str, object], nvar: int) -> np.ndarray:
    """Conserved freestream of init_flow_zero (sub_init.f90:324-354)."""
    g = float(params["gamma"]); Ma = float(params["Ma"])
    aoa = float(params["AoA"]) * math.pi / 180.0
    aos = float(params["AoS"]) * math.pi / 180.0
    if int(params["Cood_Y_UP"]) == 1:
        a_alfa, a_beta = aoa, aos
    else:
        a_alfa, a_beta = aos, aoa
    u0 = math.cos(a_alfa) * math.cos(a_beta)
    v0 = math.sin(a_alfa) * math.cos(a_beta)
    w0 = math.sin(a_beta)
    p0 = 1.0 / (g * Ma * Ma)
    q = [1.0, u0, v0, w0, p0 / (g - 1.0) + 0.5 * (u0 * u0 + v0 * v0 + w0 * w0)]
    if nvar == 6:
        q.append(5.0)
    elif nvar == 7:
        q += [10.0 * float(params["Kt_inf"]), float(params["Wt_inf"])]
    return np.array(q)


def nvar_of(params: Dict[str, object]) -> int:
    turb = int(params["Iflag_turbulence_model"]) if int(params["If_viscous"]) == 1 else 0
    return 6 if turb in (2, 21) else (7 if turb == 3 else 5)


def perturbed_state(case: Case, seed: int = 1, amp: float = 1e-3) -> List[np.ndarray]:
    """Interior conserved state per block, (nvar, nx-1, ny-1, nz-1) Fortran-ordered.

    Freestream times (1 + amp * noise) with seeded uniform noise in [-1, 1], so that every cell
    carries a non-trivial residual.  The energy is rebuilt from perturbed primitives so pressure
    stays positive.
    """
    rng = np.random.default_rng(seed)
    nv = nvar_of(case.params)
    q = freestream(case.params, nv)
    g = float(case.params["gamma"])
    out = []
    for b in case.blocks:
        nx, ny, nz = b.dims
        shp = (nx - 1, ny - 1, nz - 1)
        noise = rng.uniform(-1.0, 1.0, size=(nv,) + shp)
        d = q[0] * (1.0 + amp * noise[0])
        u = q[1] + amp * noise[1]
        v = q[2] + amp * noise[2]
        w = q[3] + amp * noise[3]
        p0 = (q[4] - 0.5 * q[0] * (q[1] ** 2 + q[2] ** 2 + q[3] ** 2)) * (g - 1.0)
        p = p0 * (1.0 + amp * noise[4])
        U = np.empty((nv,) + shp, order="F")
        U[0] = d; U[1] = d * u; U[2] = d * v; U[3] = d * w
        U[4] = p / (g - 1.0) + 0.5 * d * (u * u + v * v + w * w)
        for m in range(5, nv):
            U[m] = q[m] * (1.0 + amp * noise[m])
        out.append(np.asfortranarray(U))
    return out


def write_initial_state(path: str, states: Sequence[np.ndarray]) -> None:
    with open(path, "wb") as f:
        for U in states:
            f.write(np.asarray(U, dtype="<f8").tobytes(order="F"))
