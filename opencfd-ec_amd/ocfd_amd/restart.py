"""Restart and wall-distance files in the reference's on-disk formats.

flow3d.dat  one unformatted sequential record per block: d, u, v, w, T over i=0..nx, j=0..ny,
            k=0..nz (Fortran order, variable slowest) with T = (E - 0.5 d |u|^2) / (Cv d)
            (output_flow, sub_IO.f90:350-484)
SA3d.dat    one record per block: U(6) over the same range (NVAR = 6)
SST3d.dat   one record per block: U(6), U(7) (NVAR = 7)
Step_mess.dat  "Kstep tt" (list-directed), then a comment line
wall_dist.dat  one record per block: dw(1:nx-1, 1:ny-1, 1:nz-1) (write_dw, sub_IO.f90:573-644)

read_restart converts back to conserved variables the way read_flow_data does
(sub_IO.f90:171-338): U1 = d, U2..4 = d u, U5 = Cv d T + 0.5 d (u^2 + v^2 + w^2).
Arrays passed in and returned use the download layout of ocfd_download_U,
(NVAR, nx+7, ny+7, nz+7) indexed [m, i+3, j+3, k+3].
"""
from __future__ import annotations

import os
from typing import List, Sequence, Tuple

import numpy as np


def _record(f, payload: bytes) -> None:
    n = np.array([len(payload)], dtype="<i4").tobytes()
    f.write(n + payload + n)


def _records(path: str) -> List[bytes]:
    out = []
    with open(path, "rb") as f:
        buf = f.read()
    pos = 0
    while pos < len(buf):
        (n,) = np.frombuffer(buf, "<i4", 1, pos)
        out.append(buf[pos + 4:pos + 4 + int(n)])
        pos += 8 + int(n)
    return out


def _cv(params) -> float:
    g, Ma = float(params["gamma"]), float(params["Ma"])
    return 1.0 / (g * (g - 1.0) * Ma * Ma)


def write_restart(directory: str, params, U_blocks: Sequence[np.ndarray], dims: Sequence[Tuple[int, int, int]],
                  kstep: int, tt: float) -> None:
    """flow3d.dat (+ SA3d.dat / SST3d.dat) and Step_mess.dat from conserved states."""
    Cv = _cv(params)
    nv = U_blocks[0].shape[0]
    f99 = open(os.path.join(directory, "flow3d.dat"), "wb")
    ft = open(os.path.join(directory, "SA3d.dat" if nv == 6 else "SST3d.dat"), "wb") if nv > 5 else None
    try:
        for U, (nx, ny, nz) in zip(U_blocks, dims):
            V = U[:, 3:nx + 4, 3:ny + 4, 3:nz + 4]        # 0..n
            d = V[0]
            u, v, w = V[1] / d, V[2] / d, V[3] / d
            T = (V[4] - (0.5 * d) * ((u * u + v * v) + w * w)) / (Cv * d)
            _record(f99, b"".join(np.asarray(a, dtype="<f8").tobytes(order="F") for a in (d, u, v, w, T)))
            if ft is not None:
                _record(ft, b"".join(np.asarray(V[m], dtype="<f8").tobytes(order="F") for m in range(5, nv)))
    finally:
        f99.close()
        if ft is not None:
            ft.close()
    with open(os.path.join(directory, "Step_mess.dat"), "w") as f:
        f.write(f" {kstep:11d} {tt!r}\n ---time Step, time ----\n")


def read_restart(directory: str, params, dims: Sequence[Tuple[int, int, int]], nv: int):
    """Conserved states over 0..n per block (download layout, ghosts beyond 0..n left zero) and
    (Kstep, tt) from Step_mess.dat when present."""
    Cv = _cv(params)
    recs = _records(os.path.join(directory, "flow3d.dat"))
    tname = "SA3d.dat" if nv == 6 else "SST3d.dat"
    trecs = _records(os.path.join(directory, tname)) if nv > 5 and os.path.exists(os.path.join(directory, tname)) else None
    out = []
    for b, (nx, ny, nz) in enumerate(dims):
        shp = (nx + 1, ny + 1, nz + 1)
        a = np.frombuffer(recs[b], "<f8").reshape((5,) + shp[::-1]).transpose(0, 3, 2, 1)
        d, u, v, w, T = a
        U = np.zeros((nv, nx + 7, ny + 7, nz + 7))
        S = (slice(3, nx + 4), slice(3, ny + 4), slice(3, nz + 4))
        U[(0,) + S] = d
        U[(1,) + S] = d * u
        U[(2,) + S] = d * v
        U[(3,) + S] = d * w
        U[(4,) + S] = (Cv * d) * T + (0.5 * d) * ((u * u + v * v) + w * w)
        if nv > 5:
            if trecs is not None:
                t = np.frombuffer(trecs[b], "<f8").reshape((nv - 5,) + shp[::-1]).transpose(0, 3, 2, 1)
                for m in range(5, nv):
                    U[(m,) + S] = t[m - 5]
            elif nv == 6:
                U[(5,) + S] = 1.0 / float(params["Re"])
            else:
                U[(5,) + S] = float(params["Kt_inf"])
                U[(6,) + S] = float(params["Wt_inf"])
        out.append(U)
    step = None
    p = os.path.join(directory, "Step_mess.dat")
    if os.path.exists(p):
        with open(p) as f:
            v = f.readline().split()
            step = (int(v[0]), float(v[1].replace("D", "E").replace("d", "e")))
    return out, step


def write_wall_dist(path: str, dw_blocks: Sequence[np.ndarray]) -> None:
    """wall_dist.dat from dw(1:nx-1, 1:ny-1, 1:nz-1) per block (numpy [i-1, j-1, k-1])."""
    with open(path, "wb") as f:
        for dw in dw_blocks:
            _record(f, np.asarray(dw, dtype="<f8").tobytes(order="F"))


def read_wall_dist(path: str, dims: Sequence[Tuple[int, int, int]]) -> List[np.ndarray]:
    recs = _records(path)
    return [np.frombuffer(r, "<f8").reshape((nz - 1, ny - 1, nx - 1)).transpose(2, 1, 0)
            for r, (nx, ny, nz) in zip(recs, dims)]
