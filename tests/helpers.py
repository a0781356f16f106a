"""Shared test helpers: case catalog, oracle geometry handover and the parity norm."""
from __future__ import annotations

from collections import namedtuple

import numpy as np

from ocfd_amd import cases as C
import oracle

# name -> Case(builder, seed, amplitude, transcendental-free, extra iterations after the first)
# Transcendental-free cases (no exp/pow/tanh/acos on the path) are expected bit-exact through the
# drop-in path where the host hands over the geometry.  The HLLC C-grid case runs one iteration
# only: the reference HLLC (sub_flux_split.f90:67 uses vvr in the right star state's second
# tangential momentum) is not consistent where the face-normal velocity is negative, so a C-grid
# with HLLC diverges within a few iterations in the reference and in the oracle alike.
Case = namedtuple("Case", "builder seed amp exact more")

_CG = dict(wall_spacing=0.02)
CASES = {
    "channel_roe_muscl3": Case(lambda: C.channel_bump(25, 13, 7), 1, 1e-3, True, 2),
    "channel_hllc_weno5": Case(lambda: C.channel_bump(25, 13, 7, Iflag_Flux=3, Iflag_Scheme=7), 2, 1e-3, True, 2),
    "channel_sw_ud3": Case(lambda: C.channel_bump(25, 13, 7, Iflag_Flux=1, Iflag_Scheme=2), 3, 1e-3, True, 2),
    "channel_hll_nnd2": Case(lambda: C.channel_bump(25, 13, 7, Iflag_Flux=2, Iflag_Scheme=1), 4, 1e-3, True, 2),
    "channel_vl_muscl2c": Case(lambda: C.channel_bump(25, 13, 7, Iflag_Flux=5, Iflag_Scheme=4), 5, 1e-3, True, 2),
    "channel_ausm_omuscl2": Case(lambda: C.channel_bump(25, 13, 7, Iflag_Flux=6, Iflag_Scheme=6), 6, 1e-3, False, 2),
    "channel_roe_ud5_cons": Case(lambda: C.channel_bump(25, 13, 7, Iflag_Scheme=8, IFlag_Reconstruction=1), 7, 1e-3, True, 2),
    "channel_roe_weno5_char": Case(lambda: C.channel_bump(25, 13, 7, Iflag_Scheme=7, IFlag_Reconstruction=2), 8, 1e-3, True, 2),
    "channel_ud1_muscl2u": Case(lambda: C.channel_bump(25, 13, 7, Iflag_Scheme=3, Iflag_Flux=4), 9, 1e-3, True, 2),
    "channel_cut6": Case(lambda: C.cut_blocks(C.channel_bump(25, 13, 7), [[[9, 17], [7], []]]), 10, 1e-3, True, 2),
    "cgrid_sa_weno5_hllc": Case(lambda: C.naca_cgrid(65, 25, 9, **_CG), 2, 1e-5, False, 0),
    "cgrid_sa_weno5_roe": Case(lambda: C.naca_cgrid(65, 25, 9, Iflag_Flux=4, **_CG), 2, 1e-5, False, 2),
    "cgrid_sa_limiter": Case(lambda: C.naca_cgrid(65, 25, 9, Iflag_Flux=4, Ldmax=1.001, LSAmax=5.004, **_CG),
                             3, 1e-3, False, 2),
    "cgrid_sst_muscl3_ausm": Case(lambda: C.naca_cgrid(65, 25, 9, Iflag_turbulence_model=3, Iflag_Scheme=5,
                                                       Iflag_Flux=6, Re=3e6, Ma=0.7, **_CG), 4, 1e-5, False, 2),
    "cgrid_laminar_vl": Case(lambda: C.naca_cgrid(65, 25, 9, Iflag_turbulence_model=0, Re=1e4, Iflag_Scheme=5,
                                                  Iflag_Flux=5, **_CG), 5, 1e-5, True, 2),
    "cgrid_sa_iso_roe": Case(lambda: C.naca_cgrid(65, 25, 9, Twall=400.0, Iflag_Flux=4, **_CG), 6, 1e-5, False, 2),
    "cgrid_sa_cut12": Case(lambda: C.cut_blocks(C.naca_cgrid(65, 25, 9, Iflag_Flux=4, **_CG), [[[17, 40], [12], [5]]]),
                           7, 1e-5, False, 2),
    # inflow (bc 5) and outflow (bc 6) faces of the external-flow mode, and an extrapolated outlet
    # (bc 401): sub_boundary.f90:28-40,152-283,458-504
    "channel_inout_bc56": Case(lambda: C.channel_bump(25, 13, 7, inlet_bc=C.BC_INFLOW, outlet_bc=C.BC_OUTFLOW), 13, 1e-3, True, 2),
    "channel_extrap_bc401": Case(lambda: C.channel_bump(25, 13, 7, outlet_bc=C.BC_EXTRAPOLATE, Iflag_Flux=5), 14, 1e-3, True, 2),
    # MUT_MAX >= 0: mu_t clamped to MUT_MAX mu (limit_mut, sub_Residual.f90:1387-1408).  The values sit at
    # the median of mu_t / mu of these states (1.29 for SA, 0.01 for SST), so about half the cells clamp
    "cgrid_sa_mutmax": Case(lambda: C.naca_cgrid(65, 25, 9, Iflag_Flux=4, MUT_MAX=1.294, **_CG), 11, 1e-3, False, 2),
    "cgrid_sst_mutmax": Case(lambda: C.naca_cgrid(65, 25, 9, Iflag_turbulence_model=3, Iflag_Scheme=5, Iflag_Flux=6,
                                                  Re=3e6, Ma=0.7, MUT_MAX=0.01, **_CG), 12, 1e-5, False, 2),
}

MET = ("ix", "iy", "iz", "jx", "jy", "jz", "kx", "ky", "kz")
NRM = ("ni1", "ni2", "ni3", "nj1", "nj2", "nj3", "nk1", "nk2", "nk3")


def oracle_geometry(o: oracle.Oracle, blk: int, dims, turb: bool):
    """Block_TYPE geometry arrays of the oracle, with the reference's bounds, for ocfd_add_block."""
    nx, ny, nz = dims
    g = {}
    for nm in ("x", "y", "z"):
        g[nm] = oracle.pad_view(o.get(blk, nm), (0, 0, 0), (nx + 1, ny + 1, nz + 1))
    for nm in ("xc", "yc", "zc"):
        g[nm] = oracle.pad_view(o.get(blk, nm), (-3, -3, -3), (nx + 3, ny + 3, nz + 3))
    for nm in ("vol", "dtime_mesh") + (("dw",) if turb else ()):
        g[nm] = oracle.pad_view(o.get(blk, nm), (1, 1, 1), (nx - 1, ny - 1, nz - 1))
    face = ["Si", "Sj", "Sk"] + list(NRM) + [f"{m}{s}" for s in (1, 2, 3) for m in MET]
    if turb:
        face += [f"{m}0" for m in MET]
    for nm in face:
        g[nm] = oracle.pad_view(o.get(blk, nm), (1, 1, 1), (nx, ny, nz))
    return g


def oracle_U(o: oracle.Oracle, blk: int, nv: int):
    """U of one block as (nv, nx+7, ny+7, nz+7) indexed [m, i+3, j+3, k+3]."""
    return np.stack([o.get(blk, f"U{m + 1}").transpose(2, 1, 0) for m in range(nv)])


def oracle_field(o: oracle.Oracle, blk: int, name: str, nv: int = 0):
    """Res/DU (nv>0) or a scalar field as [m, i+1, j+1, k+1] over -1..n+1 like ocfd_download_field."""
    if nv:
        return np.stack([o.get(blk, f"{name}{m + 1}").transpose(2, 1, 0)[2:-2, 2:-2, 2:-2] for m in range(nv)])
    return o.get(blk, name).transpose(2, 1, 0)[2:-2, 2:-2, 2:-2]


def rel_err(a, b, floor_frac: float = 1e-8):
    """max |a-b| / max(|b|, tau) with tau = floor_frac * max|b| (per-cell parity norm)."""
    a = np.asarray(a, dtype=np.float64); b = np.asarray(b, dtype=np.float64)
    tau = floor_frac * np.max(np.abs(b)) if b.size else 0.0
    den = np.maximum(np.abs(b), tau if tau > 0 else 1e-300)
    return float(np.max(np.abs(a - b) / den)) if a.size else 0.0


def n_exact(a, b):
    return int(np.count_nonzero(np.asarray(a) == np.asarray(b))), int(np.asarray(a).size)
