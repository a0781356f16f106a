"""Python host of the MI355X OpenCFD-EC hot path.

Mirrors the reference host around the drop-in boundary: ``control.ec`` values go through
``set_const_para`` (sub_read_parameter.f90:371-429), Gridgen subfaces through ``Convert_bc``
(sub_convert_inp.f90:131-247), and every iteration is one call into the C ABI of
``libocfd_ec.so`` (include/ocfd_ec.h).  The library is loaded from the package's ``lib``
directory; importing this module on a machine without it raises, there is no fallback.
"""
from __future__ import annotations

import ctypes as C
import math
import os
from typing import Dict, List, Optional, Sequence

import numpy as np

from . import cases  # noqa: F401

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
LIB_PATH = os.path.join(PKG, "lib", "libocfd_ec.so")
LAP = 4

OCFD_OK, OCFD_EHIP, OCFD_EARG, OCFD_ENAN, OCFD_EUNSUPPORTED = 0, 1, 2, 3, 4
F_RES, F_DU, F_DT, F_MU, F_MUT, F_DW = 1, 2, 3, 4, 5, 6


class OcfdError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"ocfd error {code}: {msg}")
        self.code = code


class Params(C.Structure):
    _fields_ = [(n, C.c_double) for n in (
        "Ma", "Re", "gamma", "Cv", "Cp", "PrL", "PrT", "T_inf", "Twall", "AoA", "A_alfa", "A_beta",
        "P_outlet", "CFL", "dtmax", "dtmin", "dt_global", "w_LU", "MUT_MAX", "Kt_inf", "Wt_inf",
        "Ldmin", "Ldmax", "Lpmin", "Lpmax", "Lumax", "LSAmax", "Ref_S", "Ref_L")] + [
        ("Centroid", C.c_double * 3), ("Periodic_d", C.c_double * 3)] + [(n, C.c_int32) for n in (
        "NVAR", "If_viscous", "Iflag_turbulence_model", "Iflag_Scheme", "Iflag_Flux",
        "IFlag_Reconstruction", "Bound_Scheme", "Iflag_local_dt", "If_dtime_mesh",
        "IF_Scheme_Positivity", "IFLAG_LIMIT_FLOW", "Cood_Y_UP", "precision", "IF_InnerFlow", "IF_TurboMachinary")]


class BlockGeom(C.Structure):
    _fields_ = [(n, C.POINTER(C.c_double)) for n in ("x", "y", "z", "xc", "yc", "zc", "vol", "dtime_mesh", "dw",
                                                      "Si", "Sj", "Sk")] + [
        ("n", C.POINTER(C.c_double) * 9), ("met", (C.POINTER(C.c_double) * 9) * 4)]


_lib = None

EXPORTS = [
    "ocfd_create", "ocfd_destroy", "ocfd_last_error", "ocfd_add_block", "ocfd_add_block_nodes",
    "ocfd_setup_geometry", "ocfd_upload_U", "ocfd_download_U", "ocfd_download_field", "ocfd_init_flow",
    "ocfd_step_lusgs", "ocfd_residual", "ocfd_update_limit", "ocfd_bc", "ocfd_halo", "ocfd_force_vt_kw",
    "ocfd_res_norms", "ocfd_res_norms_noalloc", "ocfd_forces", "ocfd_block_status", "ocfd_get_counters", "ocfd_set_graph", "ocfd_sync",
    "ocfd_step_timed", "ocfd_set_lu_schedule", "ocfd_set_lu_tile", "ocfd_get_lu_tile", "ocfd_autotune_lu",
    "ocfd_set_lu_fail_waits", "ocfd_set_per_block",
    "ocfd_batch_create", "ocfd_batch_destroy", "ocfd_batch_add", "ocfd_batch_remove", "ocfd_batch_step",
    "ocfd_batch_set_graph", "ocfd_batch_set_lu_tile", "ocfd_batch_autotune_lu", "ocfd_batch_step_timed", "ocfd_mem_bytes",
    "ocfd_check_division",
]
LU_TILES = (4, 8, 16)
LU_PIPELINED, LU_HYPERPLANE, LU_PIPELINED_FENCED = 0, 1, 2

# stages of one iteration (OCFD_STAGE_* in include/ocfd_ec.h)
STAGES = ("pre", "inviscid", "viscous_dt", "lu_forward", "lu_backward", "update", "bc_halo")


def lib():
    """Load libocfd_ec.so (raises when the HIP library has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise OSError(f"{LIB_PATH} is missing: build it with `python opencfd-ec_amd/build.py`")
        L = C.CDLL(LIB_PATH)
        vp, i32, dp = C.c_void_p, C.c_int32, C.POINTER(C.c_double)
        ip = C.POINTER(C.c_int32)
        L.ocfd_last_error.restype = C.c_char_p
        L.ocfd_create.argtypes = [C.POINTER(Params), C.POINTER(vp)]
        L.ocfd_destroy.argtypes = [vp]
        L.ocfd_add_block.argtypes = [vp, i32, i32, i32, i32, i32, ip, C.POINTER(BlockGeom)]
        L.ocfd_add_block_nodes.argtypes = [vp, i32, i32, i32, i32, i32, ip, dp, dp, dp]
        for f in ("ocfd_setup_geometry", "ocfd_init_flow", "ocfd_residual", "ocfd_update_limit", "ocfd_bc",
                  "ocfd_halo", "ocfd_force_vt_kw", "ocfd_sync"):
            getattr(L, f).argtypes = [vp]
        L.ocfd_upload_U.argtypes = [vp, i32, dp]
        L.ocfd_download_U.argtypes = [vp, i32, dp]
        L.ocfd_download_field.argtypes = [vp, i32, i32, dp]
        L.ocfd_step_lusgs.argtypes = [vp, i32]
        L.ocfd_res_norms.argtypes = [vp, dp, dp]
        L.ocfd_res_norms_noalloc.argtypes = [vp, dp, dp]
        L.ocfd_forces.argtypes = [vp, dp]
        L.ocfd_block_status.argtypes = [vp, i32, ip, ip]
        L.ocfd_get_counters.argtypes = [vp, ip, dp]
        L.ocfd_set_graph.argtypes = [vp, i32]
        L.ocfd_step_timed.argtypes = [vp, i32, dp]
        L.ocfd_set_lu_schedule.argtypes = [vp, i32]
        L.ocfd_set_lu_tile.argtypes = [vp, i32]
        L.ocfd_set_lu_fail_waits.argtypes = [vp, C.c_int32]
        L.ocfd_set_per_block.argtypes = [vp, C.c_int32]
        L.ocfd_get_lu_tile.argtypes = [vp, ip]
        L.ocfd_autotune_lu.argtypes = [vp, i32, ip, dp]
        L.ocfd_batch_set_lu_tile.argtypes = [vp, i32]
        L.ocfd_batch_autotune_lu.argtypes = [vp, i32, ip, dp]
        L.ocfd_batch_step_timed.argtypes = [vp, i32, dp]
        L.ocfd_check_division.argtypes = [C.c_int64, C.c_uint64, i32, C.POINTER(C.c_int64)]
        L.ocfd_batch_create.argtypes = [C.POINTER(vp)]
        L.ocfd_batch_destroy.argtypes = [vp]
        L.ocfd_batch_add.argtypes = [vp, vp]
        L.ocfd_batch_remove.argtypes = [vp, vp]
        L.ocfd_batch_step.argtypes = [vp, i32]
        L.ocfd_batch_set_graph.argtypes = [vp, i32]
        L.ocfd_mem_bytes.argtypes = [vp, C.POINTER(C.c_int64)]
        _lib = L
    return _lib


def _check(code: int):
    if code != OCFD_OK:
        raise OcfdError(code, lib().ocfd_last_error().decode(errors="replace"))


def _dp(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


# ---------------------------------------------------------------------------------------------
# host restatements around the boundary

def const_params(control: Dict[str, object], precision: int = 0) -> Params:
    """set_const_para (sub_read_parameter.f90:371-429) applied to namelist values."""
    P = Params()
    g, Ma = float(control["gamma"]), float(control["Ma"])
    visc = int(control["If_viscous"])
    turb = int(control["Iflag_turbulence_model"]) if visc == 1 else 0
    P.Ma, P.Re, P.gamma = Ma, float(control["Re"]), g
    P.PrL, P.PrT, P.T_inf = float(control["PrL"]), float(control["PrT"]), float(control["T_inf"])
    P.Twall = float(control["Twall"]) / P.T_inf
    P.Lpmin = float(control["Lpmin"]) / (g * Ma * Ma)
    P.Lpmax = float(control["Lpmax"]) / (g * Ma * Ma)
    P.Ldmin, P.Ldmax, P.Lumax, P.LSAmax = (float(control[k]) for k in ("Ldmin", "Ldmax", "Lumax", "LSAmax"))
    pi = 3.1415926535897932
    P.AoA = float(control["AoA"]) * pi / 180.0
    aos = float(control["AoS"]) * pi / 180.0
    if int(control["Cood_Y_UP"]) == 1:
        P.A_alfa, P.A_beta = P.AoA, aos
    else:
        P.A_alfa, P.A_beta = aos, P.AoA
    P.Cv = 1.0 / (g * (g - 1.0) * Ma * Ma)
    P.Cp = P.Cv * g
    for k in ("P_outlet", "CFL", "dtmax", "dtmin", "dt_global", "w_LU", "MUT_MAX", "Kt_inf", "Wt_inf", "Ref_S", "Ref_L"):
        setattr(P, k, float(control[k]))
    P.Centroid[:] = [float(v) for v in control["Centroid"]]
    P.Periodic_d[:] = [float(control["Periodic_dX"]), float(control["Periodic_dY"]), float(control["Periodic_dZ"])]
    P.NVAR = 6 if turb in (2, 21) else (7 if turb == 3 else 5)
    P.If_viscous, P.Iflag_turbulence_model = visc, turb
    P.Iflag_Scheme, P.Iflag_Flux = int(control["Iflag_Scheme"]), int(control["Iflag_Flux"])
    P.IFlag_Reconstruction = int(control["IFlag_Reconstruction"])
    bs = int(control["Bound_Scheme"])
    P.Bound_Scheme = P.Iflag_Scheme if bs == -1 else bs
    P.Iflag_local_dt, P.If_dtime_mesh = int(control["Iflag_local_dt"]), int(control["If_dtime_mesh"])
    P.IF_Scheme_Positivity, P.IFLAG_LIMIT_FLOW = int(control["IF_Scheme_Positivity"]), int(control["IFLAG_LIMIT_FLOW"])
    P.Cood_Y_UP = int(control["Cood_Y_UP"])
    P.precision = precision
    P.IF_InnerFlow = int(control.get("IF_InnerFlow", 0))
    P.IF_TurboMachinary = int(control.get("IF_TurboMachinary", 0))
    return P


def convert_bc(rng: Sequence[int], bc: int, link: Optional[Sequence[int]]) -> List[int]:
    """Convert_bc (sub_convert_inp.f90:131-247): one Gridgen record -> 21-int .inc record (f_no/f_no1 unset)."""
    kb = [rng[0], rng[2], rng[4]]; ke = [rng[1], rng[3], rng[5]]
    s = [0 if kb[k] == ke[k] else (1 if kb[k] > 0 else -1) for k in range(3)]
    lo = [min(abs(kb[k]), abs(ke[k])) for k in range(3)]
    hi = [max(abs(kb[k]), abs(ke[k])) for k in range(3)]
    if s[0] == 0: face = 1 if lo[0] == 1 else 4
    elif s[1] == 0: face = 2 if lo[1] == 1 else 5
    else: face = 3 if lo[2] == 1 else 6
    rec = [lo[0], hi[0], lo[1], hi[1], lo[2], hi[2], bc, face, 0] + [0] * 12
    if bc < 0:
        kb1 = [link[0], link[2], link[4]]; ke1 = [link[1], link[3], link[5]]
        p = [0 if kb1[k] == ke1[k] else (1 if kb1[k] > 0 else -1) for k in range(3)]
        lo1 = [min(abs(kb1[k]), abs(ke1[k])) for k in range(3)]
        hi1 = [max(abs(kb1[k]), abs(ke1[k])) for k in range(3)]
        if p[0] == 0: face1 = 1 if lo1[0] == 1 else 4
        elif p[1] == 0: face1 = 2 if lo1[1] == 1 else 5
        else: face1 = 3 if lo1[2] == 1 else 6
        Lp = [0, 0, 0]
        for k in range(3):
            for j in range(3):
                if s[k] == p[j]:
                    Lp[k] = j + 1
        for k in range(3):
            if s[k] != 0:
                k1 = Lp[k] - 1
                if (ke[k] - kb[k]) * (ke1[k1] - kb1[k1]) < 0:
                    Lp[k] = -Lp[k]
            elif (face - 1) // 3 == (face1 - 1) // 3:
                Lp[k] = -Lp[k]
        rec[9:18] = [lo1[0], hi1[0], lo1[1], hi1[1], lo1[2], hi1[2], int(link[6]), face1, 0]
        rec[18:21] = Lp
    return rec


def inc_tables(case) -> List[np.ndarray]:
    """bc3d.inp -> bc3d.inc for every block (convert_inp_inc, sub_convert_inp.f90:25-122)."""
    tabs = []
    for b in case.blocks:
        recs = []
        for q, s in enumerate(b.subfaces, start=1):
            r = convert_bc(s.rng, s.bc, s.link)
            r[8] = q
            recs.append(r)
        tabs.append(recs)
    for m, recs in enumerate(tabs, start=1):
        for r in recs:
            if r[6] >= 0:
                continue
            for q1, r1 in enumerate(tabs[r[15] - 1], start=1):
                if r[9:15] == r1[0:6] and r1[15] == m:
                    r[17] = q1
                    break
            else:
                raise ValueError(f"linked subface of block {m} not found")
    return [np.array(t, dtype=np.int32) for t in tabs]


# ---------------------------------------------------------------------------------------------

class Solver:
    """One C-ABI context: blocks resident in HBM, one HIP stream."""

    def __init__(self, control: Dict[str, object], precision: int = 0):
        L = lib()
        self.control = dict(control)
        self.params = const_params(control, precision)
        self.nv = int(self.params.NVAR)
        h = C.c_void_p()
        _check(L.ocfd_create(C.byref(self.params), C.byref(h)))
        self._h = h
        self.dims: Dict[int, tuple] = {}

    def close(self):
        if getattr(self, "_h", None):
            # ocfd_destroy takes a member out of its batch first
            bt = getattr(self, "_batch", None)
            if bt is not None and self in bt.members:
                bt.members.remove(self)
            self._batch = None
            lib().ocfd_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- setup -----------------------------------------------------------------------------
    def add_block_nodes(self, no: int, x, y, z, inc21: np.ndarray):
        nx, ny, nz = x.shape
        xs = [np.asfortranarray(a, dtype=np.float64) for a in (x, y, z)]
        t = np.ascontiguousarray(inc21, dtype=np.int32)
        _check(lib().ocfd_add_block_nodes(self._h, no, nx, ny, nz, len(t), t.ctypes.data_as(C.POINTER(C.c_int32)), *[_dp(a) for a in xs]))
        self.dims[no] = (nx, ny, nz)
        self._keep = getattr(self, "_keep", []) + xs

    def setup_geometry(self):
        _check(lib().ocfd_setup_geometry(self._h))

    def add_block_geom(self, no: int, dims, inc21: np.ndarray, g: Dict[str, np.ndarray]):
        """g maps reference array names to Fortran-ordered arrays with the reference's bounds."""
        nx, ny, nz = dims
        G = BlockGeom()
        keep = []

        def ptr(name):
            a = g.get(name)
            if a is None:
                return C.POINTER(C.c_double)()
            a = np.asfortranarray(a, dtype=np.float64)
            keep.append(a)
            return _dp(a)
        for f in ("x", "y", "z", "xc", "yc", "zc", "vol", "dtime_mesh", "dw", "Si", "Sj", "Sk"):
            setattr(G, f, ptr(f))
        for q, nm in enumerate(("ni1", "ni2", "ni3", "nj1", "nj2", "nj3", "nk1", "nk2", "nk3")):
            G.n[q] = ptr(nm)
        for s in range(4):
            for q, nm in enumerate(("ix", "iy", "iz", "jx", "jy", "jz", "kx", "ky", "kz")):
                G.met[s][q] = ptr(f"{nm}{s}")
        t = np.ascontiguousarray(inc21, dtype=np.int32)
        _check(lib().ocfd_add_block(self._h, no, nx, ny, nz, len(t), t.ctypes.data_as(C.POINTER(C.c_int32)), C.byref(G)))
        self.dims[no] = (nx, ny, nz)

    # ---- state -----------------------------------------------------------------------------
    def ushape(self, no: int):
        nx, ny, nz = self.dims[no]
        return (self.nv, nx + 7, ny + 7, nz + 7)

    def upload_U(self, no: int, U: np.ndarray):
        a = np.asfortranarray(U, dtype=np.float64)
        assert a.shape == self.ushape(no), (a.shape, self.ushape(no))
        _check(lib().ocfd_upload_U(self._h, no, _dp(a)))

    def download_U(self, no: int) -> np.ndarray:
        a = np.zeros(self.ushape(no), order="F")
        _check(lib().ocfd_download_U(self._h, no, _dp(a)))
        return a

    def wall_distance(self) -> List[np.ndarray]:
        """dw(1:nx-1, 1:ny-1, 1:nz-1) of every block (numpy [i-1, j-1, k-1]), as computed on the device."""
        out = []
        for no in sorted(self.dims):
            nx, ny, nz = self.dims[no]
            out.append(np.asfortranarray(self.field(no, F_DW)[2:nx + 1, 2:ny + 1, 2:nz + 1]))
        return out

    def field(self, no: int, fid: int) -> np.ndarray:
        nx, ny, nz = self.dims[no]
        nv = self.nv if fid in (F_RES, F_DU) else 1
        a = np.zeros((nv, nx + 3, ny + 3, nz + 3), order="F")
        _check(lib().ocfd_download_field(self._h, no, fid, _dp(a)))
        return a if nv > 1 else a[0]

    def set_interior(self, states: Sequence[np.ndarray]):
        """Overwrite interior cells of every block (ghosts keep their values)."""
        for no, Ui in enumerate(states, start=1):
            U = self.download_U(no)
            nx, ny, nz = self.dims[no]
            U[:, LAP:LAP + nx - 1, LAP:LAP + ny - 1, LAP:LAP + nz - 1] = Ui
            self.upload_U(no, U)

    # ---- stepping --------------------------------------------------------------------------
    def init_flow(self): _check(lib().ocfd_init_flow(self._h))
    def bc(self): _check(lib().ocfd_bc(self._h))
    def halo(self): _check(lib().ocfd_halo(self._h))
    def residual(self): _check(lib().ocfd_residual(self._h))
    def update_limit(self): _check(lib().ocfd_update_limit(self._h))
    def force_vt_kw(self): _check(lib().ocfd_force_vt_kw(self._h))
    def step(self, n: int = 1): _check(lib().ocfd_step_lusgs(self._h, n))
    def sync(self): _check(lib().ocfd_sync(self._h))

    def step_timed(self, n: int) -> Dict[str, float]:
        """n iterations with HIP events at the stage boundaries: device ms per stage, summed."""
        ms = np.zeros(len(STAGES))
        _check(lib().ocfd_step_timed(self._h, n, _dp(ms)))
        return dict(zip(STAGES, ms.tolist()))
    def set_graph(self, on: bool): _check(lib().ocfd_set_graph(self._h, 1 if on else 0))
    def set_lu_schedule(self, schedule: int): _check(lib().ocfd_set_lu_schedule(self._h, schedule))
    def set_lu_tile(self, tj: int): _check(lib().ocfd_set_lu_tile(self._h, tj))
    def set_lu_fail_waits(self, on: bool): _check(lib().ocfd_set_lu_fail_waits(self._h, int(bool(on))))

    def set_per_block(self, on: bool):
        """Plain launch mode: every kernel once per block, eagerly, 8x8 LU-SGS tiles."""
        _check(lib().ocfd_set_per_block(self._h, int(bool(on))))

    def lu_tile(self) -> int:
        v = C.c_int32()
        _check(lib().ocfd_get_lu_tile(self._h, C.byref(v)))
        return v.value

    def autotune_lu(self, n: int = 2):
        """n iterations per LU tile shape; keeps the fastest.  Returns (tj, {tj: ms per sweep pair})."""
        best, ms = C.c_int32(), np.zeros(3)
        _check(lib().ocfd_autotune_lu(self._h, n, C.byref(best), _dp(ms)))
        return best.value, dict(zip(LU_TILES, ms.tolist()))

    def norms(self):
        mx = np.zeros(self.nv); rms = np.zeros(self.nv)
        _check(lib().ocfd_res_norms(self._h, _dp(mx), _dp(rms)))
        return mx, rms

    def norms_noalloc(self):
        """norms() through the entry point that allocates nothing (the library's fixed device
        scratch, at most 64 blocks per context)."""
        mx = np.zeros(self.nv); rms = np.zeros(self.nv)
        _check(lib().ocfd_res_norms_noalloc(self._h, _dp(mx), _dp(rms)))
        return mx, rms

    def forces(self):
        out = np.zeros(12)
        _check(lib().ocfd_forces(self._h, _dp(out)))
        return out

    def counters(self):
        """(Kstep, tt) of the context."""
        k = C.c_int32(); t = C.c_double()
        _check(lib().ocfd_get_counters(self._h, C.byref(k), C.byref(t)))
        return int(k.value), float(t.value)

    def mem_bytes(self) -> int:
        v = C.c_int64()
        _check(lib().ocfd_mem_bytes(self._h, C.byref(v)))
        return int(v.value)

    def status(self, no: int):
        a = C.c_int32(); b = C.c_int32()
        _check(lib().ocfd_block_status(self._h, no, C.byref(a), C.byref(b)))
        return a.value, b.value


class Batch:
    """Several solvers advancing together, every kernel covering all their blocks in one launch
    (ocfd_batch_*).  Members keep their own parameters, state, counters, norms and forces."""

    def __init__(self):
        h = C.c_void_p()
        _check(lib().ocfd_batch_create(C.byref(h)))
        self._h = h
        self.members: List["Solver"] = []

    def close(self):
        if getattr(self, "_h", None):
            lib().ocfd_batch_destroy(self._h)
            self._h = None
            for s in self.members:
                s._batch = None
            self.members = []

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add(self, s: "Solver"):
        _check(lib().ocfd_batch_add(self._h, s._h))
        self.members.append(s)
        s._batch = self

    def remove(self, s: "Solver"):
        _check(lib().ocfd_batch_remove(self._h, s._h))
        self.members.remove(s)
        s._batch = None

    def step(self, n: int = 1):
        _check(lib().ocfd_batch_step(self._h, n))

    def set_graph(self, on: bool):
        _check(lib().ocfd_batch_set_graph(self._h, 1 if on else 0))

    def set_lu_tile(self, tj: int):
        _check(lib().ocfd_batch_set_lu_tile(self._h, tj))

    def step_timed(self, n: int) -> Dict[str, float]:
        """n iterations; device ms per stage summed over them (HIP events)."""
        ms = np.zeros(len(STAGES))
        _check(lib().ocfd_batch_step_timed(self._h, n, _dp(ms)))
        return dict(zip(STAGES, ms.tolist()))

    def autotune_lu(self, n: int = 2):
        """n iterations per LU tile shape for every member; keeps the fastest."""
        best, ms = C.c_int32(), np.zeros(3)
        _check(lib().ocfd_batch_autotune_lu(self._h, n, C.byref(best), _dp(ms)))
        return best.value, dict(zip(LU_TILES, ms.tolist()))


def solver_from_case(case, precision: int = 0, states: Optional[Sequence[np.ndarray]] = None) -> Solver:
    """Build a solver from node coordinates only (device geometry), freestream, optional interior state."""
    s = Solver(case.params, precision)
    for no, (b, t) in enumerate(zip(case.blocks, inc_tables(case)), start=1):
        s.add_block_nodes(no, b.x, b.y, b.z, t)
    s.setup_geometry()
    s.init_flow()
    if states is not None:
        s.set_interior(states)
        s.bc()
        s.halo()
    return s


def check_division(n: int, seed: int = 1, numerator_exponent: int = 300) -> int:
    """Quotients of the shared-reciprocal division that differ from IEEE division (ocfd_check_division)."""
    m = C.c_int64()
    _check(lib().ocfd_check_division(n, seed, numerator_exponent, C.byref(m)))
    return int(m.value)
