"""ctypes wrapper of the CPU oracle (oracle/_build/liboracle.so).  TEST INFRASTRUCTURE ONLY.

Arrays come back in the padded box used by the oracle: index [k+3, j+3, i+3] for reference
indices (i, j, k) with i in [-3, nx+3].
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from typing import Dict, List, Optional, Sequence

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_build", "liboracle.so")

_RKEYS = ["Ma", "Re", "gamma", "AoA", "AoS", "P_outlet", "PrL", "PrT", "T_inf", "Twall", "CFL",
          "dtmax", "dtmin", "dt_global", "w_LU", "MUT_MAX", "Kt_inf", "Wt_inf", "Ldmin", "Ldmax",
          "Lpmin", "Lpmax", "Lumax", "LSAmax", "Ref_S", "Ref_L"]
_IKEYS = ["If_viscous", "Iflag_turbulence_model", "Iflag_Scheme", "Iflag_Flux",
          "IFlag_Reconstruction", "Bound_Scheme", "Iflag_local_dt", "If_dtime_mesh",
          "IF_Scheme_Positivity", "IFLAG_LIMIT_FLOW", "Cood_Y_UP", "Time_Method"]

_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        vp, ip, dp, cp = C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_double), C.c_char_p
        L.orc_create.restype = vp; L.orc_create.argtypes = [dp, ip]
        L.orc_destroy.argtypes = [vp]
        L.orc_add_block.argtypes = [vp, C.c_int, C.c_int, C.c_int, dp, dp, dp]
        L.orc_add_subface.argtypes = [vp, C.c_int, ip, C.c_int, ip]
        L.orc_get_inc.argtypes = [vp, C.c_int, ip]
        for f in ("orc_setup", "orc_init_flow", "orc_bc_halo", "orc_step", "orc_monitor", "orc_residual",
                  "orc_update", "orc_bc", "orc_halo_only", "orc_vt_kw"):
            getattr(L, f).argtypes = [vp]
        L.orc_set_interior_U.argtypes = [vp, C.c_int, dp]
        L.orc_block_dims.argtypes = [vp, C.c_int, ip]
        L.orc_get_array.argtypes = [vp, C.c_int, cp, dp]
        L.orc_set_array.argtypes = [vp, C.c_int, cp, dp]
        L.orc_get_state.argtypes = [vp, ip, dp, dp, dp, dp]
        L.orc_face_flux.argtypes = [C.c_int, dp, dp, C.c_double, dp]
        L.orc_face_recon.argtypes = [C.c_int, dp, dp]
        _lib = L
    return _lib


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _iptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_int))


class Oracle:
    """One oracle context holding every block of a case."""

    def __init__(self, case, setup: bool = True):
        L = lib()
        p = case.params
        rp = np.array([float(p[k]) for k in _RKEYS] + [float(v) for v in p["Centroid"]]
                      + [float(p["Periodic_dX"]), float(p["Periodic_dY"]), float(p["Periodic_dZ"])])
        ip_ = np.array([int(p[k]) for k in _IKEYS], dtype=np.int32)
        self._c = L.orc_create(_dptr(rp), _iptr(ip_))
        self.case = case
        self.nblk = len(case.blocks)
        for b in case.blocks:
            nx, ny, nz = b.dims
            xs = [np.asfortranarray(a, dtype=np.float64) for a in (b.x, b.y, b.z)]
            L.orc_add_block(self._c, nx, ny, nz, *[_dptr(a) for a in xs])
        for m, b in enumerate(case.blocks, start=1):
            for s in b.subfaces:
                rng = np.array(s.rng, dtype=np.int32)
                lk = np.array(s.link if s.link else (0,) * 7, dtype=np.int32)
                L.orc_add_subface(self._c, m, _iptr(rng), int(s.bc), _iptr(lk))
        if setup:
            if L.orc_setup(self._c) != 0:
                raise RuntimeError("oracle setup failed")

    def __del__(self):
        try:
            if self._c:
                lib().orc_destroy(self._c)
                self._c = None
        except Exception:
            pass

    # ----- state -----
    def dims(self, blk: int):
        out = np.zeros(9, dtype=np.int32)
        lib().orc_block_dims(self._c, blk, _iptr(out))
        return out

    def init_flow(self):
        lib().orc_init_flow(self._c)

    def set_interior(self, states: Sequence[np.ndarray]):
        for m, U in enumerate(states, start=1):
            a = np.asfortranarray(U, dtype=np.float64)
            lib().orc_set_interior_U(self._c, m, _dptr(a))

    def bc_halo(self):
        lib().orc_bc_halo(self._c)

    def residual(self):
        lib().orc_residual(self._c)

    def update_limit(self):
        lib().orc_update(self._c)

    def bc(self):
        lib().orc_bc(self._c)

    def halo(self):
        lib().orc_halo_only(self._c)

    def force_vt_kw(self):
        lib().orc_vt_kw(self._c)

    def step(self, n: int = 1):
        for _ in range(n):
            lib().orc_step(self._c)

    def monitor(self):
        if lib().orc_monitor(self._c) != 0:
            raise FloatingPointError("NaN in residual")
        return self.state()

    def state(self) -> Dict[str, object]:
        k = np.zeros(1, dtype=np.int32); tt = np.zeros(1); mx = np.zeros(7); rms = np.zeros(7); f = np.zeros(12)
        lib().orc_get_state(self._c, _iptr(k), _dptr(tt), _dptr(mx), _dptr(rms), _dptr(f))
        nv = int(self.dims(1)[6])
        return dict(kstep=int(k[0]), tt=float(tt[0]), res_max=mx[:nv].copy(), res_rms=rms[:nv].copy(),
                    Ft=f[:9].copy(), CL=float(f[9]), CD=float(f[10]), CS=float(f[11]))

    def get(self, blk: int, name: str) -> np.ndarray:
        d = self.dims(blk)
        NI, NJ, NK = int(d[3]), int(d[4]), int(d[5])
        if name.startswith("Surf"):
            n2 = max(NI, NJ) * max(NJ, NK) * 2
            out = np.zeros(n2)
        else:
            out = np.zeros(NI * NJ * NK)
        if lib().orc_get_array(self._c, blk, name.encode(), _dptr(out)) != 0:
            raise KeyError(name)
        if name.startswith("Surf"):
            return out
        return out.reshape(NK, NJ, NI)

    def set(self, blk: int, name: str, arr: np.ndarray):
        a = np.ascontiguousarray(arr, dtype=np.float64)
        if lib().orc_set_array(self._c, blk, name.encode(), _dptr(a)) != 0:
            raise KeyError(name)

    def inc(self, blk: int) -> np.ndarray:
        out = np.zeros(21 * 64, dtype=np.int32)
        n = lib().orc_get_inc(self._c, blk, _iptr(out))
        return out[:21 * n].reshape(n, 21)


def face_flux(iflux: int, QL, QR, gamma: float = 1.4) -> np.ndarray:
    """1-D numerical flux of the rotated primitive states (rho, u_n, u_t1, u_t2, p)."""
    ql = np.ascontiguousarray(QL, dtype=np.float64); qr = np.ascontiguousarray(QR, dtype=np.float64)
    F = np.zeros(5)
    lib().orc_face_flux(int(iflux), _dptr(ql), _dptr(qr), float(gamma), _dptr(F))
    return F


def face_recon(scheme: int, u8) -> np.ndarray:
    """(left, right) values at face i-1/2 from u(-3..4) with the given Iflag_Scheme."""
    u = np.ascontiguousarray(u8, dtype=np.float64)
    LR = np.zeros(2)
    lib().orc_face_recon(int(scheme), _dptr(u), _dptr(LR))
    return LR


def pad_view(a: np.ndarray, lo, hi):
    """Slice of a padded-box array over inclusive reference ranges lo=(i0,j0,k0), hi=(i1,j1,k1),
    returned in Fortran order [i, j, k]."""
    (i0, j0, k0), (i1, j1, k1) = lo, hi
    return a[k0 + 3:k1 + 4, j0 + 3:j1 + 4, i0 + 3:i1 + 4].transpose(2, 1, 0)
