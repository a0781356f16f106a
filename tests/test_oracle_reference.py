"""The oracle against the compiled reference (oracle/_ref/ref_driver) on every case of the catalog.

The reference's own Fortran runs the same case (same nodes, subfaces, control.ec values and
seeded interior state) through its setup and NS_Time_advance; the oracle must reproduce its
arrays after the first iteration and the state after the last one bit for bit (1e-13 relative
allowed for the libm calls, which both take from the same host library).  Skipped where the
reference build is absent.
"""
import numpy as np
import pytest

from helpers import CASES, rel_err, n_exact
from ocfd_amd import cases as C
import oracle
import refrun

pytestmark = pytest.mark.skipif(not refrun.available(), reason="oracle/_ref/ref_driver not built")

NSTEPS = 3


def _pad(o, blk, name, lo, hi):
    return oracle.pad_view(o.get(blk, name), lo, hi)


def _cmp(label, got, ref, tol=1e-13):
    e = rel_err(got, ref)
    ne, n = n_exact(got, ref)
    print(f"  {label:8s} rel={e:.2e} exact={ne}/{n}")
    assert e <= tol, f"{label}: rel {e:.3e}"


@pytest.mark.parametrize("name", sorted(CASES))
def test_oracle_matches_compiled_reference(name):
    print(f"\n[{name}]")
    cs = CASES[name]
    case = cs.builder()
    U0 = C.perturbed_state(case, seed=cs.seed, amp=cs.amp)
    nsteps = 1 + cs.more
    r = refrun.run_case(case, nsteps=nsteps, kshow=1, U_init=U0)
    nv = C.nvar_of(case.params)
    o = oracle.Oracle(case)
    o.init_flow(); o.set_interior(U0); o.bc_halo()
    # geometry as the reference's setup left it
    for n, b in enumerate(case.blocks, start=1):
        nx, ny, nz = b.dims
        g = r["geom"][n - 1]
        for nm in ("vol", "dtime_mesh") + (("dw",) if nv > 5 else ()):
            _cmp(nm, _pad(o, n, nm, (1, 1, 1), (nx - 1, ny - 1, nz - 1)), g[nm].data)
        for nm in ("Si", "Sj", "Sk", "ni1", "nj2", "nk3", "ix1", "jy2", "kz3"):
            _cmp(nm, _pad(o, n, nm, (1, 1, 1), (nx, ny, nz)), g[nm].data)
        for m in range(nv):
            _cmp(f"U0[{m + 1}]", _pad(o, n, f"U{m + 1}", (0, 0, 0), (nx, ny, nz)),
                 r["state0"][n - 1]["U"].sl((m + 1, m + 1), (0, nx), (0, ny), (0, nz))[0])
    o.step(1)
    for n, b in enumerate(case.blocks, start=1):
        nx, ny, nz = b.dims
        s1 = r["state1"][n - 1]
        lo, hi = (1, 1, 1), (nx - 1, ny - 1, nz - 1)
        for m in range(nv):
            _cmp(f"Res[{m + 1}]", _pad(o, n, f"Res{m + 1}", lo, hi), s1["Res"].sl((m + 1, m + 1), (1, nx - 1), (1, ny - 1), (1, nz - 1))[0])
            _cmp(f"DU[{m + 1}]", _pad(o, n, f"DU{m + 1}", lo, hi), s1["DU"].sl((m + 1, m + 1), (1, nx - 1), (1, ny - 1), (1, nz - 1))[0])
            _cmp(f"U1[{m + 1}]", _pad(o, n, f"U{m + 1}", (0, 0, 0), (nx, ny, nz)), s1["U"].sl((m + 1, m + 1), (0, nx), (0, ny), (0, nz))[0])
        _cmp("dt", _pad(o, n, "dt", lo, hi), s1["dt"].sl((1, nx - 1), (1, ny - 1), (1, nz - 1)))
        if "mu_t" in s1:
            _cmp("mu_t", _pad(o, n, "mu_t", lo, hi), s1["mu_t"].sl((1, nx - 1), (1, ny - 1), (1, nz - 1)))
    if nsteps > 1:
        o.step(nsteps - 1)
        for n, b in enumerate(case.blocks, start=1):
            nx, ny, nz = b.dims
            sN = r["stateN"][n - 1]
            for m in range(nv):
                _cmp(f"UN[{m + 1}]", _pad(o, n, f"U{m + 1}", (0, 0, 0), (nx, ny, nz)),
                     sN["U"].sl((m + 1, m + 1), (0, nx), (0, ny), (0, nz))[0], tol=1e-12)
            assert o.dims(n)[7] == sN["overlimit"]
    st = o.monitor()
    ks, mx, rms = r["norms"]
    np.testing.assert_allclose(st["res_max"], mx[-1], rtol=1e-12)
    np.testing.assert_allclose(st["res_rms"], rms[-1], rtol=1e-12)
