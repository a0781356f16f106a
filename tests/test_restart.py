"""Restart and wall-distance files against the reference's own writers (output_flow, write_dw).

The compiled reference runs a case and writes flow3d.dat (+ SA3d.dat / SST3d.dat), Step_mess.dat
and wall_dist.dat; ocfd_amd.restart must write the same bytes from the same arrays and read them
back to the conserved state.  Skipped where the reference build is absent.
"""
import os
import tempfile

import numpy as np
import pytest

from helpers import CASES
from ocfd_amd import cases as C
from ocfd_amd import restart as R
import refrun

pytestmark = pytest.mark.skipif(not refrun.available(), reason="oracle/_ref/ref_driver not built")


def _download_layout(arr, nv):
    """Reference U(NVAR, -3:n+3, ...) dump -> (NVAR, nx+7, ny+7, nz+7) [m, i+3, j+3, k+3]."""
    return np.ascontiguousarray(arr.data[:nv])


@pytest.mark.parametrize("name", ["cgrid_sa_weno5_roe", "cgrid_sst_muscl3_ausm", "channel_cut6"])
def test_restart_files_match_reference_writers(name):
    cs = CASES[name]
    case = cs.builder()
    U0 = C.perturbed_state(case, seed=cs.seed, amp=cs.amp)
    r = refrun.run_case(case, nsteps=2, kshow=1, U_init=U0)
    nv = C.nvar_of(case.params)
    dims = [b.dims for b in case.blocks]
    U = [_download_layout(blk["U"], nv) for blk in r["stateN"]]
    with tempfile.TemporaryDirectory() as d:
        R.write_restart(d, case.params, U, dims, kstep=2, tt=2.0 * float(case.params["dt_global"]))
        assert open(os.path.join(d, "flow3d.dat"), "rb").read() == r["flow3d.dat"]
        if nv > 5:
            tname = "SA3d.dat" if nv == 6 else "SST3d.dat"
            assert open(os.path.join(d, tname), "rb").read() == r[tname]
        # read the reference's files back: the conserved state over 0..n
        for nm in ("flow3d.dat", "SA3d.dat", "SST3d.dat", "Step_mess.dat"):
            if nm in r:
                open(os.path.join(d, nm), "wb").write(r[nm])
        Ub, step = R.read_restart(d, case.params, dims, nv)
    assert step[0] == 2
    for Ug, Ur, (nx, ny, nz) in zip(Ub, U, dims):
        S = (slice(None), slice(3, nx + 4), slice(3, ny + 4), slice(3, nz + 4))
        np.testing.assert_allclose(Ug[S], Ur[S], rtol=1e-13, atol=0)
    if nv > 5:
        with tempfile.TemporaryDirectory() as d:
            dw = [blk["dw"].data for blk in r["geom"]]
            p = os.path.join(d, "wall_dist.dat")
            R.write_wall_dist(p, dw)
            assert open(p, "rb").read() == r["wall_dist.dat"]
            back = R.read_wall_dist(p, dims)
            for a, b in zip(back, dw):
                assert np.array_equal(a, b)
