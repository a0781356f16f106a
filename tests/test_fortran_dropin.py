"""The reference's own Fortran host with the HIP library dropped in (oracle/_ref/dropin_driver).

The driver (oracle/dropin_driver.f90, linked against the reference's objects, the Fortran binding
opencfd-ec_amd/fortran/ocfd_ec.f90 and libocfd_ec.so) sets a case up with the reference's routines,
hands its Global_Var parameters, Block_TYPE geometry and bc_msg tables to the library through
iso_c_binding, advances the same state on the device and with the reference's NS_Time_advance,
and reports the differences.  Skipped where the driver has not been built
(make -f oracle/ref.mk dropin).
"""
import os
import subprocess
import tempfile

import numpy as np
import pytest

from helpers import CASES
from ocfd_amd import cases as C

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "oracle", "_ref", "dropin_driver")

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not os.access(DRIVER, os.X_OK), reason="oracle/_ref/dropin_driver not built")]


def run_dropin(case, U0, nsteps):
    with tempfile.TemporaryDirectory(prefix="ocfd_dropin_") as d:
        C.write_case(case, d)
        C.write_initial_state(os.path.join(d, "U_init.bin"), U0)
        r = subprocess.run(["bash", "-c", f"ulimit -s unlimited; {DRIVER} {nsteps}"], cwd=d, capture_output=True,
                           text=True, timeout=600)
        assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
        rows = [ln.split() for ln in open(os.path.join(d, "dropin.txt")) if ln.strip()]
        force = [ln.split() for ln in open(os.path.join(d, "force.log")) if ln.strip()]
    return rows, force


@pytest.mark.parametrize("name", ["channel_cut6", "cgrid_sa_weno5_roe", "cgrid_sst_muscl3_ausm", "cgrid_sa_limiter"])
def test_fortran_host_dropin(name):
    print(f"\n[{name}]", flush=True)
    cs = CASES[name]
    case = cs.builder()
    U0 = C.perturbed_state(case, seed=cs.seed, amp=cs.amp)
    nsteps = 1 + cs.more
    rows, force = run_dropin(case, U0, nsteps)
    res = [float(r[3]) for r in rows if r[0] == "res"]
    uu = [float(r[3]) for r in rows if r[0] == "u"]
    print(f"  Res rel max {max(res):.2e}   U rel max {max(uu):.2e}")
    assert max(res) <= 1e-11
    assert max(uu) <= 1e-9
    ng = np.array([float(v) for v in next(r for r in rows if r[0] == "norms_gpu")[1:]])
    nr = np.array([float(v) for v in next(r for r in rows if r[0] == "norms_ref")[1:]])
    np.testing.assert_allclose(ng, nr, rtol=1e-9)
    fg = np.array([float(v) for v in next(r for r in rows if r[0] == "force_gpu")[1:]])
    fr = np.array([float(v) for v in force[-1][1:7]])
    scale = max(np.max(np.abs(fr)), 1e-300)
    assert np.all(np.abs(fg - fr) <= 5.1e-8 * np.maximum(np.abs(fr), 1e-6 * scale)), (fg, fr)
    # the glue's force.log record (ocfd_host_forces, written before the reference's own) at the
    # same step and to the 8 digits of the format
    assert len(force) == 2 and force[0][0] == force[1][0]
    fl = np.array([float(v) for v in force[0][1:7]])
    assert np.all(np.abs(fl - fr) <= 1.1e-7 * np.maximum(np.abs(fr), 1e-6 * scale)), (fl, fr)
