"""Mixed precision (precision=1): reconstruction, face fluxes and the SA and SST terms in FP32;
the residual assembly, spectral radii, LU-SGS, update, limiter and norms in FP64.

The FP64 path is the drop-in and is held to the reference bit for bit elsewhere.  The mixed path
is held to what FP32 arithmetic allows:
  * one residual from the same state: the difference to the FP64 residual, over all cells, is a
    few FP32 roundings of the face fluxes (relative L2 norm per component below RES_TOL on a
    state perturbed by 1 %, where the residual is of the size of the perturbation);
  * converged force coefficients: after both paths have converged, CL and CD agree within
    FORCE_TOL relative (the target stated for the mixed form).

HLLC is left out of the residual comparison: the reference HLLC (sub_flux_split.f90:67 uses vvr in
the right star state's fourth component) is discontinuous where the contact speed S* crosses
zero, so on faces with S* within rounding of zero FP32 and FP64 pick different branches and the
face flux jumps by S_R rho (v - w), an O(1) difference that is the reference's, not rounding.

The residual of the mixed path carries the FP32 rounding of the face fluxes, about 1e-7 of the
flux magnitude, so its norm stops falling near 1e-3 of its first-iteration value on the C-grids
while the forces have converged; convergence of a mixed solve is therefore judged by its forces.
"""
import numpy as np
import pytest

import ocfd_amd as A
from ocfd_amd import cases as C
from helpers import CASES

RES_TOL = 1e-4
FORCE_TOL = 1e-4

pytestmark = pytest.mark.gpu


def _res(case, U0, precision):
    s = A.solver_from_case(case, precision=precision, states=U0)
    s.residual()
    s.sync()
    out = [s.field(no, A.F_RES) for no in range(1, len(case.blocks) + 1)]
    s.close()
    return out


@pytest.mark.parametrize("name", ["channel_roe_muscl3", "channel_sw_ud3", "channel_vl_muscl2c", "channel_roe_weno5_char",
                                  "channel_roe_ud5_cons", "cgrid_sa_weno5_roe", "cgrid_laminar_vl", "cgrid_sa_cut12",
                                  "cgrid_sst_muscl3_ausm"])
def test_mixed_residual_tracks_fp64(name):
    cs = CASES[name]
    case = cs.builder()
    U0 = C.perturbed_state(case, seed=cs.seed, amp=1e-2)
    r64, r32 = _res(case, U0, 0), _res(case, U0, 1)
    nv = r64[0].shape[0]
    for m in range(nv):
        a = np.concatenate([r[m].ravel() for r in r64])
        b = np.concatenate([r[m].ravel() for r in r32])
        rel = np.linalg.norm(b - a) / np.linalg.norm(a)
        assert rel < RES_TOL, (name, m, rel)
        assert rel > 0.0, (name, m, "mixed path gave the FP64 residual: FP32 kernels not used")


def test_mixed_converged_forces_match_fp64():
    """97x33x9 laminar NACA 0012 C-grid (Re 5000, Ma 0.5, AoA 2, MUSCL3 + Roe, CFL 5): after 10000
    iterations the FP64 residual has dropped below 1e-5 of its first value and CL, CD move by less
    than 1e-5 relative over the last 1000; the mixed path's CL and CD agree with the FP64 ones
    within FORCE_TOL."""
    case = C.naca_cgrid(97, 33, 9, Ma=0.5, AoA=2.0, Re=5e3, wall_spacing=5e-3, Iflag_turbulence_model=0,
                        Iflag_Scheme=5, Iflag_Flux=4, CFL=5.0)
    S = [A.solver_from_case(case, precision=p) for p in (0, 1)]
    for s in S:
        s.set_graph(True)
        s.step(1)
    r0 = S[0].norms()[1][0]
    for s in S:
        s.step(8999)
    f_prev = S[0].forces()
    for s in S:
        s.step(1000)
    f64, f32 = S[0].forces(), S[1].forces()
    drop = S[0].norms()[1][0] / r0
    for s in S:
        s.close()
    print(f"\nFP64 drop {drop:.2e}; CL {f64[9]:.10f} / {f32[9]:.10f}; CD {f64[10]:.10f} / {f32[10]:.10f}")
    assert drop < 1e-5, drop
    for q in (9, 10):                                   # CL, CD
        assert abs(f64[q] - f_prev[q]) < 1e-5 * abs(f64[q])       # the FP64 forces have settled
        assert abs(f32[q] - f64[q]) <= FORCE_TOL * abs(f64[q]), (q, f32[q], f64[q])


def test_batch_needs_one_precision():
    case = CASES["channel_roe_muscl3"].builder()
    a, b = A.solver_from_case(case, precision=0), A.solver_from_case(case, precision=1)
    bt = A.Batch()
    bt.add(a)
    with pytest.raises(A.OcfdError, match="precision"):
        bt.add(b)
    bt.close(); a.close(); b.close()
