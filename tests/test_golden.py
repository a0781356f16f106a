"""Golden fixtures written by the compiled reference (tests/golden/make_golden.py).

CPU: the oracle, the boundary-table conversion and the case reader against the reference's own
output.  GPU: the HIP path against the same fixtures, through the C ABI.
"""
import glob
import json
import math
import os

import numpy as np
import pytest

from ocfd_amd import cases as C
import ocfd_amd as A
import oracle

from helpers import rel_err, n_exact

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURES = sorted(glob.glob(os.path.join(HERE, "golden", "*.npz")))
NAMES = [os.path.basename(f)[:-4] for f in FIXTURES]


def load(name):
    z = np.load(os.path.join(HERE, "golden", name + ".npz"), allow_pickle=False)
    params = json.loads(str(z["params_json"]))
    params["Centroid"] = tuple(params["Centroid"])
    blocks = []
    for n in range(1, int(z["nblocks"]) + 1):
        subs = []
        for r in z[f"b{n}_sub"].tolist():
            subs.append(C.Subface(tuple(r[:6]), int(r[6]), tuple(r[7:14]) if r[6] < 0 else None))
        blocks.append(C.Block(z[f"b{n}_x"], z[f"b{n}_y"], z[f"b{n}_z"], subs))
    case = C.Case(name, blocks, params)
    U0 = [z[f"b{n}_U0"] for n in range(1, len(blocks) + 1)]
    return case, U0, z


def parse_inc(text):
    """bc3d.inc as written by convert_inp_inc (sub_convert_inp.f90:94-121): per block its dims,
    its name, the subface count, then two records per subface (9 and 12 integers)."""
    lines = [ln.split() for ln in text.splitlines()][1:]
    nblk = int(lines[0][0])
    pos, out = 1, []
    for _ in range(nblk):
        pos += 2
        nsub = int(lines[pos][0]); pos += 1
        recs = []
        for _ in range(nsub):
            recs.append([int(v) for v in lines[pos]] + [int(v) for v in lines[pos + 1]])
            pos += 2
        out.append(np.array(recs, dtype=np.int32))
    return out


def _interior(o, blk, nm, dims, nv=0):
    nx, ny, nz = dims
    if nv:
        return np.stack([oracle.pad_view(o.get(blk, f"{nm}{m + 1}"), (1, 1, 1), (nx - 1, ny - 1, nz - 1)) for m in range(nv)])
    return oracle.pad_view(o.get(blk, nm), (1, 1, 1), (nx - 1, ny - 1, nz - 1))


def _check(label, got, ref, tol):
    e = rel_err(got, ref)
    ne, n = n_exact(got, ref)
    print(f"  {label:8s} rel={e:.2e} exact={ne}/{n}")
    assert e <= tol, f"{label}: rel {e:.3e} > {tol}"


def _force_ok(got_row, ref_row):
    """force.log holds CL, CD, CS, CMx, CMy, CMz in E20.8 (0.dddddddd: 8 significant digits, so a
    rounding error up to 5e-8 relative)."""
    ref = np.asarray(ref_row, dtype=np.float64)
    scale = max(np.max(np.abs(ref)), 1e-300)
    return bool(np.all(np.abs(np.asarray(got_row) - ref) <= 5.1e-8 * np.maximum(np.abs(ref), 1e-6 * scale)))


@pytest.mark.parametrize("name", NAMES)
def test_bc_tables_match_reference_conversion(name):
    case, _, z = load(name)
    ref = parse_inc(str(z["bc3d_inc"]))
    got = A.inc_tables(case)
    assert len(got) == len(ref)
    for g, r in zip(got, ref):
        assert np.array_equal(g, r)


@pytest.mark.parametrize("name", NAMES)
def test_oracle_matches_reference(name):
    """The oracle reproduces the reference's per-cell Res, DU, dt, mu_t and state after the first
    iteration, the state after the last one, and the norms of every iteration."""
    print(f"\n[{name}]")
    case, U0, z = load(name)
    nv = C.nvar_of(case.params)
    o = oracle.Oracle(case)
    o.init_flow(); o.set_interior(U0); o.bc_halo()
    o.step(1)
    for n, b in enumerate(case.blocks, start=1):
        _check("Res1", _interior(o, n, "Res", b.dims, nv), z[f"b{n}_Res1"], 1e-13)
        _check("DU1", _interior(o, n, "DU", b.dims, nv), z[f"b{n}_DU1"], 1e-13)
        _check("dt1", _interior(o, n, "dt", b.dims), z[f"b{n}_dt1"], 1e-13)
        if f"b{n}_mut1" in z.files:
            _check("mut1", _interior(o, n, "mu_t", b.dims), z[f"b{n}_mut1"], 1e-13)
        _check("U1", _interior(o, n, "U", b.dims, nv), z[f"b{n}_U1"], 1e-13)
    rows = [o.monitor()]
    for _ in range(int(z["nsteps"]) - 1):
        o.step(1)
        rows.append(o.monitor())
    for n, b in enumerate(case.blocks, start=1):
        _check("UN", _interior(o, n, "U", b.dims, nv), z[f"b{n}_UN"], 1e-12)
        assert o.dims(n)[7] == int(z[f"b{n}_overlimit"])
    np.testing.assert_allclose([r["res_max"] for r in rows], z["norms_max"], rtol=1e-12)
    np.testing.assert_allclose([r["res_rms"] for r in rows], z["norms_rms"], rtol=1e-12)
    for r, ref in zip(rows, z["force_log"]):
        got = [r["CL"], r["CD"], r["CS"], r["Ft"][6], r["Ft"][7], r["Ft"][8]]
        assert _force_ok(got, ref[1:7]), (got, ref)


@pytest.mark.gpu
@pytest.mark.parametrize("name", NAMES)
def test_hip_path_matches_reference_fixture(name):
    """The HIP path (device geometry from the node coordinates) against the reference's output:
    per-cell residual within 1e-11 after the first iteration, the state within 1e-9 after the last,
    norms and force coefficients of the last iteration."""
    print(f"\n[{name}]", flush=True)
    case, U0, z = load(name)
    nv = C.nvar_of(case.params)
    s = A.solver_from_case(case, states=U0)
    s.step(1)
    for n, b in enumerate(case.blocks, start=1):
        nx, ny, nz = b.dims
        I = (slice(None), slice(2, nx + 1), slice(2, ny + 1), slice(2, nz + 1))
        _check("Res1", s.field(n, A.F_RES)[I], z[f"b{n}_Res1"], 1e-11)
        _check("U1", s.download_U(n)[:, 4:nx + 3, 4:ny + 3, 4:nz + 3], z[f"b{n}_U1"], 1e-11)
    s.step(int(z["nsteps"]) - 1)
    for n, b in enumerate(case.blocks, start=1):
        nx, ny, nz = b.dims
        _check("UN", s.download_U(n)[:, 4:nx + 3, 4:ny + 3, 4:nz + 3], z[f"b{n}_UN"], 1e-9)
        assert s.status(n)[0] == int(z[f"b{n}_overlimit"])
    mx, rms = s.norms()
    np.testing.assert_allclose(mx, z["norms_max"][-1], rtol=1e-9)
    np.testing.assert_allclose(rms, z["norms_rms"][-1], rtol=1e-9)
    f = s.forces()
    assert _force_ok([f[9], f[10], f[11], f[6], f[7], f[8]], z["force_log"][-1][1:7])


def test_converged_sa_fixture_oracle_matches_reference():
    """The converged SA C-grid fixture (tests/golden/make_converged.py, 30000 iterations): the
    oracle's CL, CD and CS equal the compiled reference's full-precision print within 1e-12
    relative, its moments Ft0(7:9) the reference's force.log to the log's 8 digits, and both ran
    to a residual drop below 1e-10."""
    fx = json.load(open(os.path.join(HERE, "golden", "converged_sa_cgrid_65x25x9.json")))
    ref, orc = fx["reference"], fx["oracle"]
    for k in ("CL", "CD"):
        assert abs(ref[k] - orc[k]) <= 1e-12 * abs(ref[k]), (k, ref[k], orc[k])
    assert abs(ref["CS"] - orc["CS"]) <= 1e-12 * abs(ref["CL"])
    log = ref["forcelog_last"]
    assert int(log[0]) == fx["iterations"]
    for q, m in enumerate(orc["Ft"][6:9]):
        # force.log's E20.8: an 8-digit mantissa in [0.1, 1), so half a unit of its last digit
        half_unit = 0.5 * 10.0 ** (math.floor(math.log10(abs(m))) + 1 - 8)
        assert abs(log[4 + q] - m) <= half_unit, (q, log[4 + q], m)
    assert ref["res_rms1"][-1][1] / orc["res_rms1_first"] < 1e-10
    assert orc["res_rms1_last"] / orc["res_rms1_first"] < 1e-10
