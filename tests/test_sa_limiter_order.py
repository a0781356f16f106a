"""limit_flow_SA (sub_limitflow.f90:182-231) replaces U6 above LSAmax in the reference's serial
(k, j, i) loop, each replacement reading its earlier neighbours after the pass and its later ones
before it.  The device does the flagged cells without flagged neighbours in parallel, then the rest
by wavefronts of w = i + 2j + 4k, each wavefront in parallel.  This checks, on random fields with
clustered flagged cells and cells below SAmin, that the two orders give the same values bit for
bit (a model of the pass in numpy, independent of the library and of the oracle)."""
from __future__ import annotations

import numpy as np
import pytest

SAMIN, LSAMAX = 1e-8, 5.0


def _replace(cur, orig, i, j, k):
    nx, ny, nz = cur.shape
    s0, kn = 0.0, 0
    for ka in (-1, 0, 1):
        for ja in (-1, 0, 1):
            for ia in (-1, 0, 1):
                i1, j1, k1 = i + ia, j + ja, k + ka
                if 0 <= i1 < nx and 0 <= j1 < ny and 0 <= k1 < nz:
                    later = (k1, j1, i1) > (k, j, i)
                    v = orig[i1, j1, k1] if later else cur[i1, j1, k1]
                    if SAMIN <= v <= LSAMAX:
                        s0 += v
                        kn += 1
    return LSAMAX if kn == 0 else s0 / kn


def _serial(orig):
    cur = orig.copy()
    nx, ny, nz = orig.shape
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                if cur[i, j, k] < SAMIN:
                    cur[i, j, k] = SAMIN
                if cur[i, j, k] > LSAMAX:
                    cur[i, j, k] = _replace(cur, orig, i, j, k)
    return cur


def _wavefront(orig):
    flag = orig > LSAMAX
    cur = np.where(orig < SAMIN, SAMIN, orig)
    nx, ny, nz = orig.shape
    idx = np.argwhere(flag)
    left = flag.copy()
    for i, j, k in idx:
        lo, hi = np.maximum([i - 1, j - 1, k - 1], 0), np.minimum([i + 2, j + 2, k + 2], [nx, ny, nz])
        if flag[lo[0]:hi[0], lo[1]:hi[1], lo[2]:hi[2]].sum() == 1:      # only itself: isolated
            cur[i, j, k] = _replace(cur, orig, i, j, k)
            left[i, j, k] = False
    rest = np.argwhere(left)
    if len(rest):
        lo, hi = rest.min(0), rest.max(0)
        for w in range(lo[0] + 2 * lo[1] + 4 * lo[2], hi[0] + 2 * hi[1] + 4 * hi[2] + 1):
            new = {}
            for j in range(lo[1], hi[1] + 1):
                for k in range(lo[2], hi[2] + 1):
                    i = w - 2 * j - 4 * k
                    if lo[0] <= i <= hi[0] and left[i, j, k]:
                        new[(i, j, k)] = _replace(cur, orig, i, j, k)
            for key, v in new.items():     # a wavefront's cells never read one another
                cur[key] = v
    return cur


def test_no_two_cells_of_a_wavefront_are_neighbours_and_earlier_means_smaller_w():
    for dk in (-1, 0, 1):
        for dj in (-1, 0, 1):
            for di in (-1, 0, 1):
                if (di, dj, dk) == (0, 0, 0):
                    continue
                dw = di + 2 * dj + 4 * dk
                assert dw != 0
                assert (dw < 0) == ((dk, dj, di) < (0, 0, 0))


@pytest.mark.parametrize("seed", range(12))
def test_wavefront_order_equals_serial_loop(seed):
    rng = np.random.default_rng(seed)
    shape = tuple(int(x) for x in rng.integers(2, 10, 3))
    orig = rng.uniform(0.5, 4.5, shape)
    orig[rng.random(shape) < 0.1] = 1e-9
    orig[rng.random(shape) < rng.uniform(0.1, 0.8)] = rng.uniform(5.5, 9.0)
    assert np.array_equal(_serial(orig), _wavefront(orig))
