"""ocfd_res_norms_noalloc: the residual norms through the entry point that allocates nothing (its
scratch is the library's fixed device array), against ocfd_res_norms through the C ABI.

Two calls in a row share the one scratch array of the GPU: the first on a 6-block cut channel,
the second on a smaller single-block case, so the second call finds the first one's table and
parts in the array.  Both must equal ocfd_res_norms bit for bit (the same kernels over the same
block order).  A context above the array's 64-block limit is refused with OCFD_EARG.
"""
import numpy as np
import pytest

import ocfd_amd as A
from ocfd_amd import cases as C


def _solver(case, seed):
    U0 = C.perturbed_state(case, seed=seed, amp=1e-3)
    s = A.solver_from_case(case, states=U0)
    s.step(2)
    return s


@pytest.mark.gpu
def test_norms_noalloc_equals_norms_twice_second_smaller():
    big = C.cut_blocks(C.channel_bump(49, 17, 9), [[[17, 33], [9], []]])
    small = C.channel_bump(25, 13, 7)
    assert len(big.blocks) == 6 and len(small.blocks) == 1
    for case, seed in ((big, 3), (small, 4)):
        print(f"\n[{len(case.blocks)} blocks, {case.ncell} cells]", flush=True)
        s = _solver(case, seed)
        mx0, rms0 = s.norms_noalloc()
        mx1, rms1 = s.norms()
        assert np.all(np.isfinite(rms0)) and np.all(rms0 > 0)
        assert np.array_equal(mx0, mx1) and np.array_equal(rms0, rms1)
        s.close()


@pytest.mark.gpu
def test_norms_noalloc_refuses_more_blocks_than_its_scratch():
    case = C.cut_blocks(C.channel_bump(53, 17, 5), [[[5 + 4 * q for q in range(12)], [4, 7, 10, 13], []]])
    assert len(case.blocks) == 65
    s = A.solver_from_case(case)
    with pytest.raises(A.OcfdError) as e:
        s.norms_noalloc()
    assert e.value.code == A.OCFD_EARG
    mx, rms = s.norms()
    assert np.all(np.isfinite(mx))
    s.close()
