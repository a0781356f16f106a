"""C ABI of libocfd_ec.so: the library loads and exports every entry point include/ocfd_ec.h declares."""
import ctypes
import os
import re

import pytest

import ocfd_amd as A

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    txt = open(os.path.join(ROOT, "include", "ocfd_ec.h")).read()
    return sorted(set(re.findall(r"\b(ocfd_[A-Za-z_0-9]+)\s*\(", txt)))


def test_header_and_python_agree():
    assert declared() == sorted(A.EXPORTS)


@pytest.mark.skipif(not os.path.exists(A.LIB_PATH), reason="libocfd_ec.so not built")
def test_library_exports_every_symbol():
    L = ctypes.CDLL(A.LIB_PATH)
    missing = [s for s in declared() if not hasattr(L, s)]
    assert not missing, missing


def test_params_layout_matches_header():
    # 31 doubles + 3 + 3 then 13 int32
    assert ctypes.sizeof(A.Params) == 8 * (29 + 3 + 3) + 4 * 13 + 4  # trailing pad to 8
