"""C ABI of libocfd_ec.so: the library loads and exports every entry point include/ocfd_ec.h declares."""
import ctypes
import os
import re
import subprocess

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


def test_params_layout_matches_header(tmp_path):
    """The ctypes mirror of ocfd_params has the C struct's size and every member's offset (a C
    program compiled against include/ocfd_ec.h prints them)."""
    names = [f[0] for f in A.Params._fields_]
    src = tmp_path / "layout.c"
    src.write_text("#include <stdio.h>\n#include <stddef.h>\n#include \"ocfd_ec.h\"\nint main(void) {\n"
                   '  printf("size %zu\\n", sizeof(ocfd_params));\n'
                   + "".join(f'  printf("{n} %zu\\n", offsetof(ocfd_params, {n}));\n' for n in names)
                   + "  return 0;\n}\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = dict(line.split() for line in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.splitlines())
    assert int(out["size"]) == ctypes.sizeof(A.Params)
    for n in names:
        assert int(out[n]) == getattr(A.Params, n).offset, n
