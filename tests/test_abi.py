"""The C ABI library loads and exports every entry point include/*.h declares.

No compute call is made (no GPU needed): the library is opened with dlopen
and each declared function name is looked up.
"""
import ctypes as C
import os
import re

import pytest

from .conftest import ROOT

HEADERS = ["osqp_amd.h", "osqp_amd_algebra.h", "osqp_amd_serve.h"]
DECL = re.compile(r"^\s*(?:const\s+)?(?:struct\s+|enum\s+)?[A-Za-z_][\w]*\s*\**\s+\**\s*([A-Za-z_]\w*)\s*\(", re.M)


def declared_functions(header):
    path = os.path.join(ROOT, "include", header)
    if not os.path.exists(path):
        return []
    text = open(path).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    # drop the struct linsys_solver body (function-pointer members)
    text = re.sub(r"struct linsys_solver\s*\{.*?\};", "", text, flags=re.S)
    names = set(DECL.findall(text))
    return sorted(n for n in names if n not in {"if", "while", "for", "return", "sizeof"})


def test_library_exports_declared_symbols():
    lib_path = os.path.join(ROOT, "osqp_amd", "libosqp_amd.so")
    assert os.path.exists(lib_path), "build first: python -m osqp_amd.build"
    lib = C.CDLL(lib_path)
    missing = []
    total = 0
    for h in HEADERS:
        for name in declared_functions(h):
            total += 1
            if not hasattr(lib, name):
                missing.append(f"{h}:{name}")
    assert total >= 90, f"parsed only {total} declarations"
    assert not missing, f"symbols declared but not exported: {missing}"


@pytest.mark.parametrize("header", HEADERS[:2])
def test_headers_compile_as_c(header, tmp_path):
    """The public headers are plain C (no C++ or HIP types in the signatures)."""
    import subprocess
    src = tmp_path / "t.c"
    src.write_text(f'#include "{header}"\nint main(void) {{ return 0; }}\n')
    r = subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), "-c",
                        str(src), "-o", str(tmp_path / "t.o")], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
