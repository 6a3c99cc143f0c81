"""The C-ABI library builds, loads without a GPU and exports every function
include/*.h declares (no compute calls)."""
import ctypes as C
import glob
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "dmd_amd", "libdmdx.so")
H5MD_LIB = os.path.join(ROOT, "dmd_amd", "libdmdx_h5md.so")
H5MD_HEADER = os.path.join(ROOT, "include", "dmdx_h5md.h")


def declared_functions(headers=None):
    """dmdx_* functions of include/*.h; dmdx_h5md.h belongs to libdmdx_h5md.so."""
    if headers is None:
        headers = [h for h in glob.glob(os.path.join(ROOT, "include", "*.h")) if h != H5MD_HEADER]
    names = set()
    for h in headers:
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w\s\*]*?\b(dmdx_\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return sorted(names)


def test_header_declares_entry_points():
    names = declared_functions()
    assert "dmdx_engine_create" in names and "dmdx_last_error" in names
    assert len(names) >= 15


@pytest.mark.skipif(not os.path.exists(LIB), reason="libdmdx.so not built")
def test_library_exports_every_declared_symbol():
    lib = C.CDLL(LIB)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


@pytest.mark.skipif(not os.path.exists(H5MD_LIB), reason="libdmdx_h5md.so not built (no HDF5)")
def test_h5md_library_exports_every_declared_symbol():
    C.CDLL(LIB)
    lib = C.CDLL(H5MD_LIB)
    names = declared_functions([H5MD_HEADER])
    assert "dmdx_h5md_run" in names
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


@pytest.mark.skipif(not os.path.exists(LIB), reason="libdmdx.so not built")
def test_last_error_is_empty_string_initially():
    lib = C.CDLL(LIB)
    lib.dmdx_last_error.restype = C.c_char_p
    assert isinstance(lib.dmdx_last_error(), bytes)


def test_system_desc_layout_matches_header():
    """The ctypes mirror has one field per member of dmdx_system_desc."""
    from dmd_amd import _abi
    text = open(os.path.join(ROOT, "include", "dmdx.h")).read()
    body = text[text.index("typedef struct dmdx_system_desc"):text.index("} dmdx_system_desc;")]
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    decl = []
    for line in body.splitlines()[1:]:
        line = line.strip().rstrip(";")
        if not line:
            continue
        names = line.split(None, 1)[1] if not line.startswith("const") else line.split(None, 2)[2]
        for part in names.split(","):
            decl.append(part.strip().lstrip("*").split("[")[0])
    assert decl == [f[0] for f in _abi.SystemDesc._fields_]
