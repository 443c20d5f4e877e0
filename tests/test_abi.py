"""The C ABI library loads and exports every entry point include/ declares
(no GPU needed: nothing here calls into the device)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "oxiblas_mi355x.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(oxi_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_hot_path():
    names = declared_functions()
    for required in ("oxi_dgemm", "oxi_dgetrf", "oxi_dgetrs", "oxi_dtrsm", "oxi_dlaswp", "oxi_create"):
        assert required in names


def test_library_exports_every_declared_symbol():
    import oxiblas_mi355x as ox
    if not os.path.exists(ox.LIB_PATH):
        pytest.fail(f"{ox.LIB_PATH} not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(ox.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, f"not exported: {missing}"
    # the Python binding declares exactly the header's functions
    assert sorted(ox.SIGNATURES) == declared_functions()
    ox.load()  # argtypes/restypes bind without touching the device


def test_version_string():
    import oxiblas_mi355x as ox
    L = ox.load()
    assert b"gfx950" in L.oxi_version()
