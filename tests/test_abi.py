"""The C ABI library loads (no GPU needed to load it) and exports every entry point declared
in include/fastmi.h; the Python binding declares the same set."""
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "fastmi.h")


def declared():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(fm_\w+)\s*\(", txt)))


def test_header_declares_entry_points():
    names = declared()
    for n in ("fm_ctx_create", "fm_zone_add", "fm_compute", "fm_zone_substep", "fm_apply_bc",
              "fm_residual_norms"):
        assert n in names


def test_library_exports_all_declared_symbols():
    from fast_amd import _lib
    if not os.path.exists(_lib.LIBPATH):
        pytest.fail("libfastmi.so not built: run __graft_entry__.build()")
    lib = _lib.load()
    for n in declared():
        assert hasattr(lib, n), n
    assert set(declared()) == set(_lib.SIGNATURES)


def test_missing_library_fails_loudly(tmp_path):
    from fast_amd import _lib
    with pytest.raises(_lib.FastMIError):
        _lib.load(str(tmp_path / "libfastmi.so"))
