"""CPU checks of the C ABI: the built library loads and exports every entry
point declared in include/mdx.h (no GPU calls)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "mdx.h")).read()
    return sorted(set(re.findall(r"\b(mdx_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_abi():
    syms = declared_symbols()
    for s in ("mdx_create", "mdx_nlist_get", "mdx_forces_compute", "mdx_pppm_set", "mdx_e3b_set", "mdx_run"):
        assert s in syms


def test_library_exports_every_declared_symbol():
    from md_engine_amd import _lib

    lib_path = _lib.LIB_PATH
    if not os.path.exists(lib_path):
        pytest.fail("libmdx.so is not built (run __graft_entry__.build())")
    L = ctypes.CDLL(lib_path)
    missing = [s for s in declared_symbols() if not hasattr(L, s)]
    assert not missing, missing
    assert set(declared_symbols()) == set(_lib.SYMBOLS)


def test_python_binding_declares_signatures():
    from md_engine_amd import _lib

    L = _lib.load()
    assert L.mdx_version() == 1


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    from md_engine_amd import _lib

    monkeypatch.setattr(_lib, "_lib", None)
    monkeypatch.setattr(_lib, "LIB_PATH", str(tmp_path / "nope.so"))
    with pytest.raises(_lib.MdxError):
        _lib.load()
