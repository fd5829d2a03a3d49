"""The C-ABI libraries load and export every symbol their headers declare
(no compute calls: this runs without a GPU)."""
import ctypes
import os
import re

import pytest

import lpa_amd as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    txt = open(os.path.join(ROOT, "include", header)).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(lpah?_[a-z0-9_]+)\s*\(", txt)))


@pytest.mark.parametrize("header,lib", [("lpa_amd.h", "liblpa_hip.so"), ("lpa_host.h", "liblpa_host.so")])
def test_every_declared_symbol_is_exported(header, lib):
    so = ctypes.CDLL(os.path.join(ROOT, "lpa_amd", lib))
    names = declared(header)
    assert len(names) > 10
    missing = [n for n in names if not hasattr(so, n)]
    assert not missing, missing


def test_python_binding_covers_engine_abi():
    assert set(declared("lpa_amd.h")) == set(L.engine.EXPORTS)


def test_engine_library_has_no_cpu_fallback(tmp_path, monkeypatch):
    """without the HIP library the product refuses to run"""
    monkeypatch.setattr(L.engine, "_LIB", None)
    monkeypatch.setattr(L.engine, "library_path", lambda: str(tmp_path / "missing.so"))
    with pytest.raises(L.LpaError):
        L.engine.load()
