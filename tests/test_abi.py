"""The C-ABI library loads and exports every entry point include/*.h declares (CPU only: no
compute call is made here)."""
import ctypes
import re

from conftest import ROOT


def _declared():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        names |= set(re.findall(r"\b(lfric_[a-z0-9_]+)\s*\(", text))
    return names


def test_library_exports_every_declared_symbol():
    from lfric_core_amd import _lib
    L = ctypes.CDLL(str(_lib.LIB_PATH))
    declared = _declared()
    assert len(declared) > 40
    missing = [n for n in sorted(declared) if not hasattr(L, n)]
    assert not missing, missing


def test_python_binding_covers_the_header():
    from lfric_core_amd import _lib
    assert _declared() == set(_lib.SIGNATURES)


def test_package_refuses_to_run_without_the_library(tmp_path, monkeypatch):
    import importlib
    import lfric_core_amd._lib as L
    monkeypatch.setattr(L, "LIB_PATH", tmp_path / "missing.so")
    monkeypatch.setattr(L, "_lib", None)
    try:
        L.lib()
    except ImportError:
        pass
    else:
        raise AssertionError("a missing HIP library must raise")
    importlib.reload(L)
