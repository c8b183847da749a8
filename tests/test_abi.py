"""The C ABI: every entry point declared in include/*.h is exported by the
engine library, and the Python binding resolves them (no GPU needed)."""
import ctypes
import glob
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "william_ci_amd", "libtcg.so")


def declared_symbols():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b(tcg_\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return sorted(names)


def test_header_declares_entry_points():
    names = declared_symbols()
    for must in ("tcg_create", "tcg_set_basis", "tcg_eri_build", "tcg_jk_stored", "tcg_one_electron",
                 "tcg_last_error", "tcg_destroy"):
        assert must in names, names


@pytest.mark.skipif(not os.path.exists(LIB), reason="engine library not built")
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(LIB)
    missing = [n for n in declared_symbols() if not hasattr(lib, n)]
    assert not missing, missing


@pytest.mark.skipif(not os.path.exists(LIB), reason="engine library not built")
def test_python_binding_resolves():
    from william_ci_amd import engine
    engine._lib = None
    engine.lib()


def test_engine_refuses_without_library(monkeypatch, tmp_path):
    from william_ci_amd import engine
    monkeypatch.setattr(engine, "LIB_PATH", str(tmp_path / "missing.so"))
    monkeypatch.setattr(engine, "_lib", None)
    with pytest.raises(engine.EngineError):
        engine.lib()
