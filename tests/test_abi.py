"""The C ABI (CPU only, no compute calls): libsoftbox_hip.so loads, exports every function that
include/softbox_hip.h declares, and the ctypes binding gives each one the header's argument count."""
import ctypes as C
import os
import re

import pytest

from softbox_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "softbox_hip.h")


def _declarations():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    decls = {}
    for m in re.finditer(r"\b(?:int|void|size_t|const\s+char\s*\*)\s*\**\s*(sbx_\w+)\s*\(([^;{]*?)\)\s*;", text,
                         flags=re.S):
        name, args = m.group(1), " ".join(m.group(2).split())
        decls[name] = 0 if args in ("", "void") else len(args.split(","))
    return decls


def test_header_parses():
    d = _declarations()
    assert len(d) > 50 and "sbx_net_step" in d and "sbx_grid_reachable" in d


def test_library_exports_every_declared_symbol():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libsoftbox_hip.so not built")
    lib = C.CDLL(_lib.LIB_PATH)
    missing = [name for name in _declarations() if not hasattr(lib, name)]
    assert not missing, f"declared but not exported: {missing}"


def test_binding_matches_header_arity():
    d = _declarations()
    unbound = sorted(set(d) - set(_lib._SIGS))
    assert not unbound, f"declared but not bound in softbox_amd/_lib.py: {unbound}"
    extra = sorted(set(_lib._SIGS) - set(d))
    assert not extra, f"bound but not declared: {extra}"
    wrong = {k: (len(_lib._SIGS[k]), d[k]) for k in d if len(_lib._SIGS[k]) != d[k]}
    assert not wrong, f"argument count (binding, header) differs: {wrong}"


def test_load_raises_without_library(monkeypatch, tmp_path):
    """The product path has no fallback: a missing library is an error."""
    monkeypatch.setattr(_lib, "LIB_PATH", str(tmp_path / "absent.so"))
    monkeypatch.setattr(_lib, "_lib", None)
    with pytest.raises(RuntimeError):
        _lib.load()
