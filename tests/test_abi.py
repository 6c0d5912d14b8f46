"""The C ABI: libuph_gpu.so loads and exports every entry point include/uph.h declares.

No GPU work is done here (no call that reaches the device)."""
import ctypes
import os
import re

import pytest

from conftest import ROOT
from uph_amd import solver


def _declared():
    with open(os.path.join(ROOT, "include", "uph.h")) as f:
        text = f.read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\(", text)
    skip = {"if", "for", "while", "return", "sizeof"}
    return sorted({n for n in names if n not in skip})


def test_header_declares_the_seam():
    d = _declared()
    assert "SearchNeighbors" in d
    for name in ("uph_create", "uph_step", "uph_get_state", "uph_destroy", "uph_last_error"):
        assert name in d


def test_library_exports_every_declared_symbol():
    if not os.path.exists(solver.LIB_PATH):
        pytest.fail("libuph_gpu.so is not built: run __graft_entry__.build()")
    lib = ctypes.CDLL(solver.LIB_PATH)
    missing = [n for n in _declared() if not hasattr(lib, n)]
    assert not missing, missing
    # the Python mirror binds exactly the declared entry points
    assert sorted(solver.SIGNATURES) == _declared()


def test_struct_layout_matches_header():
    # uph_scene: 3 ints, 6 pointers, then doubles/ints as declared; ctypes pads like C
    s = solver.UphScene
    assert s.pos.offset == 16 and s.box.offset + 32 == ctypes.sizeof(s)
    assert ctypes.sizeof(solver.UphOpts) == 24


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    monkeypatch.setattr(solver, "LIB_PATH", str(tmp_path / "nope.so"))
    monkeypatch.setattr(solver, "_lib", None)
    with pytest.raises(RuntimeError):
        solver.lib()
