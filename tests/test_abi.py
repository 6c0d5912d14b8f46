"""The C ABI: libuph_gpu.so loads and exports every entry point include/uph.h declares.

No GPU work is done here (no call that reaches the device)."""
import ctypes
import os
import re

import pytest

from conftest import ROOT
from uph_amd import solver


def _declared():
    text = ""
    for h in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if h.endswith(".h"):
            with open(os.path.join(ROOT, "include", h)) as f:
                text += f.read()
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


def test_struct_layout_matches_header(tmp_path):
    """Every field of the ctypes mirrors sits at the offset gcc gives the C struct
    of include/uph.h, and the sizes agree."""
    import subprocess
    mirrors = {"uph_scene": solver.UphScene, "uph_opts": solver.UphOpts, "uph_diag": solver.UphDiag}
    src = ['#include <stdio.h>', '#include <stddef.h>',
           f'#include "{os.path.join(ROOT, "include", "uph.h")}"', "int main(void) {"]
    for cname, py in mirrors.items():
        src.append(f'  printf("{cname} size %zu\\n", sizeof({cname}));')
        for fname, _ in py._fields_:
            src.append(f'  printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    src.append("  return 0; }")
    (tmp_path / "layout.c").write_text("\n".join(src) + "\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-o", str(exe), str(tmp_path / "layout.c")], check=True)
    got = {}
    for line in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n"):
        if line:
            c, f, v = line.split()
            got[(c, f)] = int(v)
    for cname, py in mirrors.items():
        assert got[(cname, "size")] == ctypes.sizeof(py), cname
        for fname, _ in py._fields_:
            assert got[(cname, fname)] == getattr(py, fname).offset, (cname, fname)


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    monkeypatch.setattr(solver, "LIB_PATH", str(tmp_path / "nope.so"))
    monkeypatch.setattr(solver, "_lib", None)
    with pytest.raises(RuntimeError):
        solver.lib()


def _system_members():
    """Member names of uph_system in declaration order."""
    with open(os.path.join(ROOT, "include", "uph_system.h")) as f:
        text = f.read()
    body = text[text.index("typedef struct uph_system {"):text.index("} uph_system;")]
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S).split("{", 1)[1]
    names = []
    for decl in body.split(";"):
        decl = decl.strip()
        if not decl:
            continue
        decl = re.sub(r"^(const\s+)?(int|double)\s+", "", decl)
        names += [p.strip().lstrip("*").strip() for p in decl.split(",")]
    return names


REF_SYSTEM_H = "/root/reference/src/lib/System.h"


@pytest.mark.skipif(not os.path.exists(REF_SYSTEM_H), reason="reference tree not present")
def test_system_view_is_binary_compatible_with_the_reference(tmp_path):
    """Every member of uph_system sits at the offset the reference's System has."""
    import subprocess
    names = _system_members()
    assert len(names) > 150 and names[0] == "dim" and names[-1] == "HowManyNeighbors"
    src = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{REF_SYSTEM_H}"',
           f'#include "{os.path.join(ROOT, "include", "uph_system.h")}"', "int main(void) {",
           '  printf("%zu %zu\\n", sizeof(System), sizeof(uph_system));']
    for n in names:
        src.append(f'  printf("{n} %zu %zu\\n", offsetof(System, {n}), offsetof(uph_system, {n}));')
    src.append("  return 0; }")
    (tmp_path / "abi.c").write_text("\n".join(src) + "\n")
    exe = tmp_path / "abi"
    subprocess.run(["gcc", "-o", str(exe), str(tmp_path / "abi.c")], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    a, b = out[0].split()
    assert a == b, ("sizeof", a, b)
    bad = [l for l in out[1:] if l and l.split()[1] != l.split()[2]]
    assert not bad, bad
    assert len([l for l in out[1:] if l]) == len(names)
