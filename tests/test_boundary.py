"""The C-ABI library: builds, loads without a GPU, exports every declared symbol.

No compute call is made here (these tests run on machines without a GPU).
"""
import ctypes
import re
import subprocess
from pathlib import Path

import pytest

import mpskit_amd
from mpskit_amd import _lib

ROOT = Path(__file__).resolve().parents[1]


def declared_symbols():
    text = (ROOT / "include" / "mpskit_hip.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mk_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_binding_table():
    assert declared_symbols() == sorted(_lib.EXPORTED)


def test_library_exports_every_declared_symbol():
    if not _lib.LIB_PATH.exists():
        pytest.skip("library not built (run python mpskit.jl_amd/build.py)")
    out = subprocess.run(["nm", "-D", "--defined-only", str(_lib.LIB_PATH)], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r"\bT (mk_[a-z0-9_]+)$", out, flags=re.M))
    missing = [s for s in declared_symbols() if s not in exported]
    assert not missing, missing


def test_library_loads_and_reports_version():
    if not _lib.LIB_PATH.exists():
        pytest.skip("library not built")
    lib = _lib.load()
    assert lib.mk_version() >= 100
    assert isinstance(lib.mk_last_error(), bytes)


def test_errors_map_to_exceptions():
    if not _lib.LIB_PATH.exists():
        pytest.skip("library not built")
    lib = _lib.load()
    # null-handle argument errors are detected on the host, before any GPU work
    st = lib.mk_ctx_set_stream(None, None)
    assert st == 2
    with pytest.raises(_lib.MKError) as ei:
        _lib.check(st)
    assert "null context" in str(ei.value)
    with pytest.raises(_lib.SpaceMismatch):
        _lib.check(1)
    h = ctypes.c_void_p()
    assert lib.mk_apply(None, None, None, 1.0, 0.0, 0) == 2


def test_package_has_no_oracle_dependency():
    """The product package must not import the oracle (no CPU fallback)."""
    for p in (ROOT / "mpskit.jl_amd").rglob("*.py"):
        src = p.read_text()
        assert "oracle" not in re.sub(r"#.*", "", src).replace('"""', ""), p


def test_integration_ccall_shims_match_the_header():
    """Every Julia ccall shown in INTEGRATION.md passes as many arguments as the
    header declares for that entry point."""
    hdr = re.sub(r"/\*.*?\*/", "", (ROOT / "include" / "mpskit_hip.h").read_text(), flags=re.S)
    decl = {m.group(1): len([a for a in m.group(2).split(",") if a.strip() and a.strip() != "void"])
            for m in re.finditer(r"\b(mk_\w+)\s*\(([^;()]*?)\);", hdr, flags=re.S)}
    text = (ROOT / "INTEGRATION.md").read_text()
    checked = 0
    for m in re.finditer(r"ccall\(\(:(mk_\w+),\s*\w+\),\s*\w+,\s*", text):
        rest = text[m.end():]
        assert rest.startswith("("), m.group(1)
        depth, end = 0, 0
        for k, ch in enumerate(rest):
            depth += ch == "("
            depth -= ch == ")"
            if depth == 0:
                end = k
                break
        types = [t for t in rest[1:end].split(",") if t.strip()]
        assert m.group(1) in decl, m.group(1)
        assert len(types) == decl[m.group(1)], (m.group(1), len(types), decl[m.group(1)])
        checked += 1
    assert checked >= 10
