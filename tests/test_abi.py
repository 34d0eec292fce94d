"""The C-ABI library loads and exports every entry point include/sailor_hip.h
declares (no GPU work is launched)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            src = open(os.path.join(ROOT, "include", fn)).read()
            src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
            names |= set(re.findall(r"\b(sailor_[a-z0-9_]+)\s*\(", src))
    return names


def test_header_declares_entry_points():
    names = declared_symbols()
    assert "sailor_integrate" in names and "sailor_octree_build" in names
    assert len(names) >= 15


def test_library_exports_every_declared_symbol():
    from sailor_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.fail("libsailor_hip.so is not built (run __graft_entry__.build())")
    L = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in sorted(declared_symbols()) if not hasattr(L, n)]
    assert not missing, f"not exported: {missing}"


def test_python_binding_covers_header():
    from sailor_amd import _lib
    missing = sorted(declared_symbols() - set(_lib._SIGS))
    assert not missing, f"no ctypes signature for {missing}"


def test_abi_version_and_error_string():
    from sailor_amd import _lib
    L = _lib.lib()
    assert L.sailor_abi_version() >= 1
    assert isinstance(L.sailor_last_error(), bytes)


def test_argument_errors_are_reported_not_fatal():
    """Shape errors come back as a status + message (no exit(), no GPU work)."""
    from sailor_amd import _lib
    with pytest.raises(_lib.SailorError, match="S must be"):
        _lib.call("sailor_sort_samplings", None, 4, 0, None, None, None)
    with pytest.raises(_lib.SailorError, match="not a multiple"):
        _lib.call("sailor_octree_build", None, None, None, None, 1, 30, 8, None, None, 0, 1,
                  None, None, None, None, 0, None)


def _header_params():
    """name -> list of parameter kinds ('P', 'I', 'F', 'SZ', 'LL') parsed from include/*.h."""
    out = {}
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if not fn.endswith(".h"):
            continue
        src = open(os.path.join(ROOT, "include", fn)).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        for m in re.finditer(r"\b(sailor_[a-z0-9_]+)\s*\(([^)]*)\)\s*;", src):
            kinds = []
            params = [p.strip() for p in m.group(2).split(",") if p.strip()]
            if params == ["void"]:
                params = []
            for p in params:
                if "*" in p:
                    kinds.append("P")
                elif p.startswith("size_t"):
                    kinds.append("SZ")
                elif p.startswith("long long"):
                    kinds.append("LL")
                elif p.startswith("float"):
                    kinds.append("F")
                elif p.startswith("int"):
                    kinds.append("I")
                else:
                    raise AssertionError(f"unclassified parameter {p!r} of {m.group(1)}")
            out[m.group(1)] = kinds
    return out


def test_ctypes_signatures_match_header():
    from sailor_amd import _lib
    kind = {_lib.P: "P", _lib.I: "I", _lib.F: "F", _lib.SZ: "SZ", _lib.LL: "LL"}
    hdr = _header_params()
    bad = []
    for name, params in hdr.items():
        got = [kind[a] for a in _lib._SIGS[name]]
        if got != params:
            bad.append(f"{name}: header {params} vs ctypes {got}")
    assert not bad, "\n".join(bad)
