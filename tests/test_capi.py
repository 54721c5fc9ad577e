"""The C ABI (include/blond_mi355x.h) on the CPU: the library loads, exports
every declared entry point with a matching ctypes binding, and rejects bad
arguments with an error code and message before any device work."""
import ctypes as ct
import os
import re

import pytest

from blond_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "blond_mi355x.h")


def declared():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(blmi_\w+)\s*\(", src)))


def test_library_loads():
    lib = _lib.load()
    assert lib.blmi_version() >= 100


def test_every_declared_symbol_is_exported_and_bound():
    lib = _lib.load()
    names = declared()
    assert len(names) >= 30
    for n in names:
        assert hasattr(lib, n), f"{n} declared but not exported"
        assert n in _lib.SIGNATURES, f"{n} has no ctypes signature"
    extra = set(_lib.SIGNATURES) - set(names)
    assert not extra, f"bound but not declared: {extra}"


def test_header_arity_matches_bindings():
    src = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    for name, args in _lib.SIGNATURES.items():
        m = re.search(r"\b" + name + r"\s*\(([^)]*)\)", src)
        assert m, name
        params = [p for p in m.group(1).split(",") if p.strip() and p.strip() != "void"]
        assert len(params) == len(args), f"{name}: header {len(params)} vs binding {len(args)}"


def test_argument_errors_without_gpu():
    lib = _lib.load()
    rc = lib.blmi_kick(None, None, 1, 1.0, None, None, None, -5, 0.0, None)
    assert rc == 1 and b"negative" in lib.blmi_last_error()
    rc = lib.blmi_drift(None, None, 7, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 10, None)
    assert rc == 1 and b"solver" in lib.blmi_last_error()
    rc = lib.blmi_histogram(None, None, 1.0, 0.0, 10, 0, None)
    assert rc == 1 and b"null profile" in lib.blmi_last_error()
    buf = (ct.c_double * 10)()
    rc = lib.blmi_histogram(None, buf, 1.0, 0.0, 10, 0, None)
    assert rc == 1 and b"cut_right" in lib.blmi_last_error()
    rc = lib.blmi_linear_interp_kick(None, None, None, None, 1.0, 1, 0, 0.0, None)
    assert rc == 1 and b"2 slices" in lib.blmi_last_error()
    assert lib.blmi_engine_create(10, 1, 0, 0) is None
    assert b"engine_create" in lib.blmi_last_error()
    assert lib.blmi_engine_track(None, 1, 1, None) == 1


def test_missing_library_fails_loudly(tmp_path):
    with pytest.raises(_lib.LibraryNotAvailable):
        _lib.load(str(tmp_path / "nope.so"))
