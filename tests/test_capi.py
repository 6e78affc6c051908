"""The C-ABI library: it loads without a GPU and exports exactly what
include/ec2d.h declares (CPU; no compute call is made)."""
import ctypes as C
import os
import re

import pytest

from ec2d import capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "ec2d.h")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return re.findall(r"^\s*(?:int|long long)\s+(ec2d_\w+)\s*\(", text, flags=re.M)


def test_header_declares_entry_points():
    names = declared()
    assert len(names) >= 30
    assert len(names) == len(set(names))


@pytest.mark.skipif(not os.path.exists(capi.LIB_PATH), reason="HIP library not built")
def test_library_exports_every_declared_symbol():
    lib = C.CDLL(capi.LIB_PATH)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_header():
    assert sorted(declared()) == sorted(capi.SIGNATURES)


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    monkeypatch.setattr(capi, "LIB_PATH", str(tmp_path / "absent.so"))
    monkeypatch.setattr(capi, "_lib", None)
    with pytest.raises(capi.Ec2dError):
        capi.load()
