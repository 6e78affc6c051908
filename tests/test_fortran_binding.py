"""The Fortran binding (include/ec2d_bind.f90) against the C-ABI (CPU).

ec2d_bind is what the reference's Fortran driver `use`s to call the library.
Its ec2d_control must have the C struct's size and every member's offset, and
it must declare a bind(C) interface, with as many arguments, for every entry
point of include/ec2d.h.  flang builds the module and a probe library
(__graft_entry__.build()); without flang's output the layout test skips.
"""
import ctypes as C
import os
import re

import pytest

from ec2d import build as B
from ec2d import capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PROBE = os.path.join(B.FBUILD, "libec2d_layout.so")


@pytest.mark.skipif(not os.path.exists(PROBE), reason="Fortran binding not built (__graft_entry__.build())")
def test_control_layout_matches_c():
    L = C.CDLL(PROBE)
    nf, nc = C.c_int(), C.c_int()
    of, oc = (C.c_int * 32)(), (C.c_int * 32)()
    L.fprobe_layout(C.byref(nf), of)
    L.cprobe_layout(C.byref(nc), oc)
    assert nf.value == nc.value == C.sizeof(capi.Control), (nf.value, nc.value, C.sizeof(capi.Control))
    assert list(of) == list(oc), (list(of), list(oc))
    # and the ctypes mirror the Python host passes
    names = [f[0] for f in capi.Control._fields_]
    assert [getattr(capi.Control, n).offset for n in names] == list(oc)[:len(names)]
    # the last double after 18 doubles and 14 ints: 8-byte aligned at 200
    assert oc[30] == 200 and nc.value == 208


def _c_functions():
    text = open(os.path.join(ROOT, "include", "ec2d.h")).read()
    text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
    out = {}
    for m in re.finditer(r"\b(?:int|long long)\s+(ec2d_\w+)\s*\(([^)]*)\)\s*;", text):
        args = m.group(2).strip()
        out[m.group(1)] = 0 if args in ("", "void") else args.count(",") + 1
    return out


def _f_interfaces():
    text = open(os.path.join(ROOT, "include", "ec2d_bind.f90")).read()
    text = re.sub(r"&\s*\n\s*", " ", text)
    out = {}
    for m in re.finditer(r"function\s+(ec2d_\w+)\s*\(([^)]*)\)\s*bind\(C,\s*name=\"(\w+)\"\)", text, flags=re.I):
        assert m.group(1) == m.group(3), m.group(0)
        args = m.group(2).strip()
        out[m.group(1)] = 0 if not args else args.count(",") + 1
    return out


def test_every_entry_point_has_a_fortran_interface():
    c, f = _c_functions(), _f_interfaces()
    assert len(c) >= 50
    missing = sorted(set(c) - set(f))
    assert not missing, missing
    extra = sorted(set(f) - set(c))
    assert not extra, extra
    wrong = {k: (c[k], f[k]) for k in c if c[k] != f[k]}
    assert not wrong, wrong
