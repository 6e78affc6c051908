"""Build of the HIP library libec2d_hip.so (gfx950) from csrc/, and of the
Fortran binding (include/ec2d_bind.f90) with what tests drive through it.

Flags: -ffp-contract=off keeps every multiply and add separately rounded, as
the reference Fortran evaluates them; packed FP32 instructions are disabled.
"""
from __future__ import annotations

import os
import subprocess
import sys

PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "libec2d_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = ["--offload-arch=gfx950", "-O3", "-ffp-contract=off", "-fPIC", "-shared", "-std=c++17",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def build_lib(verbose: bool = False) -> str:
    tmp = LIB + ".tmp"
    cmd = [HIPCC, *FLAGS, "-o", tmp, os.path.join(CSRC, "ec2d_lib.hip")]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed for libec2d_hip.so")
    if verbose:
        sys.stderr.write(r.stderr)
    os.replace(tmp, LIB)
    return LIB


ROOT = os.path.dirname(PKG)
FBUILD = os.path.join(PKG, "build", "fortran")
FLANG = os.environ.get("FLANG", "/opt/rocm/llvm/bin/flang")


def _run(cmd, what):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"{what} failed")


def build_fortran() -> str:
    """The binding module (ec2d_bind.mod / .o) with flang, and the two test
    programs that use it: tests/fortran/layout_probe (the Fortran and C
    layouts of ec2d_control, libec2d_layout.so) and tests/fortran/fdriver.f90
    (a Fortran caller of the library, ec2d_fdriver).  Needs libec2d_hip.so.
    Returns the build folder."""
    os.makedirs(FBUILD, exist_ok=True)
    inc = os.path.join(ROOT, "include")
    tf = os.path.join(ROOT, "tests", "fortran")
    bind_o = os.path.join(FBUILD, "ec2d_bind.o")
    _run([FLANG, "-O2", "-fPIC", "-module-dir", FBUILD, "-c", os.path.join(inc, "ec2d_bind.f90"), "-o", bind_o],
         "flang (ec2d_bind.f90)")
    probe_o = os.path.join(FBUILD, "layout_probe.o")
    _run([FLANG, "-O2", "-fPIC", "-module-dir", FBUILD, "-I", FBUILD, "-c", os.path.join(tf, "layout_probe.f90"),
          "-o", probe_o], "flang (layout_probe.f90)")
    probe_c = os.path.join(FBUILD, "layout_probe_c.o")
    _run(["gcc", "-O2", "-fPIC", "-c", os.path.join(tf, "layout_probe.c"), "-o", probe_c], "gcc (layout_probe.c)")
    _run([FLANG, "-shared", "-o", os.path.join(FBUILD, "libec2d_layout.so"), probe_o, probe_c],
         "flang (libec2d_layout.so)")
    drv_o = os.path.join(FBUILD, "fdriver.o")
    _run([FLANG, "-O2", "-module-dir", FBUILD, "-I", FBUILD, "-c", os.path.join(tf, "fdriver.f90"), "-o", drv_o],
         "flang (fdriver.f90)")
    _run([FLANG, "-o", os.path.join(FBUILD, "ec2d_fdriver"), drv_o, bind_o, "-L" + PKG, "-lec2d_hip",
          "-Wl,-rpath," + PKG], "flang (ec2d_fdriver)")
    return FBUILD


if __name__ == "__main__":
    print(build_lib(verbose=True))
    print(build_fortran())
