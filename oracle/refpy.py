"""ctypes driver for oracle/_ref/libec2d_ref.so (the flang build of the
unmodified reference plus ref_harness.f90).  Test infrastructure only.

The reference keeps its whole state in module globals, so one process holds at
most one live reference case; ``RefCase(dir)`` re-runs the reference set-up in
``dir`` (it chdir()s there for the duration of each call that reads or writes
files, because the reference opens its files by relative name).
"""
from __future__ import annotations

import contextlib
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_ref", "libec2d_ref.so")

# array selectors of ref_harness.f90:pick
ARRAYS = {
    "U": 1, "Res": 2, "dt": 3, "dU": 4, "Lci": 5, "Lcj": 6, "Lvi": 7, "Lvj": 8, "Amu": 9, "Amu_t": 10,
    "vol": 11, "si": 12, "sj": 13, "ni1": 14, "ni2": 15, "nj1": 16, "nj2": 17, "x": 18, "y": 19,
    "x1": 20, "y1": 21, "dw": 22, "Un": 23, "QF": 24, "deltU": 25,
    "Akx": 26, "Aky": 27, "Aix": 28, "Aiy": 29, "Ajac": 30, "Un1": 31,
}
OPS = {
    "residual": 1, "lusgs": 2, "update": 3, "bc": 4, "exchange": 5, "clip": 6, "advance": 7,
    "mg2": 8, "mg3": 9, "euler1": 10, "rk3": 11, "restrict": 12, "prolong": 13, "force": 14,
    "add_force": 15, "delt_u": 16, "set_un": 17, "new_u": 18, "advance_lusgs": 19, "main_step": 20,
    "main_steps": 21, "output": 22, "output_res": 23,
}

_lib = None


def available() -> bool:
    return os.path.exists(LIB_PATH)


def lib():
    global _lib
    if _lib is None:
        _lib = C.CDLL(LIB_PATH)
        i = C.c_int
        _lib.ref_setup.argtypes = []
        _lib.ref_num_mesh.restype = i
        _lib.ref_mesh_info.argtypes = [i, C.POINTER(i), C.POINTER(i), C.POINTER(i), C.POINTER(i)]
        _lib.ref_block_info.argtypes = [i, i, C.POINTER(i), C.POINTER(i), C.POINTER(i), C.POINTER(i)]
        _lib.ref_bc_msg.argtypes = [i, i, i, C.POINTER(i)]
        _lib.ref_size.argtypes = [i, i, i]
        _lib.ref_size.restype = i
        _lib.ref_bounds.argtypes = [i, i, i, C.POINTER(i), C.POINTER(i), C.POINTER(i)]
        _lib.ref_get.argtypes = [i, i, i, C.c_void_p, i]
        _lib.ref_set.argtypes = [i, i, i, C.c_void_p, i]
        _lib.ref_res_norms.argtypes = [i, C.c_void_p, C.c_void_p]
        _lib.ref_params.argtypes = [C.c_void_p, C.c_void_p]
        _lib.ref_set_option.argtypes = [i, i, i]
        _lib.ref_set_real.argtypes = [i, C.c_double]
        _lib.ref_op.argtypes = [i, i, i, i]
    return _lib


@contextlib.contextmanager
def _cwd(path):
    old = os.getcwd()
    os.chdir(path)
    try:
        yield
    finally:
        os.chdir(old)


class RefCase:
    def __init__(self, case_dir: str):
        self.dir = os.path.abspath(case_dir)
        L = lib()
        stale = os.path.join(self.dir, "dist_wall.dat")
        if os.path.exists(stale):
            os.remove(stale)
        with _cwd(self.dir):
            L.ref_setup()
        self.num_mesh = L.ref_num_mesh()

    # ---- structure ----------------------------------------------------
    def mesh_info(self, nmesh=1):
        a, b, c = C.c_int(), C.c_int(), C.c_int()
        info = (C.c_int * 4)()
        lib().ref_mesh_info(nmesh, C.byref(a), C.byref(b), C.byref(c), info)
        return {"nblock": a.value, "nvar": b.value, "kstep": c.value, "scheme": info[0], "flux": info[1],
                "recon": info[2], "turb": info[3]}

    def block_info(self, nmesh, mblock):
        a, b, c, d = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        lib().ref_block_info(nmesh, mblock, C.byref(a), C.byref(b), C.byref(c), C.byref(d))
        return a.value, b.value, c.value, d.value

    def bc(self, nmesh, mblock):
        nsub = self.block_info(nmesh, mblock)[2]
        out = []
        for k in range(1, nsub + 1):
            v = (C.c_int * 9)()
            lib().ref_bc_msg(nmesh, mblock, k, v)
            out.append(list(v))
        return out

    def params(self):
        r = np.zeros(20)
        iv = np.zeros(20, dtype=np.int32)
        lib().ref_params(r.ctypes.data, iv.ctypes.data)
        return r, iv

    # ---- arrays -------------------------------------------------------
    def get(self, name, nmesh=1, mblock=1):
        """Array `name` of block mblock as numpy, C-ordered reversed dims
        (j, i) or (j, i, m), i.e. the Fortran memory order."""
        L = lib()
        w = ARRAYS[name]
        n = L.ref_size(nmesh, mblock, w)
        if n == 0:
            return None
        lb = (C.c_int * 3)()
        ub = (C.c_int * 3)()
        rk = C.c_int()
        L.ref_bounds(nmesh, mblock, w, lb, ub, C.byref(rk))
        shape = [ub[k] - lb[k] + 1 for k in range(rk.value)]
        buf = np.empty(n)
        L.ref_get(nmesh, mblock, w, buf.ctypes.data, n)
        return buf.reshape(shape[::-1])

    def set(self, name, arr, nmesh=1, mblock=1):
        L = lib()
        w = ARRAYS[name]
        n = L.ref_size(nmesh, mblock, w)
        a = np.ascontiguousarray(arr, dtype=np.float64).ravel()
        assert a.size == n, (name, a.size, n)
        L.ref_set(nmesh, mblock, w, a.ctypes.data, n)

    def norms(self, nmesh=1):
        rmax = np.zeros(6)
        rrms = np.zeros(6)
        lib().ref_res_norms(nmesh, rmax.ctypes.data, rrms.ctypes.data)
        return rmax, rrms

    # ---- calls --------------------------------------------------------
    def op(self, name, a=1, b=1, c=0):
        with _cwd(self.dir):
            lib().ref_op(OPS[name], a, b, c)

    def set_option(self, which, val, nmesh=1):
        lib().ref_set_option(nmesh, which, val)

    def set_real(self, which, val):
        lib().ref_set_real(which, float(val))
