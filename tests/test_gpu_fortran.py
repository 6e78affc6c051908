"""A Fortran caller of the library, GPU against the oracle.

ec2d_fdriver (tests/fortran/fdriver.f90, built by flang with the binding
module include/ec2d_bind.f90) is a Fortran main program that calls the
library as the reference's driver would: the control struct through the
bind(C) type, every geometry array allocated in the reference's bounds and
passed by c_loc, then ec2d_init_flow, the main loop (ec2d_ns_time_advance per
pass, or ec2d_run for all of them), ec2d_res_norms, ec2d_time and
ec2d_get_array(EC2D_U).  This test writes the host geometry of a case to the
driver's input file, runs it, and compares the U it wrote with the oracle's.
"""
import os
import subprocess

import numpy as np
import pytest

import oracle as O
from ec2d import build as B
from ec2d import cases
from ec2d import solver
from helpers import first_diff, naca_case, vortex_case

pytestmark = pytest.mark.gpu

DRIVER = os.path.join(B.FBUILD, "ec2d_fdriver")

CTL_D1 = ("Ma", "Re", "gamma", "AoA", "Pr", "Cv", "Cp")
CTL_D2 = ("t_end", "dt_global", "CFL", "dtmax", "dtmin", "P_outlet", "T_inf", "Twall", "vt_inf", "Kt_inf", "Wt_inf")
CTL_I = ("If_viscous", "Iflag_turbulence_model", "Iflag_init", "Iflag_local_dt", "Time_Method", "Iflag_Scheme",
         "Iflag_Flux", "IFlag_Reconstruction", "Num_Mesh", "Nvar")


def _write_input(path, case, nsteps, mode):
    k = case.c_control()
    U0 = case.read_restart() if case.ctl.Iflag_init == 1 else None
    with open(path, "wb") as f:
        blocks = case.levels[0]
        f.write(np.array([len(blocks), nsteps, mode], np.int32).tobytes())
        f.write(np.array([getattr(k, n) for n in CTL_D1 + CTL_D2], np.float64).tobytes())
        f.write(np.array([getattr(k, n) for n in CTL_I] + list(k.Pre_Step_Mesh) + [k.Nstep_Inner_Limit],
                         np.int32).tobytes())
        f.write(np.array([k.Res_Inner_Limit], np.float64).tobytes())
        for mb, g in enumerate(blocks):
            has_dw = g.dw is not None
            f.write(np.array([g.nx, g.ny, len(g.bc), 0, int(has_dw), int(U0 is not None)], np.int32).tobytes())
            f.write(np.array(g.bc, np.int32).ravel().tobytes())
            for n in ("x", "y", "x1", "y1", "vol", "si", "sj", "ni1", "ni2", "nj1", "nj2"):
                f.write(np.ascontiguousarray(getattr(g, n), np.float64).tobytes())
            if has_dw:
                f.write(np.ascontiguousarray(g.dw, np.float64).tobytes())
            if U0 is not None:
                f.write(np.ascontiguousarray(U0[mb], np.float64).tobytes())


def _run_driver(d, case, nsteps, mode, tmp_path):
    fin, fout = str(tmp_path / "in.bin"), str(tmp_path / "out.bin")
    _write_input(fin, case, nsteps, mode)
    print(f"ec2d_fdriver {d} steps {nsteps} mode {mode}", flush=True)
    r = subprocess.run([DRIVER, fin, fout], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and "fdriver OK" in r.stdout, r.stdout + r.stderr
    raw = np.fromfile(fout, dtype=np.uint8)
    out, off = [], 0
    nv = case.ctl.nvar
    for g in case.levels[0]:
        n = nv * (g.nx + 3) * (g.ny + 3)
        out.append(np.frombuffer(raw[off:off + 8 * n].tobytes(), np.float64).reshape(g.ny + 3, g.nx + 3, nv))
        off += 8 * n
    norms = np.frombuffer(raw[off:off + 96].tobytes(), np.float64)
    kstep = int(np.frombuffer(raw[off + 96:off + 100].tobytes(), np.int32)[0])
    return out, norms[:6], norms[6:], kstep


@pytest.mark.skipif(not os.path.exists(DRIVER), reason="Fortran driver not built (__graft_entry__.build())")
@pytest.mark.parametrize("kind,mode", [("euler_cgrid", 1), ("euler_multiblock", 2), ("sa_cgrid", 2)])
def test_fortran_caller_matches_oracle(kind, mode, tmp_path):
    nsteps = 3
    if kind == "euler_cgrid":
        d = naca_case("f_euler", cases.control_euler_naca(iters=10))
    elif kind == "euler_multiblock":
        # 6 blocks, vortex start read from flow2d.dat (put before ec2d_init_flow)
        d = vortex_case("f_mb", cases.control_vortex(iters=10))
    else:
        d = naca_case("f_sa", cases.control_sa_naca(iters=10, Re=1e5, Ma=0.5), ds1=1e-4)
    case = solver.Case.from_dir(d)
    U, rmax, rrms, kstep = _run_driver(d, case, nsteps, mode, tmp_path)
    ref = O.OracleCase(d)
    ref.init_flow()
    for _ in range(nsteps):
        ref.main_step()
    assert kstep == nsteps
    nv = case.ctl.nvar
    for mb, u in enumerate(U, start=1):
        if kind.startswith("euler"):
            diff = first_diff(u, ref.block(1, mb).U)
            assert diff is None, f"block {mb}: {diff}"
        else:
            np.testing.assert_allclose(u, ref.block(1, mb).U, rtol=1e-11, atol=1e-300)
    np.testing.assert_allclose(rrms[:nv], ref.norms(1)[1][:nv], rtol=1e-12 if kind.startswith("euler") else 1e-10)
