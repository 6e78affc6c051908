"""The LU-SGS-smoothed V-cycle (the extension the C2 workload runs) pinned
against the reference build's own subroutines, and by the FAS fixed point.

The reference stops on LU-SGS combined with multigrid
(opencfd_ec2d_v1.5.4.f90:216-221).  The extension keeps the structure of
NS_3stge_multigrid (sub_NS_multigrid.f90:66-142) and replaces each coarse
level's RK3 step (NS_Time_advance_RK3, sub_NS_singlegrid.f90:115-176, which
adds the forcing after each residual) by Set_Un followed by one LU-SGS step on
the forced residual:

    Comput_Residual_one_mesh(n); Add_force_function(n);
    du_LU_SGS_2D(n, every block, 0); U = U + dU; Boundary_condition_onemesh(n);
    update_buffer_onemesh(n)

and the fine level's step by NS_Time_advance(1) with LU-SGS (with the
nu~ / k / omega clip, force_vt_kw).  Here the whole cycle is driven through
the reference build's subroutines one call at a time (oracle/_ref, every call
a reference subroutine except the three-line update loop that the reference
keeps inline in NS_Time_advance_LU_SGS), and after every call the oracle's
arrays on that level must equal the reference's bit for bit; the oracle's
composite driver (OracleCase.main_step, the sequence the GPU runs) must then
equal the whole call sequence.

The FAS fixed point pins the composition independently of the reference: from
a fine state whose residual is zero to round-off, every coarse forcing cancels
the coarse residual of the restricted state, so one V-cycle changes U only at
round-off.  A forcing term with the wrong sign, Set_Un after the smoother, or
a correction prolonged from the wrong level moves U by far more.
"""
import numpy as np
import pytest

import oracle as O
import refpy
from ec2d import cases
from ec2d import formats as F
from helpers import first_diff, naca_case

needs_ref = pytest.mark.skipif(not refpy.available(), reason="reference build (oracle/_ref) not present")

LEVEL_ARRAYS = {1: ("U", "Res", "dU", "deltU"), 2: ("U", "Res", "dU", "Un", "QF", "deltU"),
                3: ("U", "Res", "dU", "Un", "QF", "deltU")}


def _vcycle_calls(oc, R, nm, nblock):
    """The extension's V-cycle as (label, level, oracle call, reference call)."""
    L = O.lib()
    M = oc.M

    def both(label, lev, fo, fr):
        return (label, lev, fo, fr)

    def lusgs_step(n):
        out = [both(f"residual({n})", n, lambda: oc.residual(n), lambda: R.op("residual", n))]
        if n > 1:
            out.append(both(f"add_force({n})", n, lambda: L.o_add_force(M(n)), lambda: R.op("add_force", n)))
        for mb in range(1, nblock + 1):
            out.append(both(f"lusgs({n},{mb})", n, lambda mb=mb: oc.lusgs(n, mb, 0.0),
                            lambda mb=mb: R.op("lusgs", n, mb, 0)))
        out.append(both(f"update({n})", n, lambda: oc.update(n),
                        lambda: [R.op("update", n, mb) for mb in range(1, nblock + 1)]))
        out.append(both(f"bc({n})", n, lambda: oc.bc(n), lambda: R.op("bc", n)))
        out.append(both(f"exchange({n})", n, lambda: oc.exchange(n), lambda: R.op("exchange", n)))
        return out

    def restrict(n):
        return [both(f"interpolation2h({n},{n + 1},2)", n + 1, lambda: L.o_interp2h(M(n), M(n + 1), 2),
                     lambda: R.op("restrict", n, n + 1, 2)),
                both(f"interpolation2h({n},{n + 1},1)", n + 1, lambda: L.o_interp2h(M(n), M(n + 1), 1),
                     lambda: R.op("restrict", n, n + 1, 1)),
                both(f"bc({n + 1})", n + 1, lambda: oc.bc(n + 1), lambda: R.op("bc", n + 1)),
                both(f"exchange({n + 1})", n + 1, lambda: oc.exchange(n + 1), lambda: R.op("exchange", n + 1)),
                both(f"residual({n + 1})", n + 1, lambda: oc.residual(n + 1), lambda: R.op("residual", n + 1)),
                both(f"comput_force_function({n + 1})", n + 1, lambda: L.o_force_function(M(n + 1)),
                     lambda: R.op("force", n + 1)),
                both(f"set_un({n + 1})", n + 1, lambda: L.o_set_un(M(n + 1)), lambda: R.op("set_un", n + 1))]

    def correct(n):
        return [both(f"comput_delt_u({n})", n, lambda: L.o_delt_u(M(n)), lambda: R.op("delt_u", n)),
                both(f"prolong_u({n},{n - 1},2)", n - 1, lambda: L.o_prolong(M(n), M(n - 1), 2),
                     lambda: R.op("prolong", n, n - 1, 2)),
                both(f"comput_new_u({n - 1})", n - 1, lambda: L.o_new_u(M(n - 1)), lambda: R.op("new_u", n - 1))]

    calls = lusgs_step(1)
    calls.append(both("force_vt_kw(1)", 1, lambda: oc.clip(1), lambda: R.op("clip", 1)))
    calls.append(both("residual(1)", 1, lambda: oc.residual(1), lambda: R.op("residual", 1)))
    calls += restrict(1) + lusgs_step(2)
    if nm >= 3:
        calls.append(both("residual(2)", 2, lambda: oc.residual(2), lambda: R.op("residual", 2)))
        calls.append(both("add_force(2)", 2, lambda: L.o_add_force(M(2)), lambda: R.op("add_force", 2)))
        calls += restrict(2) + lusgs_step(3) + correct(3)
    calls += correct(2)
    calls.append(both("bc(1)", 1, lambda: oc.bc(1), lambda: R.op("bc", 1)))
    calls.append(both("exchange(1)", 1, lambda: oc.exchange(1), lambda: R.op("exchange", 1)))
    return calls


def _mg_case(name, sa, nm, fdm=()):
    if sa:
        c = cases.control_sa_naca(iters=10, Re=1e5, Ma=0.5, num_mesh=nm, time_method=F.TIME_RK3)
    else:
        c = cases.control_euler_naca(iters=10)
        c.Num_Mesh, c.Time_Method = nm, F.TIME_RK3
    c.CFL = 2.0
    # the reference reads an RK3 V-cycle (it refuses LU-SGS with multigrid);
    # its Time_Method switches to LU-SGS after the set-up, as the oracle's
    return naca_case(name, c, ds1=1e-4 if sa else 1e-3, fdm=fdm)


@needs_ref
@pytest.mark.parametrize("sa,nm,fdm", [(False, 2, ()), (False, 3, ()), (True, 3, (1,))],
                         ids=["euler-2", "euler-3", "sa-fdweno5-3"])
def test_lusgs_vcycle_pieces_match_reference(sa, nm, fdm):
    d = _mg_case(f"r_vc_{int(sa)}_{nm}", sa, nm, fdm)
    R = refpy.RefCase(d)
    R.set_option(4, F.TIME_LU_SGS)
    oc = O.OracleCase(d)
    oc.init_flow()
    oc.ctl.Time_Method = F.TIME_LU_SGS
    oc.P.Time_Method = F.TIME_LU_SGS
    nblock = R.mesh_info(1)["nblock"]
    delt_set = {1}  # levels whose deltU holds values (the reference allocates coarse deltU unset)
    for cycle in range(2):
        for label, lev, fo, fr in _vcycle_calls(oc, R, nm, nblock):
            fo()
            fr()
            if label.startswith("comput_delt_u"):
                delt_set.add(lev)
            for mb in range(1, nblock + 1):
                for a in LEVEL_ARRAYS[lev]:
                    got = getattr(oc.block(lev, mb), a)
                    want = R.get(a, lev, mb)
                    if got is None or want is None or (a == "deltU" and lev not in delt_set):
                        continue
                    if a == "deltU" and lev > 1:
                        # coarse deltU is allocated without a value; comput_delt_U
                        # writes 0:nx x 0:ny (sub_NS_multigrid.f90:207-208)
                        got, want = got[1:-1, 1:-1], want[1:-1, 1:-1]
                    diff = first_diff(got, want)
                    assert diff is None, f"cycle {cycle}, after {label}: level {lev} block {mb} {a}: {diff}"
    # the oracle's composite driver performs the same call sequence
    oc2 = O.OracleCase(d)
    oc2.init_flow()
    oc2.ctl.Time_Method = F.TIME_LU_SGS
    oc2.P.Time_Method = F.TIME_LU_SGS
    for _ in range(2):
        oc2.main_step()
    for lev in range(1, nm + 1):
        for mb in range(1, nblock + 1):
            for a in LEVEL_ARRAYS[lev]:
                got, want = getattr(oc2.block(lev, mb), a), R.get(a, lev, mb)
                if got is None or want is None:
                    continue
                if a == "deltU" and lev > 1:
                    got, want = got[1:-1, 1:-1], want[1:-1, 1:-1]
                diff = first_diff(got, want)
                assert diff is None, f"composite driver, level {lev} block {mb} {a}: {diff}"


def _converged_fine_state(nm):
    """A 64 x 16-cell C-grid Euler case (first-order upwind, LU-SGS) run on
    its fine level alone until the density residual has dropped to round-off;
    returns the oracle case set up for the nm-level LU-SGS V-cycle on it."""
    c = cases.control_euler_naca(iters=10, cfl=20.0)
    c.Iflag_Scheme = F.SCHEME_UD1
    d1 = naca_case("o_fas_fine", c)
    single = O.OracleCase(d1)
    single.init_flow()
    r0 = None
    for k in range(6000):
        single.main_step()
        rr = single.norms(1)[1][0]
        r0 = rr if r0 is None else r0
        if rr < 1e-13 * r0:
            break
    assert rr < 1e-12 * r0, f"fine level did not converge: drop {rr / r0:.2e} after {k + 1} passes"
    c.Num_Mesh, c.Time_Method = nm, F.TIME_LU_SGS
    d = naca_case(f"o_fas_{nm}", c)
    mg = O.OracleCase(d)
    mg.init_flow()
    mg.block(1, 1).U[...] = single.block(1, 1).U
    return mg


@pytest.mark.parametrize("nm", [2, 3])
def test_fas_fixed_point(nm):
    mg = _converged_fine_state(nm)
    U0 = mg.block(1, 1).U.copy()
    mg.main_step()
    U1 = mg.block(1, 1).U
    scale = np.abs(U0[2:-2, 2:-2]).max(axis=(0, 1))
    change = np.abs(U1[2:-2, 2:-2] - U0[2:-2, 2:-2]).max(axis=(0, 1)) / scale
    assert change.max() < 1e-11, f"one V-cycle from a converged state moved U by {change}"
    # the coarse levels were exercised: their restricted states carry a
    # nonzero residual that the forcing cancels
    rc = np.abs(mg.block(2, 1).QF).max()
    assert rc > 1e-6, rc


def test_fas_fixed_point_detects_a_forcing_sign_error():
    """The fixed-point check has teeth: with Add_force_function subtracting
    the forcing instead of adding it (a mistake planted in the composition on
    the CPU), the same V-cycle moves U far from the converged state."""
    mg = _converged_fine_state(2)
    U0 = mg.block(1, 1).U.copy()
    L = O.lib()
    for b in mg.meshes[1]["blocks"]:
        b.QF[...] = 0.0
    # the V-cycle of OracleCase.mg3 with the sign of QF flipped after
    # comput_force_function
    mg.advance_lusgs(1)
    mg.clip(1)
    mg.residual(1)
    L.o_interp2h(mg.M(1), mg.M(2), 2)
    L.o_interp2h(mg.M(1), mg.M(2), 1)
    mg.bc(2)
    mg.exchange(2)
    mg.residual(2)
    L.o_force_function(mg.M(2))
    for b in mg.meshes[1]["blocks"]:
        b.QF[...] = -b.QF
    L.o_set_un(mg.M(2))
    mg.advance_lusgs(2, with_forcing=True)
    L.o_delt_u(mg.M(2))
    L.o_prolong(mg.M(2), mg.M(1), 2)
    L.o_new_u(mg.M(1))
    mg.bc(1)
    mg.exchange(1)
    U1 = mg.block(1, 1).U
    scale = np.abs(U0[2:-2, 2:-2]).max(axis=(0, 1))
    change = np.abs(U1[2:-2, 2:-2] - U0[2:-2, 2:-2]).max(axis=(0, 1)) / scale
    assert change.max() > 1e-6, change
