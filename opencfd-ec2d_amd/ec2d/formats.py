"""Readers and writers for the OpenCFD-EC2D case files.

The formats are the reference's own, so that a case directory can be run by
either the reference Fortran or this package:

* ``control.in``  six list-directed records, each preceded by a comment line
  (opencfd_ec2d_v1.5.4.f90:197-212).
* ``Mesh2d.dat``  block count, (NI, NJ) per block, then x(i,j) and y(i,j)
  per block with i fastest (sub_init.f90:122-127, 176).
* ``bc2d_1.in``   two comment lines, block count, then per block two comment
  lines, the subface count, a comment line and one 9-integer record per
  subface ``f_no face ist iend jst jend neighb subface orient``
  (sub_init.f90:228-249).
* ``flow2d.dat``  restart file: a title line, then per block a zone line and
  (nx+1)(ny+1) records ``x y d u v T ...`` over j=0..ny, i=0..nx
  (sub_init.f90:509-523).
* ``FDM.in``      optional list of blocks computed by finite differences
  (sub_Finite_Difference.f90:287-305).
"""
from __future__ import annotations

import os
from dataclasses import dataclass, field, fields
from typing import List, Sequence

import numpy as np

# Codes of the reference (opencfd_ec2d_v1.5.4.f90:42-47, 50).
SCHEME_UD1, SCHEME_NND2, SCHEME_UD3, SCHEME_WENO3, SCHEME_MUSCL2, SCHEME_MUSCL3, SCHEME_OMUSCL2 = range(7)
FLUX_SW, FLUX_HLL, FLUX_HLLC, FLUX_ROE, FLUX_VANLEER, FLUX_AUSM = 1, 2, 3, 4, 5, 6
RECON_ORIGINAL, RECON_CONSERVATIVE, RECON_CHARACTERISTIC = 0, 1, 2
BC_WALL, BC_FARFIELD, BC_SYMMETRY, BC_INLET, BC_OUTLET = -10, -20, -30, -40, -50
TIME_LU_SGS, TIME_EULER1, TIME_RK3, TIME_DUAL_LU_SGS = 0, 1, 3, -1
TURB_NONE, TURB_BL, TURB_SA, TURB_SST = 0, 1, 2, 3
# the learned closure standing in for SA's transport (no reference counterpart): 4 equations, mu_t from the MLP
TURB_CLOSURE = 4
LAP = 2
PI = 3.1415926535897932


def _fnum(tok: str) -> float:
    """Parse one Fortran list-directed real (``1.d-6``, ``10.``, ``2.31d0``)."""
    return float(tok.replace("d", "e").replace("D", "e"))


def _ftok(x: float) -> str:
    """Format a double so that a Fortran list-directed read recovers it exactly."""
    return repr(float(x)).replace("e", "E")


@dataclass
class Control:
    """The records of ``control.in`` (names follow opencfd_ec2d_v1.5.4.f90:203-211)."""

    Ma: float = 0.5
    Re: float = 1.0e6
    gamma: float = 1.4
    AoA: float = 0.0            # degrees, as written in the file
    Pr: float = 0.72
    t_end: float = 1.0
    Kstep_save: int = 1000000
    If_viscous: int = 0
    Iflag_turbulence_model: int = TURB_NONE
    Iflag_init: int = 0
    Iflag_local_dt: int = 1
    dt_global: float = 2.0 ** -10
    CFL: float = 5.0
    dtmax: float = 1.0
    dtmin: float = 1.0e-10
    Time_Method: int = TIME_LU_SGS
    P_outlet: float = -1.0
    T_inf: float = 288.15
    Twall: float = -1.0
    vt_inf: float = 3.0
    Kt_inf: float = 1.0e-8
    Wt_inf: float = 1.0e-6
    Iflag_Scheme: int = SCHEME_MUSCL3
    Iflag_Flux: int = FLUX_ROE
    IFlag_Reconstruction: int = RECON_ORIGINAL
    Kstep_show: int = 1000000
    Num_Mesh: int = 1
    Num_Threads: int = 1
    Nstep_Inner_Limit: int = 20
    Res_Inner_Limit: float = 1.0e-6
    Pre_Step_Mesh: List[int] = field(default_factory=lambda: [0, 0, 0])

    # -- derived quantities, as read_parameter computes them (:222-233) ----
    @property
    def nvar(self) -> int:
        if self.Iflag_turbulence_model == TURB_SST:
            return 6
        if self.Iflag_turbulence_model == TURB_SA:
            return 5
        return 4

    @property
    def aoa_rad(self) -> float:
        return self.AoA * PI / 180.0

    @property
    def Cv(self) -> float:
        return 1.0 / (self.gamma * (self.gamma - 1.0) * self.Ma * self.Ma)

    @property
    def Cp(self) -> float:
        return self.Cv * self.gamma

    @property
    def Twall_nd(self) -> float:
        return self.Twall / self.T_inf

    def iterations(self) -> int:
        """Outer iterations the reference main loop makes: tt += dt_global while tt < t_end."""
        tt, n = 0.0, 0
        while tt < self.t_end:
            tt = tt + self.dt_global
            n += 1
        return n


_RECORDS = [
    ("Ma Re gamma AoA Pr t_end Kstep_save If_viscous Iflag_turbulence_model Iflag_init",
     ["Ma", "Re", "gamma", "AoA", "Pr", "t_end", "Kstep_save", "If_viscous", "Iflag_turbulence_model", "Iflag_init"]),
    ("Iflag_local_dt dt_global CFL dtmax dtmin Time_Method P_outlet T_inf Twall vt_inf Kt_inf Wt_inf",
     ["Iflag_local_dt", "dt_global", "CFL", "dtmax", "dtmin", "Time_Method", "P_outlet", "T_inf", "Twall",
      "vt_inf", "Kt_inf", "Wt_inf"]),
    ("Iflag_Scheme Iflag_Flux IFlag_Reconstruction Kstep_show",
     ["Iflag_Scheme", "Iflag_Flux", "IFlag_Reconstruction", "Kstep_show"]),
    ("Num_Mesh Num_Threads Nstep_Inner_Limit Res_Inner_Limit",
     ["Num_Mesh", "Num_Threads", "Nstep_Inner_Limit", "Res_Inner_Limit"]),
]


def write_control(path: str, c: Control) -> None:
    types = {f.name: f.type for f in fields(Control)}
    lines = []
    for header, names in _RECORDS:
        lines.append(" " + header)
        vals = []
        for n in names:
            v = getattr(c, n)
            vals.append(str(int(v)) if types[n] in ("int", int) else _ftok(v))
        lines.append(" " + "  ".join(vals))
    lines.append(" Pre_Step_Mesh(1:Num_Mesh)")
    lines.append(" " + "  ".join(str(int(v)) for v in c.Pre_Step_Mesh[: c.Num_Mesh]))
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def read_control(path: str) -> Control:
    with open(path, encoding="latin-1") as f:
        raw = f.read().splitlines()
    c = Control()
    types = {f.name: f.type for f in fields(Control)}
    rec = 1  # record k lives on line 2k (a comment line precedes each one)
    for _, names in _RECORDS:
        toks = raw[rec].split()
        for n, t in zip(names, toks):
            v = _fnum(t)
            setattr(c, n, int(v) if types[n] in ("int", int) else v)
        rec += 2
    toks = raw[rec].split()
    pre = [int(_fnum(t)) for t in toks[: c.Num_Mesh]]
    c.Pre_Step_Mesh = (pre + [0, 0, 0])[:3]
    return c


def write_mesh2d(path: str, xs: Sequence[np.ndarray], ys: Sequence[np.ndarray]) -> None:
    """xs[m], ys[m] are (ny, nx) arrays of node coordinates (index [j-1, i-1])."""
    with open(path, "w") as f:
        f.write(f"{len(xs)}\n")
        f.write(" ".join(f"{x.shape[1]} {x.shape[0]}" for x in xs) + "\n")
        for x, y in zip(xs, ys):
            for arr in (x, y):
                flat = np.ascontiguousarray(arr).ravel()  # i fastest, then j
                for k in range(0, flat.size, 4):
                    f.write(" ".join(_ftok(v) for v in flat[k:k + 4]) + "\n")


def _tokens(path: str) -> List[str]:
    with open(path, encoding="latin-1") as f:
        return f.read().split()


def read_mesh2d(path: str):
    tok = _tokens(path)
    nb = int(tok[0])
    dims = [(int(tok[1 + 2 * m]), int(tok[2 + 2 * m])) for m in range(nb)]
    pos = 1 + 2 * nb
    xs, ys = [], []
    for nx, ny in dims:
        n = nx * ny
        x = np.array([_fnum(t) for t in tok[pos:pos + n]]).reshape(ny, nx)
        pos += n
        y = np.array([_fnum(t) for t in tok[pos:pos + n]]).reshape(ny, nx)
        pos += n
        xs.append(x)
        ys.append(y)
    return xs, ys


def write_bcin(path: str, bcs: Sequence[Sequence[Sequence[int]]], title: str = "ec2d case") -> None:
    """bcs[m] is the list of 9-integer subface records of block m+1."""
    out = [f" 2D BC file for multi-block mesh:  {title}", " # Blocks", f"  {len(bcs)}"]
    for m, subs in enumerate(bcs):
        out.append(f" Block  {m + 1}")
        out.append(" Subface")
        out.append(f"  {len(subs)}")
        out.append(" f_no, face, istart,iend, jstart, jend,  neighb, subface orient")
        for s in subs:
            out.append(" " + " ".join(f"{int(v):6d}" for v in s))
    with open(path, "w") as f:
        f.write("\n".join(out) + "\n")


def read_bcin(path: str) -> List[List[List[int]]]:
    with open(path, encoding="latin-1") as f:
        lines = f.read().splitlines()
    nb = int(lines[2].split()[0])
    pos = 3
    bcs = []
    for _ in range(nb):
        pos += 2
        nsub = int(lines[pos].split()[0])
        pos += 2
        subs = []
        for _ in range(nsub):
            subs.append([int(v) for v in lines[pos].split()[:9]])
            pos += 1
        bcs.append(subs)
    return bcs


def write_fdm_in(path: str, blocks: Sequence[int]) -> None:
    lines = [" Blocks using the finite-difference method", " (hybrid FVM/FDM)", " ---", " NB_FDM",
             f" {len(blocks)}", " block numbers"] + [f" {b}" for b in blocks]
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def read_fdm_in(path: str) -> List[int]:
    if not os.path.exists(path):
        return []
    with open(path, encoding="latin-1") as f:
        lines = f.read().splitlines()
    nb = int(lines[4].split()[0])
    return [int(lines[6 + k].split()[0]) for k in range(nb)]


def write_flow2d(path: str, x1s, y1s, prims) -> None:
    """Restart file read by init_flow_read (sub_init.f90:509-523).

    prims[m] is an array (ny+1, nx+1, 4) of (d, u, v, T) over j=0..ny, i=0..nx
    (the interior plus one ghost ring), x1s/y1s the matching cell centres.
    """
    with open(path, "w") as f:
        f.write("variables=x,y,d,u,v,T,p\n")
        for x1, y1, q in zip(x1s, y1s, prims):
            nyp, nxp = q.shape[:2]
            f.write(f" zone  i= {nxp} j= {nyp}\n")
            for j in range(nyp):
                for i in range(nxp):
                    f.write(" ".join(_ftok(v) for v in (x1[j, i], y1[j, i], *q[j, i, :4])) + "\n")


# ---- the reference's result files (output / output_Res, opencfd_ec2d_v1.5.4.f90:266-371) ----

def fortran_f(x: float, w: int = 20, d: int = 10) -> str:
    """Fortran Fw.d editing: correctly rounded, right-justified, asterisks when
    the value does not fit."""
    t = f"{float(x):{w}.{d}f}"
    return "*" * w if len(t) > w else t


def fortran_e(x: float, w: int, d: int = 10) -> str:
    """Fortran Ew.d editing: 0.dddd...E+ee with d correctly rounded digits
    (a three-digit exponent drops the E, as the standard prescribes)."""
    x = float(x)
    if x == 0.0 or x != x or x in (float("inf"), float("-inf")):
        if x != x or x in (float("inf"), float("-inf")):
            return f"{'NaN' if x != x else ('-Infinity' if x < 0 else 'Infinity'):>{w}}"
        mant, e10 = "0" * d, 0
    else:
        m, e = f"{abs(x):.{d - 1}e}".split("e")
        mant, e10 = m.replace(".", ""), int(e) + 1
    sign = "-" if (x < 0 or (x == 0.0 and str(x).startswith("-"))) else ""
    ex = f"E{'-' if e10 < 0 else '+'}{abs(e10):02d}" if abs(e10) <= 99 else f"{'-' if e10 < 0 else '+'}{abs(e10):03d}"
    t = f"{sign}0.{mant}{ex}"
    return "*" * w if len(t) > w else f"{t:>{w}}"


def write_output(dirpath: str, nmesh: int, kstep: int, blocks, nvar: int, Cv: float, gamma: float, Ma: float,
                 Re: float) -> List[str]:
    """Result files of `output(nMesh)` (opencfd_ec2d_v1.5.4.f90:286-371), byte
    for byte: flow2d.dat (nMesh 1; flow2d-<nMesh>.dat otherwise) with
    x, y, d, u, v, T, p, Amut*Re; for nMesh 1 also flow2d-<Kstep>.dat with
    x, y, d, u, v, T, p; SA2d.dat (5 equations) or SST2d.dat (6).  blocks: per
    block a dict of U (ny+3, nx+3, nvar), Amu_t (ny+3, nx+3), x1, y1
    (ny+3, nx+3), all over -1..n+1; rows j = 0..ny, i = 0..nx are written.
    Returns the paths written."""
    def prims(b):
        U = b["U"][1:-1, 1:-1]
        d1 = U[..., 0]
        u1 = U[..., 1] / d1
        v1 = U[..., 2] / d1
        T1 = (U[..., 3] - 0.5 * d1 * (u1 * u1 + v1 * v1)) / (Cv * d1)
        return d1, u1, v1, T1, d1 * T1 / (gamma * Ma * Ma)

    def zone(b):
        ny1, nx1 = b["U"].shape[0] - 2, b["U"].shape[1] - 2
        return f"  zone i=  {nx1}  j=  {ny1}\n"

    def write(path, header, cols_of):
        with open(path, "w") as f:
            f.write(f" {header}\n")
            for b in blocks:
                f.write(zone(b))
                cols = [np.asarray(c).ravel() for c in cols_of(b)]
                for row in zip(*cols):
                    f.write("".join(fortran_f(v) for v in row) + "\n")
        return path

    out = []
    xy = lambda b: (b["x1"][1:-1, 1:-1], b["y1"][1:-1, 1:-1])  # noqa: E731
    if nmesh == 1:
        out.append(write(os.path.join(dirpath, f"flow2d-{kstep:07d}.dat"), "variables=x,y,d,u,v,T,p",
                         lambda b: (*xy(b), *prims(b))))
    name = "flow2d.dat" if nmesh == 1 else f"flow2d-{nmesh:1d}.dat"
    out.append(write(os.path.join(dirpath, name), "variables=x,y,d,u,v,T,p,Amut",
                     lambda b: (*xy(b), *prims(b), b["Amu_t"][1:-1, 1:-1] * Re)))
    if nvar == 5:
        out.append(write(os.path.join(dirpath, "SA2d.dat"), "variables=x,y,vt",
                         lambda b: (*xy(b), b["U"][1:-1, 1:-1, 4])))
    if nvar == 6:
        out.append(write(os.path.join(dirpath, "SST2d.dat"), "variables=x,y,Kt,Wt",
                         lambda b: (*xy(b), b["U"][1:-1, 1:-1, 4], b["U"][1:-1, 1:-1, 5])))
    return out


def append_residual(dirpath: str, kstep: int, res_max, res_rms, nvar: int) -> str:
    """One line of Residual.dat as output_Res writes it (opencfd_ec2d_v1.5.4.f90:266-281):
    (I8,13E24.10) of Kstep, Res_max(1:6), Res_rms(1:NVAR)."""
    vals = list(res_max[:6]) + list(res_rms[:nvar])
    path = os.path.join(dirpath, "Residual.dat")
    with open(path, "a") as f:
        f.write(f"{kstep:8d}" + "".join(fortran_e(v, 24) for v in vals) + "\n")
    return path


def _list_digits(x: float):
    """(digits, e) with x = 0.<digits> * 10**e for a positive finite double,
    as flang's list-directed output picks them: the fewest significant digits
    n for which n-digit decimals lie strictly between x and its neighbours'
    midpoints, and of those the middle one (the lower of two), which is not
    always the one nearest to x (1.6954791700008828 for the double nearest
    1.69547917000088288475)."""
    import decimal
    import math

    with decimal.localcontext() as ctx:
        ctx.prec = 1200  # a double's exact decimal expansion has at most 767 digits
        D = decimal.Decimal
        ex = D(x)
        up = D(math.ulp(x))
        # below a power of two the spacing halves (normal range only)
        dn = up / 2 if (math.frexp(x)[0] == 0.5 and x >= 2.0 ** -1021) else up
        lo, hi = ex - dn / 2, ex + up / 2
        adj = ex.adjusted()
        for n in range(1, 18):
            q = D(1).scaleb(adj - n + 1)
            a, b = lo / q, hi / q
            cmin = a.to_integral_value(decimal.ROUND_CEILING)
            cmax = b.to_integral_value(decimal.ROUND_FLOOR)
            if cmin == a:
                cmin += 1
            if cmax == b:
                cmax -= 1
            if cmin <= cmax:
                t = (((cmin + cmax) // 2) * q).normalize().as_tuple()
                digits = "".join(map(str, t.digits))
                return digits, len(digits) + t.exponent
    raise ValueError(x)


def fortran_list_real(x: float) -> str:
    """One REAL(8) item of list-directed output as flang's runtime edits it.

    The digits are those of `_list_digits`.  The form is F when x, rounded to
    one significant digit, lies in [0.1, 1e15): no leading zero (".5"), a
    trailing point on whole numbers ("12.", "0.").  Otherwise it is 1P E with
    a two-digit minimum exponent ("1.E-05", "2.5E+300").  Zero keeps its sign
    ("-0.")."""
    import decimal

    x = float(x)
    if x != x:
        return "NaN"
    sign = "-" if np.signbit(x) else ""
    if x in (float("inf"), float("-inf")):
        return sign + "Inf"
    if x == 0.0:
        return sign + "0."
    digits, e = _list_digits(abs(x))
    with decimal.localcontext() as ctx:
        ctx.prec = 1200
        ax = decimal.Decimal(abs(x))
        r = ax.quantize(decimal.Decimal(1).scaleb(ax.adjusted()), rounding=decimal.ROUND_HALF_EVEN)
        f_form = decimal.Decimal("0.1") <= r < decimal.Decimal("1e15")
    if f_form:
        if e <= 0:
            return sign + "." + "0" * (-e) + digits
        return sign + digits[:e].ljust(e, "0") + "." + digits[e:]
    return sign + digits[0] + "." + digits[1:] + "E" + ("-" if e - 1 < 0 else "+") + f"{abs(e - 1):02d}"


def fortran_list_line(*items) -> str:
    """A list-directed record: a leading blank, items separated by one blank;
    strings as they are, integers in minimal width, reals as fortran_list_real."""
    out = []
    for v in items:
        if isinstance(v, str):
            out.append(v)
        elif isinstance(v, (int, np.integer)):
            out.append(str(int(v)))
        else:
            out.append(fortran_list_real(v))
    return " " + " ".join(out)


def write_wall_cp(dirpath: str, case_dir: str = None) -> str:
    """Cp.dat of the reference's wall-pressure post-processor (read_flow-1.2.f90:24-53):
    Ma from control.in, the blocks' sizes from Mesh2d.dat, the subfaces from
    bc2d_1.in (all in case_dir, default dirpath) and the flow from dirpath's
    flow2d.dat (the F20.10 file `output` writes).  For every wall subface
    (bc -10) a zone of x1(i,j) and -(p(i,j)-p00)*2 over its node-index range
    ibegin..iend, jbegin..jend, j outer; p00 = 1/(1.4 Ma^2) (gamma fixed at 1.4
    in the tool).  The cell-centred values are read at node indices, as the
    tool does."""
    case_dir = case_dir or dirpath
    with open(os.path.join(case_dir, "control.in"), encoding="latin-1") as f:
        f.readline()
        ma = _fnum(f.readline().split()[0])
    p00 = 1.0 / (1.4 * ma * ma)
    tok = _tokens(os.path.join(case_dir, "Mesh2d.dat"))
    dims = [(int(tok[1 + 2 * m]), int(tok[2 + 2 * m])) for m in range(int(tok[0]))]
    bcs = read_bcin(os.path.join(case_dir, "bc2d_1.in"))
    lines = [fortran_list_line("variable=x,Cp")]
    with open(os.path.join(dirpath, "flow2d.dat")) as f:
        f.readline()
        flows = []
        for nx, ny in dims:
            f.readline()
            # the tool reads the first 7 items of each record (x1, y1, d, u, v, T, p)
            q = np.array([[_fnum(t) for t in f.readline().split()[:7]] for _ in range((nx + 1) * (ny + 1))])
            flows.append(q.reshape(ny + 1, nx + 1, 7))
    for subs, q in zip(bcs, flows):
        for sb in subs:
            if sb[6] != BC_WALL:
                continue
            ib, ie, jb, je = sb[2], sb[3], sb[4], sb[5]
            lines.append(fortran_list_line("zone i=", (je - jb + 1) * (ie - ib + 1)))
            for j in range(jb, je + 1):
                for i in range(ib, ie + 1):
                    lines.append(fortran_list_line(q[j, i, 0], -(q[j, i, 6] - p00) * 2.0))
    path = os.path.join(dirpath, "Cp.dat")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")
    return path


def fortran_pow_int(x, n: int):
    """x**n for a constant integer n as flang evaluates it: ((x*x)*x)*... ."""
    r = x
    for _ in range(n - 1):
        r = r * x
    return r
