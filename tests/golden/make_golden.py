"""Write the golden fixtures of tests/test_golden.py from the reference build.

    python tests/golden/make_golden.py [fixture names]

Runs the flang build of the unmodified reference (oracle/_ref, built by
oracle/Makefile.ref from the reference sources) on small generated cases and
stores, per case, the input files and the reference's state after a few
main-loop passes in full FP64 precision.  The reference itself writes its
outputs with 10 significant digits only, so the harness (oracle/ref_harness.f90)
reads its arrays directly.
"""
import hashlib
import os
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.join(ROOT, "opencfd-ec2d_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "tests"))

from ec2d import cases  # noqa: E402
from ec2d import formats as F  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))
FILES = ("control.in", "Mesh2d.dat", "bc2d_1.in", "FDM.in", "flow2d.dat")


def specs():
    """(name, builder, steps, arrays); builder(dir) writes the case."""
    def naca(ctl, nx=65, ny=17, ds1=1e-3, fdm=()):
        def b(d):
            x, y, bc, _ = cases.naca_cgrid(nx, ny, ds1=ds1)
            cases.write_case(d, ctl, [x], [y], [bc], fdm_blocks=fdm)
        return b

    def vortex(ctl):
        def b(d):
            from ec2d import geometry as G
            xs, ys, bcs, field, _ = cases.vortex_multiblock(seed=3, nbx=3, nby=2, wmin=9, wmax=21, hmin=9, hmax=17,
                                                           h=0.25)
            lev = G.build_levels(xs, ys, bcs, 1)
            flow = ([g.x1[1:-1, 1:-1] for g in lev[0]], [g.y1[1:-1, 1:-1] for g in lev[0]],
                    [field(g.x1[1:-1, 1:-1], g.y1[1:-1, 1:-1]) for g in lev[0]])
            cases.write_case(d, ctl, xs, ys, bcs, flow=flow)
        return b

    out = []
    out.append(("euler_roe_muscl3_lusgs", naca(cases.control_euler_naca(iters=10)), 3, ("U", "Res", "dt")))
    c = cases.control_sa_naca(iters=10, Re=1e5, Ma=0.5)
    out.append(("sa_lusgs", naca(c, ds1=1e-4), 3, ("U", "Res", "dt", "Amu_t")))
    c = cases.control_sa_naca(iters=10, Re=1e5, Ma=0.5)
    c.Iflag_turbulence_model, c.Twall = F.TURB_SST, 300.0
    out.append(("sst_isothermal_lusgs", naca(c, ds1=1e-4), 2, ("U", "Res", "dt", "Amu_t")))
    out.append(("vortex_multiblock_ausm", vortex(cases.control_vortex(iters=10)), 3, ("U", "Res", "dt")))
    c = cases.control_euler_naca(iters=10)
    c.Time_Method, c.Num_Mesh, c.CFL = F.TIME_RK3, 3, 0.3
    out.append(("euler_rk3_mg3", naca(c), 2, ("U", "Res")))
    c = cases.control_sa_naca(iters=10, Re=1e5, Ma=0.5)
    out.append(("sa_fdm_weno5", naca(c, ds1=1e-4, fdm=(1,)), 2, ("U", "Res", "dt")))
    c = cases.control_sa_naca(iters=10, Re=1e5, Ma=0.5)
    c.Iflag_turbulence_model = F.TURB_BL
    out.append(("baldwin_lomax_lusgs", naca(c, ds1=1e-4), 3, ("U", "Res", "dt", "Amu_t")))
    out.append(("rae2822_sample_sst_isothermal", rae2822_sample, 2, ("U", "Res", "dt", "Amu_t")))
    return out


REF = "/root/reference"
DIGESTS = {"rae2822_sample_sst_isothermal"}


def rae2822_sample(d):
    """The reference's own sample case (RAE2822, 3 blocks with a reversed
    wake-cut interface, SST, Van Leer flux, characteristic MUSCL3, LU-SGS at
    CFL 10): its Mesh2d.dat as shipped, its bc2d_1.in and control.in as
    read_parameter reads them, rewritten without the comment text, with one
    change: an isothermal wall (Twall 288.15 K, the free-stream temperature)
    in place of the adiabatic one, whose SST omega ghost reads an unset
    density (sub_boundary.f90:379)."""
    import shutil

    shutil.copy(os.path.join(REF, "Mesh2d.dat"), os.path.join(d, "Mesh2d.dat"))
    F.write_bcin(os.path.join(d, "bc2d_1.in"), F.read_bcin(os.path.join(REF, "bc2d_1.in")))
    c = F.read_control(os.path.join(REF, "control.in"))
    c.Twall = 288.15
    F.write_control(os.path.join(d, "control.in"), c)


def dual_fixture(name, steps=3):
    """Dual-time LU-SGS (Time_Method -1), Euler with early inner exits.  The
    reference's first-call flags (SAVE variables) need a process that has not
    run a residual yet, so tests/ref_dual_run.py runs it in a new one; the
    inner-loop indices come from the reference's own "Inner step" lines."""
    import subprocess

    from helpers import dual_case

    d = dual_case("golden_dual", False, 3.5e-3, 8)
    files = {f: open(os.path.join(d, f)).read() for f in FILES if os.path.exists(os.path.join(d, f))}
    with tempfile.TemporaryDirectory() as t:
        out = os.path.join(t, "ref.npz")
        r = subprocess.run([sys.executable, os.path.join(ROOT, "tests", "ref_dual_run.py"), d, str(steps), out],
                           capture_output=True, text=True, check=True)
        z = np.load(out)
        data = {f"file:{k}": np.array(v) for k, v in files.items()}
        data["steps"] = np.array(steps)
        data["nblock"] = np.array(1)
        data["res_rms"] = np.array([z[f"rms_{k}"] for k in range(steps)])
        data["res_max"] = np.array([z[f"max_{k}"] for k in range(steps)])
        data["inner"] = np.array([int(line.split()[-1]) for line in r.stdout.splitlines() if "Inner step" in line])
        for a in ("U", "Res", "dt"):
            data[f"{a}:1"] = z[f"{a}_{steps - 1}"]
    np.savez_compressed(os.path.join(HERE, f"{name}.npz"), **data)
    print(name, "steps", steps, "inner", data["inner"])


def main():
    import refpy

    if not refpy.available():
        raise SystemExit("reference build missing: make -C oracle -f Makefile.ref")
    only = set(sys.argv[1:])  # fixture names to write (default: all)
    if not only or "euler_dual_time_lusgs" in only:
        dual_fixture("euler_dual_time_lusgs")
    for name, build, steps, arrays in specs():
        if only and name not in only:
            continue
        with tempfile.TemporaryDirectory() as d:
            build(d)
            files = {f: open(os.path.join(d, f)).read() for f in FILES if os.path.exists(os.path.join(d, f))}
            R = refpy.RefCase(d)
            nblock = R.mesh_info(1)["nblock"]
            rms, rmax = [], []
            for _ in range(steps):
                R.op("main_step")
                mx, rm = R.norms(1)
                rms.append(rm)
                rmax.append(mx)
            data = {f"file:{k}": np.array(v) for k, v in files.items()}
            if name in DIGESTS:
                # large case: input files as bytes, arrays as SHA-256 of their
                # FP64 bytes (bit-exact parity needs no more) plus their shapes
                data = {f"file:{k}": np.frombuffer(v.encode(), np.uint8) for k, v in files.items()}
            data["steps"] = np.array(steps)
            data["nblock"] = np.array(nblock)
            data["res_rms"] = np.array(rms)
            data["res_max"] = np.array(rmax)
            for mb in range(1, nblock + 1):
                for a in arrays:
                    arr = R.get(a, 1, mb)
                    if name in DIGESTS:
                        data[f"sha256:{a}:{mb}"] = np.array(hashlib.sha256(np.ascontiguousarray(arr).tobytes()).hexdigest())
                        data[f"shape:{a}:{mb}"] = np.array(arr.shape)
                    else:
                        data[f"{a}:{mb}"] = arr
            np.savez_compressed(os.path.join(HERE, f"{name}.npz"), **data)
            print(name, "steps", steps, "blocks", nblock, "rms", rms[-1][:4])


if __name__ == "__main__":
    main()
