"""Reference side of the dual-time tests, run in a fresh process.

The reference keeps two first-call flags in SAVE variables (Sfac of
Comput_Residual_one_mesh, sub_Residual.f90:18,28-34; alfa1 of Dual_time_LU_SGS,
sub_NS_singlegrid.f90:241-247), so its dual-time stepping depends on whether
the process has run a residual before.  This script sets the case up in a new
process, runs main-loop passes and saves U, Un, Un1, Res, dt and Res_rms after
each (and Res_max); the reference prints its inner-iteration counts ("Inner
step ...") on stdout.

    python ref_dual_run.py CASE_DIR NSTEPS OUT.npz
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "opencfd-ec2d_amd"))

import refpy  # noqa: E402


def main():
    d, nsteps, out = sys.argv[1], int(sys.argv[2]), sys.argv[3]
    R = refpy.RefCase(d)
    arrs = {}
    for k in range(nsteps):
        R.op("main_step")
        for a in ("U", "Un", "Un1", "Res", "dt"):
            arrs[f"{a}_{k}"] = R.get(a)
        arrs[f"max_{k}"], arrs[f"rms_{k}"] = R.norms(1)
    np.savez(out, **arrs)


if __name__ == "__main__":
    main()
