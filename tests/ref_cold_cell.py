"""Reference side of the cold-cell test, run in a fresh process: the
reference stops the process when a cell's temperature falls below 1e-5
(sub_Residual.f90:143-146), printing "T <1.d-5 !" with the cell's i, j, block
and T.  This script sets the case up, gives the listed interior cells of one
block zero internal energy (T = 0) and evaluates the residual.

    python ref_cold_cell.py CASE_DIR BLOCK I1 J1 [I2 J2 ...]
"""
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "opencfd-ec2d_amd"))

import refpy  # noqa: E402


def main():
    d, mb = sys.argv[1], int(sys.argv[2])
    cells = [(int(a), int(b)) for a, b in zip(sys.argv[3::2], sys.argv[4::2])]
    R = refpy.RefCase(d)
    U = R.get("U", 1, mb)
    for i, j in cells:
        q = U[j + 1, i + 1]
        q[3] = 0.5 * (q[1] ** 2 + q[2] ** 2) / q[0]
    R.set("U", U, 1, mb)
    sys.stdout.flush()
    R.op("residual", 1)
    print("no stop")


if __name__ == "__main__":
    main()
