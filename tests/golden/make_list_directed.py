"""Write tests/golden/flang_list_directed.npz: how flang's runtime writes
REAL(8) values under list-directed output (`write(u,*) x`), the editing the
reference's wall-Cp tool uses for Cp.dat (read_flow-1.2.f90:44-47).

    python tests/golden/make_list_directed.py

Compiles a four-line Fortran program of this file (not reference code) with
/opt/rocm/llvm/bin/flang, feeds it the bit patterns below and stores each
pattern with the text flang printed.  The patterns: seeded random finite
doubles over the whole exponent range, values in [1, 2), subnormals, powers of
two and ten and their neighbours, and the edges of the F/E switch (0.1, 1e15).
"""
import os
import struct
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
FLANG = os.environ.get("FLANG", "/opt/rocm/llvm/bin/flang")
PROGRAM = """program ld
  integer(8) :: b
  real(8) :: x
  do
    read(*, *, end=9) b
    x = transfer(b, x)
    write(*, *) x
  end do
9 continue
end program
"""


def bits(x: float) -> int:
    return struct.unpack("<q", struct.pack("<d", x))[0]


def patterns(seed=20260):
    rng = np.random.default_rng(seed)
    out = [int(v) for v in rng.integers(0, 0x7FEFFFFFFFFFFFFF, 1500, dtype=np.int64)]
    out += [int(v) | 0x3FF0000000000000 for v in rng.integers(0, 1 << 52, 1000, dtype=np.int64)]
    out += [int(v) for v in rng.integers(1, 1 << 52, 100, dtype=np.int64)]  # subnormals
    special = [0.0, 5e-324, 2.0 ** -1022, 2.0 ** -1074 * 3, 1.0, 0.1, 0.01, 0.099, 0.0995, 0.09999999999999999,
               0.095, 0.094999, 1e14, 1e15, 9.5e14, 9.49e14, 9.99e14, 1e16, 1e300, 1e-300, 12.0, 1 / 3]
    for e in range(-60, 61, 3):
        special += [2.0 ** e, 10.0 ** e]
    for v in list(special):
        if v > 0:
            special += [np.nextafter(v, 0.0), np.nextafter(v, np.inf)]
    out += [bits(v) for v in special]
    out += [(b | (1 << 63)) - (1 << 64) for b in out[::7]]  # negatives (as signed 64-bit integers)
    return out


def main():
    pats = patterns()
    with tempfile.TemporaryDirectory() as tmp:
        src = os.path.join(tmp, "ld.f90")
        exe = os.path.join(tmp, "ld")
        with open(src, "w") as f:
            f.write(PROGRAM)
        subprocess.run([FLANG, src, "-o", exe], check=True, cwd=tmp)
        res = subprocess.run([exe], input="\n".join(str(b) for b in pats) + "\n", capture_output=True, text=True,
                             check=True)
    lines = res.stdout.splitlines()
    assert len(lines) == len(pats)
    text = np.array([ln.strip() for ln in lines], dtype="S32")
    np.savez_compressed(os.path.join(HERE, "flang_list_directed.npz"), bits=np.array(pats, dtype=np.int64), text=text)
    print(f"{len(pats)} values")


if __name__ == "__main__":
    main()
