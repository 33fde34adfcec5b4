"""Writes tests/golden/fftw_r2r_ref.npz: FFTW-computed DCT/DST (FFTW's
REDFT/RODFT kinds) of x = 0, 1, ..., n-1 in double precision, taken from
the reference data scipy ships with its fftpack tests
(scipy/fftpack/tests/fftw_double_ref.npz, computed with FFTW 3 by the
scipy authors).  Keys `dct_<type>_<n>` / `dst_<type>_<n>`; scipy's DCT/DST
type t is FFTW's REDFT/RODFT kind: type 1 = xx00, 2 = xx10, 3 = xx01,
4 = xx11.

    python tests/golden/gen_r2r_golden.py
"""
import os

import numpy as np
import scipy

SIZES = (2, 3, 4, 8, 12, 15, 16, 17, 32, 64, 128)


def main():
    src = os.path.join(os.path.dirname(scipy.__file__), "fftpack", "tests",
                       "fftw_double_ref.npz")
    d = np.load(src)
    out = {}
    for fn in ("dct", "dst"):
        for t in (1, 2, 3, 4):
            for n in SIZES:
                key = "%s_%d_%d" % (fn, t, n)
                if key in d.files:
                    out[key] = np.asarray(d[key], dtype=np.float64)
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "fftw_r2r_ref.npz")
    np.savez_compressed(dst, **out)
    print("wrote %d vectors to %s" % (len(out), dst))


if __name__ == "__main__":
    main()
