"""Writes the committed golden fixtures used by the CPU tests.

  init_input.npz  values of pfft_init_input_* (api/api-basic.c:60-121,
                  254-264) for several shapes/starts, evaluated here from the
                  closed form with Python integers and IEEE doubles;
  dft_kat.npz     known-answer DFTs: seeded inputs and their DFTs by direct
                  summation in long double with exact integer reduction of
                  j*k mod n (no FFT algorithm involved).

The reference ships no golden data (its tests only check roundtrips), so
these fixtures come from the mathematics.  Run from the repo root:
    python tests/golden/gen_golden.py
"""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


# ---- pfft_init_input_* closed form ---------------------------------------
def plain_index(n, kvec):
    k = 0
    for t in range(len(n)):
        k += k * n[t] + kvec[t]
    return k


def c_mod(a, b):
    r = abs(a) % abs(b)
    return r if a >= 0 else -r


def init_scalar(n, ind):
    per = [c_mod(ind[t], n[t]) for t in range(len(n))]
    m = float(plain_index(n, per))
    if m == 0:
        return complex(1500.0, 1250.0)
    return complex(1000.0 / (2 * m), 1000.0 / (2 * m + 1))


def init_array(n, ln, ls, kind):
    out = []
    for flat in range(int(np.prod(ln))):
        kl = []
        r = flat
        for t in reversed(range(len(n))):
            kl.append(r % ln[t])
            r //= ln[t]
        kl = kl[::-1]
        kg = [kl[t] + ls[t] for t in range(len(n))]
        km = [n[t] - kg[t] for t in range(len(n))]
        d1 = init_scalar(n, kg)
        if kind == "real":
            out.append(d1.real if kg[-1] < n[-1] else 0.0)
        elif kind == "complex":
            out.append(d1)
        else:
            d2 = init_scalar(n, km)
            out.append(0.5 * (d1 + d2.conjugate()))
    return np.array(out)


INIT_CASES = [
    ("complex", [4, 5, 6], [4, 5, 6], [0, 0, 0]),
    ("complex", [6, 4, 8], [6, 4, 8], [-3, -2, -4]),      # shifted starts
    ("hermitian", [4, 6, 5], [4, 6, 3], [0, 0, 0]),
    ("hermitian", [8, 6, 4], [8, 6, 3], [-4, -3, -2]),
    ("real", [4, 5, 6], [4, 5, 8], [0, 0, 0]),            # padded rows: 2*(6/2+1)
    ("real", [3, 4, 5], [3, 4, 5], [0, 0, 0]),
    ("complex", [3, 4, 2, 5], [3, 4, 2, 5], [0, 0, 0, 0]),
]


# ---- brute-force DFT known answers ----------------------------------------
def dft_direct(x, sign):
    """y_k = sum_j x_j exp(sign 2 pi i j k / n), long double, j*k reduced mod n."""
    n = len(x)
    j = np.arange(n)
    y = np.zeros(n, dtype=np.clongdouble)
    two_pi = np.longdouble(2) * np.arccos(np.longdouble(-1))
    for k in range(n):
        e = (j * k) % n
        ang = two_pi * e.astype(np.longdouble) / np.longdouble(n)
        w = np.cos(ang) + sign * 1j * np.sin(ang)
        y[k] = np.sum(x.astype(np.clongdouble) * w)
    return y


KAT_SIZES = [1, 2, 3, 4, 5, 6, 7, 8, 9, 12, 15, 16, 27, 29, 31, 32, 64]


def main():
    d = {}
    for i, (kind, n, ln, ls) in enumerate(INIT_CASES):
        v = init_array(n, ln, ls, kind)
        d["case%d_kind" % i] = np.array(kind)
        d["case%d_n" % i] = np.array(n)
        d["case%d_ln" % i] = np.array(ln)
        d["case%d_ls" % i] = np.array(ls)
        d["case%d_val" % i] = v
    np.savez_compressed(os.path.join(HERE, "init_input.npz"), **d)

    rng = np.random.default_rng(20261202)
    k = {}
    for n in KAT_SIZES:
        x = rng.uniform(-1, 1, n) + 1j * rng.uniform(-1, 1, n)
        k["x%d" % n] = x
        for sign, nm in ((-1, "fwd"), (+1, "bwd")):
            y = dft_direct(x, sign)
            k["y%d_%s_re" % (n, nm)] = y.real.astype(np.float64)
            k["y%d_%s_im" % (n, nm)] = y.imag.astype(np.float64)
            # residual beyond double precision, kept so a test can check
            # the double-rounded values against long double as well
            k["y%d_%s_re_lo" % (n, nm)] = (y.real - y.real.astype(np.float64)).astype(np.float64)
            k["y%d_%s_im_lo" % (n, nm)] = (y.imag - y.imag.astype(np.float64)).astype(np.float64)
    np.savez_compressed(os.path.join(HERE, "dft_kat.npz"), **k)
    print("wrote init_input.npz (%d cases), dft_kat.npz (%d sizes)" % (len(INIT_CASES),
                                                                      len(KAT_SIZES)))


if __name__ == "__main__":
    main()
