"""Pins of the both-shifted real transforms (CPU, oracle).

With PFFT_SHIFTED_IN | PFFT_SHIFTED_OUT the reference computes the centred
DFT (doc/features.tex:36-70): inputs indexed k in [-ni/2, ni/2), outputs
l in [-no/2, no/2).  For r2c the stored half of the last dimension is
l2 in [-no2/2, 0] (the r2c output's local start is -no2/2 and its extent
no2/2+1, kernel/partrafo.c:185-190,294-299); for c2r the input half is
k2 in [-ni2/2, 0].  Three identities fix these paths independently of the
oracle's own formulas:
  * r2c equals the direct centred sum on the stored half;
  * c2r equals the direct centred sum over the Hermitian extension of its
    half input (the self-paired planes k2 = 0 and, when not embedded,
    k2 = -n2/2 enter through their Hermitian projection, FFTW's c2r
    convention);
  * for Hermitian input the c2c and c2r transforms agree after the
    forward step (the comparison of tests/simple_check_c2r_c2c_shifted.c
    and simple_check_c2r_c2c_ousam_shifted.c, at their sizes and with
    their test data), and r2c equals c2c of the real input on the stored
    half.
A planted mistake (the stored half taken as l2 in [0, no2/2]) must fail.
"""
import numpy as np
import pytest

import pfft_amd as P
from oracle import semantics as S

SI, SO = P.SHIFTED_IN, P.SHIFTED_OUT
TWO_PI = np.longdouble(2) * np.arccos(np.longdouble(-1))


def cplx(flat):
    return flat[0::2] + 1j * flat[1::2]


def centred_sum(x, ks, ls, n, sign):
    """y[l] = sum_k x[k] exp(sign 2 pi i k.l / n) over the index lists"""
    y = np.asarray(x, dtype=np.clongdouble)
    for d in range(len(n)):
        E = np.outer(np.asarray(ls[d]), np.asarray(ks[d])).astype(np.longdouble)
        W = np.exp(sign * 1j * TWO_PI * E / np.longdouble(n[d]))
        y = np.moveaxis(np.tensordot(W, np.moveaxis(y, d, 0), axes=(1, 0)), 0, d)
    return y


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a - b, dtype=np.clongdouble)) /
                 np.linalg.norm(np.asarray(b, dtype=np.clongdouble)))


def centred(a):
    return np.arange(-(a // 2), a - a // 2)


CASES = [([4, 4, 6], [4, 4, 6], [4, 4, 6]),        # simple_check_c2r_c2c_shifted.c:32
         ([6, 6, 6], [4, 4, 4], [6, 6, 6]),        # simple_check_c2r_c2c_ousam_shifted.c:32-33
         ([8, 6, 10], [4, 6, 6], [8, 6, 10]),
         ([8, 6, 10], [8, 6, 10], [6, 4, 8]),
         ([6, 8, 12], [4, 4, 8], [4, 6, 6])]


def r2c_centred(x_real, n, ni, no):
    ks = [centred(a) for a in ni]
    ls = [centred(no[0]), centred(no[1]), np.arange(-(no[2] // 2), 1)]
    return centred_sum(x_real, ks, ls, n, -1)


def c2r_centred(X_half, n, ni, no, sign=+1):
    """The half input (k in centred ni on dims 0, 1; k2 in [-ni2/2, 0])
    embedded in the centred n grid, extended Hermitian over the full grid
    (X(-k) = conj X(k); the self-paired planes k2 = 0 and k2 = -n2/2 enter
    through their Hermitian projection), then the direct centred sum to
    l in no."""
    n0, n1, n2 = n
    full = np.zeros(n, dtype=np.clongdouble)
    o0, o1, o2 = n0 // 2 - ni[0] // 2, n1 // 2 - ni[1] // 2, n2 // 2 - ni[2] // 2
    full[o0:o0 + ni[0], o1:o1 + ni[1], o2:o2 + ni[2] // 2 + 1] = X_half
    # position of -k on the centred grid of extent a (even)
    neg = lambda a: (a - np.arange(a)) % a
    n0i, n1i = neg(n0), neg(n1)
    h = n2 // 2
    for j in range(1, h):
        # column k2 = +j from the conjugate of column k2 = -j at (-k0, -k1)
        full[:, :, h + j] = np.conj(full[np.ix_(n0i, n1i, [h - j])][:, :, 0])
    for p in (h, 0):   # k2 = 0 and k2 = -n2/2
        pl = full[:, :, p].copy()
        full[:, :, p] = 0.5 * (pl + np.conj(pl[np.ix_(n0i, n1i)]))
    ks = [centred(a) for a in n]
    ls = [centred(a) for a in no]
    return centred_sum(full, ks, ls, n, sign)


@pytest.mark.parametrize("n,ni,no", CASES)
def test_both_shifted_r2c_is_centred_dft_on_stored_half(n, ni, no):
    spec = S.PlanSpec("r2c", n, ni, no, flags=SI | SO)
    x = np.random.default_rng(4).standard_normal(int(np.prod(ni)))
    y = cplx(S.execute_flat(spec, x)).reshape(no[0], no[1], no[2] // 2 + 1)
    want = r2c_centred(x.reshape(ni), n, ni, no)
    assert rel(y, want) < 1e-15, (n, ni, no)


@pytest.mark.parametrize("n,ni,no", CASES)
def test_both_shifted_c2r_is_centred_inverse_dft(n, ni, no):
    spec = S.PlanSpec("c2r", n, ni, no, sign=+1, flags=SI | SO)
    rng = np.random.default_rng(5)
    shp = (ni[0], ni[1], ni[2] // 2 + 1)
    X = rng.uniform(-1, 1, shp) + 1j * rng.uniform(-1, 1, shp)
    flat = np.empty(2 * X.size)
    flat[0::2], flat[1::2] = X.real.ravel(), X.imag.ravel()
    y = S.execute_flat(spec, flat).reshape(no)
    want = c2r_centred(X, n, ni, no)
    assert np.abs(want.imag).max() < 1e-12 * np.abs(want).max()
    assert rel(y, want.real) < 1e-15, (n, ni, no)


def semirandom(N, k0, k1, k2):
    """test data of tests/simple_check_c2r_c2c_shifted.c:122-145"""
    if k0 == (N[0] + 1) // 2:
        return semirandom(N, -k0, k1, k2)
    if k1 == (N[1] + 1) // 2:
        return semirandom(N, k0, -k1, k2)
    if k2 == (N[2] + 1) // 2:
        return semirandom(N, k0, k1, -k2)
    di = lambda i: 1000.0 / (1 if i == 0 else i)
    l0 = k0 + N[0] if k0 else 1
    l1 = k1 - 2 * N[1] if k1 else 1
    l2 = k2 + N[2] if k2 else 1
    re = di(l0 + l1 * l2) + 3 * di(l0 * l1 + l2) + di(l2 * l0 + l1)
    im = 3 * di(l0 + l1 * l1) + di(l1 + l2 * l2) - di(l2 + l0 * l0)
    return re + 1j * im


def hermitian_input(n, ni):
    """init_input of the reference test: f(k) + conj f(-k) over the centred
    input range (simple_check_c2r_c2c_shifted.c:148-155).  When the input
    is embedded (ni < n) its edge k = -ni/2 has no partner +ni/2 inside the
    input, so those slices are zeroed to keep the embedded array Hermitian
    on the n grid (the identity below needs that)."""
    ks = [centred(a) for a in ni]
    x = np.empty(ni, dtype=np.complex128)
    for a0, k0 in enumerate(ks[0]):
        for a1, k1 in enumerate(ks[1]):
            for a2, k2 in enumerate(ks[2]):
                x[a0, a1, a2] = semirandom(n, k0, k1, k2) + np.conj(semirandom(n, -k0, -k1, -k2))
    for d in range(3):
        if ni[d] < n[d]:
            sl = [slice(None)] * 3
            sl[d] = 0
            x[tuple(sl)] = 0
    return x


def flat_c(X):
    f = np.empty(2 * X.size)
    f[0::2], f[1::2] = X.real.ravel(), X.imag.ravel()
    return f


def c2c_vs_c2r_forward(n, ni, execute):
    """c2c FORWARD of the Hermitian full input and c2r FORWARD of its stored
    half (k2 in [-ni2/2, 0]), both SHIFTED_IN | SHIFTED_OUT; returns
    (c2c output, c2r output)."""
    x = hermitian_input(n, ni)
    h = ni[2] // 2
    half = x[:, :, 0:h + 1]
    sc = S.PlanSpec("c2c", n, ni, n, sign=-1, flags=SI | SO)
    sr = S.PlanSpec("c2r", n, ni, n, sign=-1, flags=SI | SO)
    yc = cplx(execute(sc, flat_c(x))).reshape(n)
    yr = execute(sr, flat_c(half)).reshape(n)
    return yc, yr


@pytest.mark.parametrize("n,ni", [([4, 4, 6], [4, 4, 6]), ([6, 6, 6], [4, 4, 4]),
                                  ([8, 6, 10], [6, 6, 8])])
def test_c2c_and_c2r_agree_after_forward_on_hermitian_input(n, ni):
    yc, yr = c2c_vs_c2r_forward(n, ni, S.execute_flat)
    scale = np.abs(yc).max()
    assert np.abs(yc.imag).max() < 1e-13 * scale
    assert np.abs(yc.real - yr).max() < 1e-13 * scale


@pytest.mark.parametrize("n,ni,no", CASES)
def test_both_shifted_r2c_is_c2c_of_real_input(n, ni, no):
    x = np.random.default_rng(6).standard_normal(int(np.prod(ni)))
    sr = S.PlanSpec("r2c", n, ni, no, flags=SI | SO)
    sc = S.PlanSpec("c2c", n, ni, no, flags=SI | SO)
    yr = cplx(S.execute_flat(sr, x)).reshape(no[0], no[1], no[2] // 2 + 1)
    xc = np.zeros(2 * x.size)
    xc[0::2] = x
    yc = cplx(S.execute_flat(sc, xc)).reshape(no)
    # stored half l2 in [-no2/2, 0] = the first no2/2+1 columns of the centred output
    assert rel(yr, yc[:, :, :no[2] // 2 + 1]) < 1e-15


def test_planted_wrong_half_is_caught():
    """the stored half taken as l2 in [0, no2/2] must fail the pin"""
    n = ni = no = [4, 4, 6]
    spec = S.PlanSpec("r2c", n, ni, no, flags=SI | SO)
    x = np.random.default_rng(4).standard_normal(int(np.prod(ni)))
    y = cplx(S.execute_flat(spec, x)).reshape(4, 4, 4)
    ks = [centred(a) for a in ni]
    ls = [centred(4), centred(4), np.arange(0, 4)]
    wrong = centred_sum(x.reshape(ni), ks, ls, n, -1)
    assert rel(y, wrong) > 1e-3
