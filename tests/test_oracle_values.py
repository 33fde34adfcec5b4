"""Pinning the oracle (CPU).

The reference holds no golden outputs, so the oracle's values are pinned
against what the mathematics fixes:
  * brute-force DFT sums (tests/golden/dft_kat.npz) for the 1-D kernels;
  * separable brute force for a small 3-D transform;
  * the centred DFT of doc/features.tex:36-70 for both-shifted plans;
  * zero-padding / slicing closed forms for oversampled and pruned plans;
  * FFTW's c2r definition (Hermitian extension, Im of DC/Nyquist ignored);
and the reference's quirks (single-sided shift guard, the upper-end r2c
truncation, the plain_index test data) are checked against closed forms
restated from the cited source lines.  A planted mistake in the oracle
must fail these checks (test_planted_*).
"""
import os

import numpy as np
import pytest

import pfft_amd as P
from oracle import layout as L
from oracle import semantics as S

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
TI, TO, SI, SO = P.TRANSPOSED_IN, P.TRANSPOSED_OUT, P.SHIFTED_IN, P.SHIFTED_OUT


def cplx(flat):
    return flat[0::2] + 1j * flat[1::2]


def brute_dft(x, sign, axis):
    """Direct DFT along one axis in long double with exact reduction."""
    x = np.moveaxis(np.asarray(x, dtype=np.clongdouble), axis, -1)
    n = x.shape[-1]
    j = np.arange(n)
    two_pi = np.longdouble(2) * np.arccos(np.longdouble(-1))
    e = np.outer(j, j) % n
    ang = two_pi * e.astype(np.longdouble) / np.longdouble(n)
    W = np.cos(ang) + sign * 1j * np.sin(ang)
    y = x @ W.T
    return np.moveaxis(y, -1, axis)


def brute_dftn(x, sign):
    for ax in range(x.ndim):
        x = brute_dft(x, sign, ax)
    return x


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a - b, dtype=np.clongdouble)) /
                 max(np.linalg.norm(np.asarray(b, dtype=np.clongdouble)), 1e-300))


# ---------------------------------------------------------------- 1-D kernels
@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 6, 7, 8, 9, 12, 15, 16, 27, 29, 31, 32, 64])
def test_oracle_dfts_match_known_answers(n):
    k = np.load(os.path.join(GOLD, "dft_kat.npz"))
    x = k["x%d" % n].astype(np.clongdouble)
    for sign, nm in ((-1, "fwd"), (+1, "bwd")):
        want = (k["y%d_%s_re" % (n, nm)].astype(np.longdouble) + k["y%d_%s_re_lo" % (n, nm)]) + \
            1j * (k["y%d_%s_im" % (n, nm)].astype(np.longdouble) + k["y%d_%s_im_lo" % (n, nm)])
        got = S.dft_c2c(x, 0, sign)
        assert rel(got, want) < 1e-17, (n, sign)


@pytest.mark.parametrize("n", [2, 3, 4, 5, 6, 8, 15, 16, 31, 32])
def test_oracle_r2c_c2r_definitions(n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n)
    full = brute_dft(x, -1, 0)
    got = S.dft_r2c(x.astype(np.longdouble), 0, n)
    assert rel(got, full[: n // 2 + 1]) < 1e-17
    # c2r: Hermitian extension of the half spectrum, Im of DC and Nyquist ignored
    X = rng.standard_normal(n // 2 + 1) + 1j * rng.standard_normal(n // 2 + 1)
    ext = np.zeros(n, dtype=np.clongdouble)
    ext[: n // 2 + 1] = X
    ext[0] = X[0].real
    if n % 2 == 0:
        ext[n // 2] = X[n // 2].real
    for k in range(1, (n + 1) // 2):
        ext[n - k] = np.conj(X[k])
    want = brute_dft(ext, +1, 0).real
    got = S.dft_c2r(X.astype(np.clongdouble), 0, n)
    assert rel(got, want) < 1e-17


# ---------------------------------------------------------------- whole plans
def run(kind, n, flags=0, ni=None, no=None, sign=-1, howmany=1, rnk_pm=2, skip=None, seed=3):
    spec = S.PlanSpec(kind, n, ni, no, howmany, sign, flags, rnk_pm, skip)
    rng = np.random.default_rng(seed)
    nin = int(np.prod(spec.in_memory_shape())) * (1 if spec.in_is_real() else 2)
    x = rng.uniform(-1, 1, nin)
    return spec, x, S.execute_flat(spec, x)


def test_c2c_3d_matches_separable_brute_force():
    n = [4, 3, 5]
    spec, x, y = run("c2c", n)
    want = brute_dftn(cplx(x).reshape(n), -1)
    assert rel(cplx(y).reshape(n), want) < 1e-15
    spec, x, y = run("c2c", n, sign=+1)
    assert rel(cplx(y).reshape(n), brute_dftn(cplx(x).reshape(n), +1)) < 1e-15


@pytest.mark.parametrize("rnk_pm,perm", [(2, (1, 2, 0)), (1, (1, 0, 2))])
def test_transposed_output_layouts(rnk_pm, perm):
    n = [4, 6, 5]
    _, x, y0 = run("c2c", n, rnk_pm=rnk_pm)
    _, _, y1 = run("c2c", n, flags=TO, rnk_pm=rnk_pm)
    a = cplx(y0).reshape(n)
    b = cplx(y1).reshape([n[p] for p in perm])
    assert np.array_equal(np.transpose(a, perm), b)
    # TRANSPOSED_IN reads that layout back
    spec = S.PlanSpec("c2c", n, sign=+1, flags=TI, rnk_pm=rnk_pm)
    z = cplx(S.execute_flat(spec, y1)).reshape(n)
    assert rel(z / np.prod(n), cplx(x).reshape(n)) < 1e-15


def test_both_shifted_is_centred_dft():
    """doc/features.tex:36-70: with SHIFTED_IN|SHIFTED_OUT the plan computes
    g_l = sum_{k=-ni/2}^{ni/2-1} gh_k exp(-2 pi i k l / n), l in [-no/2, no/2)."""
    for n, ni, no in (([4, 4, 6], [4, 4, 6], [4, 4, 6]), ([6, 6, 6], [4, 4, 4], [4, 4, 4]),
                      ([8, 6, 10], [4, 6, 6], [6, 2, 8])):
        spec, x, y = run("c2c", n, flags=SI | SO, ni=ni, no=no)
        xin = cplx(x).reshape(ni)
        out = np.zeros(no, dtype=np.clongdouble)
        two_pi = np.longdouble(2) * np.arccos(np.longdouble(-1))
        # direct triple sum over centred indices
        ks = [np.arange(-a // 2, a // 2) for a in ni]
        ls = [np.arange(-b // 2, b // 2) for b in no]
        for d in range(3):
            E = np.outer(ls[d], ks[d]).astype(np.longdouble)
            W = np.exp(-1j * two_pi * E / np.longdouble(n[d]))
            xin = np.moveaxis(np.tensordot(W, np.moveaxis(xin, d, 0), axes=(1, 0)), 0, d)
        out = xin
        assert rel(cplx(y).reshape(no), out) < 1e-15, (n, ni, no)


def test_oversampled_and_pruned_unshifted():
    n, ni, no = [8, 6, 10], [5, 4, 6], [8, 3, 7]
    spec, x, y = run("c2c", n, ni=ni, no=no)
    xp = np.zeros(n, dtype=np.clongdouble)
    xp[:5, :4, :6] = cplx(x).reshape(ni)               # zeros appended (ousample.c:241-245)
    full = brute_dftn(xp, -1)
    assert rel(cplx(y).reshape(no), full[:8, :3, :7]) < 1e-15   # first no kept (ousample.c:251-255)


def test_r2c_pruned_keeps_upper_half_spectrum():
    """kernel/ousample.c:262-272: r2c truncation keeps the last no/2+1 of
    the n/2+1 half-spectrum values along the last dimension."""
    n, no = [4, 4, 16], [4, 4, 8]
    spec, x, y = run("r2c", n, no=no)
    full = brute_dftn(x.reshape(n), -1)[:, :, : 16 // 2 + 1]
    kept = full[:, :, (16 // 2 + 1) - (8 // 2 + 1):]
    assert rel(cplx(y).reshape(4, 4, 5), kept) < 1e-15


def test_c2r_embed_places_data_at_upper_end():
    """kernel/ousample.c:294-304: c2r embedding puts the ni/2+1 inputs at
    half-spectrum positions [n/2 - ni/2, n/2]."""
    n, ni = [4, 16], [4, 8]
    spec = S.PlanSpec("c2r", n, ni=ni, sign=+1, rnk_pm=1)
    rng = np.random.default_rng(1)
    X = rng.uniform(-1, 1, 4 * 5 * 2)
    y = S.execute_flat(spec, X)
    half = np.zeros((4, 9), dtype=np.clongdouble)
    half[:, 4:] = cplx(X).reshape(4, 5)
    half = brute_dft(half, +1, 0)
    want = np.stack([S.dft_c2r(half[i], 0, 16) for i in range(4)])
    assert rel(y.reshape(4, 16), want) < 1e-15


def test_single_sided_shift_guard():
    """api/api-basic.c:1219-1227, 1272-1275: without the partner flag the
    +-1 modulation only touches indices below n/2 (the reference quirk)."""
    n = [4, 6]
    g = [np.where(np.arange(a) < a // 2, (-1.0) ** np.arange(a), 1.0) for a in n]
    spec, x, y = run("c2c", n, flags=SO, rnk_pm=1)
    xin = cplx(x).reshape(n) * np.outer(g[0], g[1])
    assert rel(cplx(y).reshape(n), brute_dftn(xin, -1)) < 1e-15
    spec, x, y = run("c2c", n, flags=SI, rnk_pm=1)
    want = brute_dftn(cplx(x).reshape(n), -1) * np.outer(g[0], g[1])
    assert rel(cplx(y).reshape(n), want) < 1e-15


def test_conjugated_directions():
    """c2r FORWARD and r2c BACKWARD (kernel/partrafo.c:426-443): the same
    transform with the other sign."""
    n = [4, 6]
    spec, X, y = run("c2r", n, sign=-1, rnk_pm=1)
    half = cplx(X).reshape(4, 4)
    t = brute_dft(half, -1, 0)
    want = np.stack([np.real(brute_dft(np.concatenate(
        [[t[i, 0].real], t[i, 1:3], [t[i, 3].real], np.conj(t[i, 1:3][::-1])]), -1, 0))
        for i in range(4)])
    assert rel(y.reshape(n), want) < 1e-15
    spec, x, y = run("r2c", n, sign=+1, rnk_pm=1)
    want = brute_dftn(x.reshape(n), +1)[:, :4]
    assert rel(cplx(y).reshape(4, 4), want) < 1e-15


def test_howmany_is_interleaved_independent_transforms():
    n = [4, 5, 6]
    spec, x, y = run("c2c", n, howmany=3)
    xs = cplx(x).reshape(n + [3])
    ys = cplx(y).reshape(n + [3])
    for c in range(3):
        assert rel(ys[..., c], brute_dftn(xs[..., c], -1)) < 1e-15


def test_skipped_dimension_is_not_transformed():
    n = [4, 5, 6]
    spec, x, y = run("c2c", n, skip=[0, 1, 0])
    xin = cplx(x).reshape(n)
    want = brute_dft(brute_dft(xin, -1, 0), -1, 2)
    assert rel(cplx(y).reshape(n), want) < 1e-15


# ---------------------------------------------------------------- test data
def test_init_input_tools_match_fixtures():
    """pfft_init_input_* (host arrays) against the closed form of
    api/api-basic.c:60-121 (tests/golden/init_input.npz)."""
    import ctypes
    g = np.load(os.path.join(GOLD, "init_input.npz"))
    Lb = P.lib()
    i = 0
    while "case%d_kind" % i in g:
        kind = str(g["case%d_kind" % i])
        n, ln, ls = g["case%d_n" % i], g["case%d_ln" % i], g["case%d_ls" % i]
        want = g["case%d_val" % i]
        rnk = len(n)
        if kind == "real":
            buf = np.zeros(int(np.prod(ln)))
            Lb.pfft_init_input_real(rnk, P.ivec(n), P.ivec(ln), P.ivec(ls), buf.ctypes.data)
            assert np.array_equal(buf, want), i
            err = Lb.pfft_check_output_real(rnk, P.ivec(n), P.ivec(ln), P.ivec(ls),
                                            buf.ctypes.data, P.COMM_WORLD)
        else:
            buf = np.zeros(2 * int(np.prod(ln)))
            f = Lb.pfft_init_input_complex if kind == "complex" else \
                Lb.pfft_init_input_complex_hermitian
            f(rnk, P.ivec(n), P.ivec(ln), P.ivec(ls), buf.ctypes.data)
            assert np.array_equal(cplx(buf), want), i
            chk = Lb.pfft_check_output_complex if kind == "complex" else \
                Lb.pfft_check_output_complex_hermitian
            err = chk(rnk, P.ivec(n), P.ivec(ln), P.ivec(ls), buf.ctypes.data, P.COMM_WORLD)
        assert err == 0.0
        i += 1
    assert i >= 7


def test_check_output_reports_max_abs_error():
    # api/api-basic.c:148-189: the global max of |data - expected|
    n = [3, 4, 5]
    buf = np.zeros(2 * 60)
    Lb = P.lib()
    Lb.pfft_init_input_complex(3, P.ivec(n), P.ivec(n), P.ivec([0, 0, 0]), buf.ctypes.data)
    buf[2 * 17] += 0.25
    buf[2 * 40 + 1] -= 0.5
    err = Lb.pfft_check_output_complex(3, P.ivec(n), P.ivec(n), P.ivec([0, 0, 0]),
                                       buf.ctypes.data, P.COMM_WORLD)
    assert err == 0.5


# ---------------------------------------------------------------- planted mistakes
def test_planted_wrong_sign_is_caught(monkeypatch):
    """A transposed-sign DFT planted in the oracle fails the known-answer
    check, so the checks above have teeth."""
    orig = S.dft_c2c
    monkeypatch.setattr(S, "dft_c2c", lambda a, axis, sign: orig(a, axis, -sign))
    spec, x, y = run("c2c", [4, 3, 5])
    want = brute_dftn(cplx(x).reshape(4, 3, 5), -1)
    assert rel(cplx(y).reshape(4, 3, 5), want) > 1e-3


def test_planted_guard_drop_is_caught(monkeypatch):
    """Dropping the temp < n/2 guard (i.e. using the textbook modulation)
    changes single-sided shifted results."""
    orig = S.twiddle_input_factors

    def textbook(spec):
        facs = orig(spec)
        return [np.where(np.arange(f.size) % 2, -1.0, 1.0) for f in facs]
    spec, x, y = run("c2c", [4, 6], flags=SO, rnk_pm=1)
    monkeypatch.setattr(S, "twiddle_input_factors", textbook)
    spec2, x2, y2 = run("c2c", [4, 6], flags=SO, rnk_pm=1)
    assert rel(cplx(y2), cplx(y)) > 1e-3
