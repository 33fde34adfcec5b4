"""The oracle's r2r transforms (tests only; no device needed).

Pinned by FFTW's own outputs: tests/golden/fftw_r2r_ref.npz holds the
FFTW-computed REDFT/RODFT of x = 0..n-1 (written by
tests/golden/gen_r2r_golden.py from the FFTW reference data scipy ships),
and by the defining sums of FFTW's r2r kinds evaluated directly (which also
pin the DHT, absent from that data)."""
import os

import numpy as np
import pytest

from oracle import semantics as S

GOLD = os.path.join(os.path.dirname(__file__), "golden", "fftw_r2r_ref.npz")
# scipy/FFTW name -> FFTW kind number
KIND = {("dct", 1): 3, ("dct", 3): 4, ("dct", 2): 5, ("dct", 4): 6,
        ("dst", 1): 7, ("dst", 3): 8, ("dst", 2): 9, ("dst", 4): 10}


def brute(kind, x):
    """FFTW's definitions of the r2r kinds (doc 'What FFTW Really
    Computes'), as O(n^2) sums in long double."""
    x = np.asarray(x, dtype=np.longdouble)
    n = len(x)
    j = np.arange(n, dtype=np.longdouble)
    pi = np.longdouble(np.pi)
    y = np.empty(n, dtype=np.longdouble)
    for k in range(n):
        if kind == 2:
            y[k] = np.sum(x * (np.cos(2 * pi * j * k / n) + np.sin(2 * pi * j * k / n)))
        elif kind == 3:
            y[k] = x[0] + (-1) ** k * x[n - 1] + 2 * np.sum(
                x[1:n - 1] * np.cos(pi * j[1:n - 1] * k / (n - 1)))
        elif kind == 4:
            y[k] = x[0] + 2 * np.sum(x[1:] * np.cos(pi * j[1:] * (k + 0.5) / n))
        elif kind == 5:
            y[k] = 2 * np.sum(x * np.cos(pi * (j + 0.5) * k / n))
        elif kind == 6:
            y[k] = 2 * np.sum(x * np.cos(pi * (j + 0.5) * (k + 0.5) / n))
        elif kind == 7:
            y[k] = 2 * np.sum(x * np.sin(pi * (j + 1) * (k + 1) / (n + 1)))
        elif kind == 8:
            y[k] = (-1) ** k * x[n - 1] + 2 * np.sum(
                x[:n - 1] * np.sin(pi * (j[:n - 1] + 1) * (k + 0.5) / n))
        elif kind == 9:
            y[k] = 2 * np.sum(x * np.sin(pi * (j + 0.5) * (k + 1) / n))
        elif kind == 10:
            y[k] = 2 * np.sum(x * np.sin(pi * (j + 0.5) * (k + 0.5) / n))
    return y


def test_oracle_matches_fftw_vectors():
    g = np.load(GOLD)
    assert len(g.files) >= 80
    worst = 0.0
    for key in g.files:
        fn, t, n = key.split("_")
        kind = KIND[(fn, int(t))]
        x = np.arange(int(n), dtype=np.float64)
        got = S.dft_r2r(x, 0, kind)
        want = g[key]
        err = np.max(np.abs(got - want)) / max(np.max(np.abs(want)), 1.0)
        worst = max(worst, err)
        assert err < 1e-13, (key, err)
    assert worst < 1e-13


@pytest.mark.parametrize("kind", range(2, 11))
@pytest.mark.parametrize("n", [2, 3, 5, 8, 12])
def test_oracle_matches_definitions(kind, n):
    x = np.random.default_rng(kind * 100 + n).standard_normal(n)
    want = brute(kind, x)
    got = S.dft_r2r(x, 0, kind)
    assert np.max(np.abs(got - np.asarray(want, dtype=np.float64))) < 1e-12 * max(1.0, n)


def test_dht_is_its_own_inverse_up_to_n():
    x = np.random.default_rng(1).standard_normal(17)
    y = S.dft_r2r(S.dft_r2r(x, 0, 2), 0, 2)
    assert np.max(np.abs(y / 17 - x)) < 1e-13


def test_planted_kind_swap_is_detected():
    """Swapping REDFT10 and REDFT01 (DCT-II / DCT-III, transposes of each
    other) must fail the FFTW vectors."""
    g = np.load(GOLD)
    x = np.arange(8, dtype=np.float64)
    assert np.max(np.abs(S.dft_r2r(x, 0, 4) - g["dct_2_8"])) > 1.0


def test_multidim_r2r_plan_is_separable():
    """A 3-d r2r plan of the oracle applies each dimension's kind along that
    dimension (FFTW r2r plans of rank > 1)."""
    n = [4, 6, 5]
    kinds = [5, 7, 2]
    spec = S.PlanSpec("r2r", n, kinds=kinds)
    x = np.random.default_rng(3).standard_normal(int(np.prod(n)))
    got = S.execute_flat(spec, x).reshape(n)
    a = x.reshape(n)
    for d, k in enumerate(kinds):
        a = np.apply_along_axis(lambda v: np.asarray(brute(k, v), dtype=np.float64), d, a)
    assert np.max(np.abs(got - a)) < 1e-11
