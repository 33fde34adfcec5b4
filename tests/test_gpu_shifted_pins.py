"""GPU parity of the both-shifted real transforms against the identities of
tests/test_shifted_pins.py (direct centred sums, the c2c-vs-c2r agreement
of tests/simple_check_c2r_c2c_shifted.c after the forward step), and the
differentiation configuration's forward transform (ni = 128^3 real input
zero-padded to n = 256^3, SHIFTED_IN | SHIFTED_OUT) at full size against
the FP64 oracle."""
import os

import numpy as np
import pytest

import pfft_amd as P
from oracle import semantics as S
from gpu_util import spec_of, run_gpu, rel_l2
from test_shifted_pins import (CASES, SI, SO, c2c_vs_c2r_forward, c2r_centred, flat_c,
                               r2c_centred)

pytestmark = pytest.mark.gpu


def gpu_exec(prec):
    def ex(spec, x):
        return run_gpu(spec, x, prec=prec)
    return ex


@pytest.mark.parametrize("prec,tol", [("f64", 1e-12), ("f32", 2e-6)])
@pytest.mark.parametrize("n,ni,no", CASES)
def test_gpu_both_shifted_r2c_is_centred_dft(n, ni, no, prec, tol):
    spec = spec_of("r2c", n, ni, no, flags=SI | SO)
    x = np.random.default_rng(4).standard_normal(int(np.prod(ni)))
    if prec == "f32":
        x = x.astype(np.float32).astype(np.float64)
    got = run_gpu(spec, x, prec=prec)
    want = r2c_centred(x.reshape(ni), n, ni, no)
    assert rel_l2(got, flat_c(np.asarray(want, dtype=np.complex128))) < tol


@pytest.mark.parametrize("prec,tol", [("f64", 1e-12), ("f32", 2e-6)])
@pytest.mark.parametrize("n,ni,no", CASES)
def test_gpu_both_shifted_c2r_is_centred_inverse_dft(n, ni, no, prec, tol):
    spec = spec_of("c2r", n, ni, no, sign=+1, flags=SI | SO)
    rng = np.random.default_rng(5)
    shp = (ni[0], ni[1], ni[2] // 2 + 1)
    X = rng.uniform(-1, 1, shp) + 1j * rng.uniform(-1, 1, shp)
    flat = flat_c(X)
    if prec == "f32":
        flat = flat.astype(np.float32).astype(np.float64)
        X = flat[0::2].reshape(shp) + 1j * flat[1::2].reshape(shp)
    got = run_gpu(spec, flat, prec=prec)
    want = c2r_centred(X, n, ni, no).real.astype(np.float64).ravel()
    assert rel_l2(got, want) < tol


@pytest.mark.parametrize("n,ni", [([4, 4, 6], [4, 4, 6]), ([6, 6, 6], [4, 4, 4]),
                                  ([8, 6, 10], [6, 6, 8]), ([32, 16, 24], [32, 16, 24])])
def test_gpu_c2c_and_c2r_agree_after_forward(n, ni):
    yc, yr = c2c_vs_c2r_forward(n, ni, gpu_exec("f64"))
    scale = np.abs(yc).max()
    assert np.abs(yc.imag).max() < 1e-12 * scale
    assert np.abs(yc.real - yr).max() < 1e-12 * scale
    # and both agree with the oracle's c2r
    _, yr_o = c2c_vs_c2r_forward(n, ni, S.execute_flat)
    assert np.abs(yr - yr_o).max() < 1e-12 * scale


@pytest.mark.parametrize("prec,tol", [("f64", 1e-12), ("f32", 2e-6)])
def test_gpu_cfg5_forward_full_size(prec, tol):
    """ni = 128^3 real N(0,1) zero-padded to n = 256^3, both shifted: the
    whole stored half spectrum against the FP64 oracle."""
    n, ni = [256] * 3, [128] * 3
    spec = spec_of("r2c", n, ni, n, flags=SI | SO)
    x = np.random.default_rng(20261202).standard_normal(128 ** 3)
    if prec == "f32":
        x = x.astype(np.float32).astype(np.float64)
    got = run_gpu(spec, x, prec=prec)
    workers = S.WORKERS
    S.WORKERS = min(16, len(os.sched_getaffinity(0)))
    try:
        want = S.execute_flat(spec, x, precision=np.float64)
    finally:
        S.WORKERS = workers
    assert rel_l2(got, want) < tol
