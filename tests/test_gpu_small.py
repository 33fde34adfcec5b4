"""GPU parity of the register-resident short-length c2c kernel (lengths
2^a 3^b 5^c up to 128, and the per-pass mid lengths 144..960) against the
oracle, in both of its forms (one line per lane, and G lanes per line with
an LDS exchange for the longer lengths): every instantiated length, with
contiguous lines (the innermost dimension, staged through LDS) and strided
lines (an outer dimension, lanes over adjacent lines), ragged tiles, embed /
trunc and shifts on the short dimension.  The per-pass launch is reached
for the lengths that have no power-of-two kernel; the batch launch (one
plan per batch) reaches the short-length kernel for every length up to
128."""
import numpy as np
import pytest

import pfft_amd as P
from gpu_util import spec_of, random_input, run_gpu, oracle_out, rel_l2, comm_for, n_scalars

pytestmark = pytest.mark.gpu

SI, SO, TO = P.SHIFTED_IN, P.SHIFTED_OUT, P.TRANSPOSED_OUT
LENGTHS = [2, 3, 4, 5, 6, 8, 9, 10, 12, 15, 16, 18, 20, 24, 25, 27, 30, 32, 36, 40, 45, 48, 50,
           54, 60, 64, 72, 80, 90, 96, 100, 108, 120, 128]


def tol(prec):
    return 1e-12 if prec == "f64" else 2e-6


def run_batch(spec, x, prec):
    """One plan executed through the batch launch."""
    rdt = np.float64 if prec == "f64" else np.float32
    esz = np.dtype(rdt).itemsize
    nin, nout = n_scalars(spec, "in"), n_scalars(spec, "out")
    din, dout = P.DeviceBuffer(nin * esz), P.DeviceBuffer(nout * esz)
    plan = P.Plan(spec.kind, spec.n, din.ptr, dout.ptr, spec.ni, spec.no, spec.howmany,
                  spec.sign, spec.flags, comm=comm_for(spec.rnk_pm), prec=prec)
    assert plan.ok
    din.upload(np.asarray(x, dtype=rdt))
    P.execute_batch([plan])
    out = dout.download(rdt, nout).astype(np.float64)
    plan.destroy()
    din.free()
    dout.free()
    return out


def cases():
    for L in LENGTHS:
        for prec in ("f32", "f64"):
            if not (L in (45, 90) and prec == "f64"):   # 45 values per lane: complex64 only
                yield L, prec


@pytest.mark.parametrize("L,prec", list(cases()))
def test_small_lengths(L, prec):
    # the short dimension outermost (strided lines) and innermost (contiguous
    # lines); 7 lines per pass of the middle extent keeps the tiles ragged
    spec = spec_of("c2c", [L, 7, L])
    x = random_input(spec, L)
    if prec == "f32":
        x = x.astype(np.float32).astype(np.float64)
    want = oracle_out(spec, x)
    assert rel_l2(run_batch(spec, x, prec), want) <= tol(prec), "batch"
    nonpow2 = L & (L - 1) != 0 or L < 64
    if nonpow2:
        assert rel_l2(run_gpu(spec, x, prec=prec), want) <= tol(prec), "per pass"


@pytest.mark.parametrize("L", [24, 48, 60, 80, 96, 120, 128])
@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_small_embed_trunc_shift(L, prec):
    # shifts need even n, ni and no in every dimension
    for ni, no, flags in (([L - 5, 3, L - 7], [L - 3, 3, L - 9], 0),
                          ([L - 4, 4, L - 6], [L - 2, 4, L - 8], SI | SO),
                          ([L, 6, L], [L, 6, L], SI | SO | TO)):
        spec = spec_of("c2c", [L, ni[1], L], ni=ni, no=no, flags=flags, sign=+1)
        x = random_input(spec, 3)
        if prec == "f32":
            x = x.astype(np.float32).astype(np.float64)
        want = oracle_out(spec, x)
        assert rel_l2(run_batch(spec, x, prec), want) <= tol(prec), (ni, no, flags)
        assert rel_l2(run_gpu(spec, x, prec=prec), want) <= tol(prec), (ni, no, flags)


MID = [144, 160, 192, 200, 216, 240, 288, 320, 384, 400, 432, 480, 576, 640, 768, 800, 864, 960]
MID_F64 = set(MID)


def _kernels(spec, prec):
    esz = 16 if prec == "f64" else 8
    d = P.DeviceBuffer(esz * n_scalars(spec, "in") // 2 + esz * n_scalars(spec, "out") // 2)
    pl = P.Plan(spec.kind, spec.n, d.ptr, d.ptr, spec.ni, spec.no, spec.howmany, spec.sign,
                spec.flags | P.ESTIMATE, comm=comm_for(spec.rnk_pm), prec=prec)
    assert pl.ok
    lines = pl.describe()
    pl.destroy()
    d.free()
    return lines


@pytest.mark.parametrize("L", MID)
@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_mid_lengths(L, prec):
    # lengths above 128 run the split form through the per-pass launch
    # (complex128 only for the lengths whose split fits its registers)
    spec = spec_of("c2c", [L, 3, L], sign=(-1 if L % 3 else +1))
    x = random_input(spec, L)
    if prec == "f32":
        x = x.astype(np.float32).astype(np.float64)
    want = oracle_out(spec, x)
    # PFFT_ESTIMATE: no timing-based choice, the split form runs where it exists
    assert rel_l2(run_gpu(spec, x, prec=prec, plan_flags_extra=P.ESTIMATE), want) <= tol(prec)
    if prec == "f32" or L in MID_F64:
        names = [ln for ln in _kernels(spec, prec) if (" n=%d " % L) in ln]
        assert names and all(ln.endswith("kernel=small") for ln in names), names


@pytest.mark.parametrize("L", [144, 200, 480, 960])
def test_mid_embed_trunc_shift(L):
    for ni, no, flags in (([L - 5, 3, L - 7], [L - 3, 3, L - 9], 0),
                          ([L - 4, 4, L - 6], [L - 2, 4, L - 8], SI | SO)):
        spec = spec_of("c2c", [L, ni[1], L], ni=ni, no=no, flags=flags)
        x = random_input(spec, 5).astype(np.float32).astype(np.float64)
        want = oracle_out(spec, x)
        got = run_gpu(spec, x, prec="f32", plan_flags_extra=P.ESTIMATE)
        assert rel_l2(got, want) <= 2e-6, (ni, no, flags)


def test_small_kernel_named_in_describe():
    c = comm_for(2)
    d = P.DeviceBuffer(8 * 2 * 96 * 7 * 96)
    pl = P.Plan("c2c", [96, 7, 96], d.ptr, d.ptr, comm=c, prec="f32")
    assert pl.ok
    assert sum(ln.endswith("kernel=small") for ln in pl.describe()) == 2, pl.describe()
    pl.destroy()
    d.free()
