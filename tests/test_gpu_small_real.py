"""GPU parity of even-length r2c / c2r passes computed by the split form of
the short-length kernel (the half-length complex line of length nf = n/2
for nf in the short (above 16) and mid length lists, split in one LDS
chunk)
against the oracle: real lines contiguous and strided (transposed layouts),
embed / trunc and shifts on the real dimension, both precisions, and the
adjoint passes (Hermitian weights) through dot-product identities.  Plans
use PFFT_ESTIMATE so that no timing-based choice replaces the kernel."""
import numpy as np
import pytest

import pfft_amd as P
from gpu_util import (spec_of, random_input, run_gpu, oracle_out, rel_l2, out_mask, comm_for,
                      n_scalars)

pytestmark = pytest.mark.gpu

TO, TI, SI, SO, EST = (P.TRANSPOSED_OUT, P.TRANSPOSED_IN, P.SHIFTED_IN, P.SHIFTED_OUT,
                       P.ESTIMATE)


def covered_lengths(f64):
    """Every half length the dispatcher sends to the split real kernel (the
    library's own rule, pfk_small_real_length)."""
    import ctypes
    L = P.lib()
    L.pfk_small_real_length.restype = ctypes.c_int
    L.pfk_small_real_length.argtypes = [ctypes.c_int, ctypes.c_int]
    return [nf for nf in range(1, 4097) if L.pfk_small_real_length(nf, int(f64))]


F32 = covered_lengths(False)
F64 = covered_lengths(True)


def tol(prec):
    return 1e-12 if prec == "f64" else 2e-6


def real_kernel(spec, prec):
    """Kernel names of the plan's r2c / c2r passes."""
    esz = 16 if prec == "f64" else 8
    d = P.DeviceBuffer(esz * (n_scalars(spec, "in") + n_scalars(spec, "out")))
    pl = P.Plan(spec.kind, spec.n, d.ptr, d.ptr, spec.ni, spec.no, spec.howmany, spec.sign,
                spec.flags | EST, comm=comm_for(spec.rnk_pm), prec=prec)
    assert pl.ok
    names = [ln.split("kernel=")[-1] for ln in pl.describe()
             if (" %s " % spec.kind) in ln]
    pl.destroy()
    d.free()
    return names


def check(kind, n, prec, flags=0, ni=None, no=None, sign=None):
    if sign is None:
        sign = +1 if kind == "c2r" else -1
    spec = spec_of(kind, n, ni=ni, no=no, sign=sign, flags=flags)
    x = random_input(spec, sum(n))
    if prec == "f32":
        x = x.astype(np.float32).astype(np.float64)
    want = oracle_out(spec, x)
    got = run_gpu(spec, x, prec=prec, plan_flags_extra=EST)
    return rel_l2(got, want, out_mask(spec))


def cases():
    for nf in F32:
        yield nf, "f32"
    for nf in F64:
        yield nf, "f64"


def test_covered_length_lists():
    """the enumeration finds the lengths of the dispatcher's lists,
    including the mid lengths and the whole-chunk shapes 64 and 128"""
    for lst in (F32, F64):
        for nf in (18, 24, 64, 128, 144, 160, 200, 216, 480, 960):
            assert nf in lst, nf
    assert 40 in F32 and 80 in F64 and 864 in F32


@pytest.mark.parametrize("nf,prec", list(cases()))
def test_real_split_lengths(nf, prec):
    n = [3, 5, 2 * nf]
    for kind, flags in (("r2c", 0), ("r2c", TO), ("c2r", 0), ("c2r", TI)):
        spec = spec_of(kind, n, sign=(+1 if kind == "c2r" else -1), flags=flags)
        names = real_kernel(spec, prec)
        if nf & (nf - 1) == 0 and nf >= 64:
            # plain power-of-two real lines run the power-of-two kernels;
            # embedded ones (below) take the split kernel
            assert names and all(k.startswith("fast") for k in names), (kind, flags, names)
        else:
            assert all(k == "small" for k in names), (kind, flags, names)
        err = check(kind, n, prec, flags)
        assert err <= tol(prec), (kind, flags, err)


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("nf", [64, 128])
def test_real_split_whole_chunk_shapes(nf, prec):
    """64 and 128: the split kernel runs them when the power-of-two kernel
    declines the pass (embedded real lines, or the adjoint's weights)."""
    L = 2 * nf
    for kind, ni, no in (("r2c", [3, 5, L - 6], [3, 5, L]), ("c2r", [3, 5, L], [3, 5, L - 10])):
        n = [3, 5, L]
        spec = spec_of(kind, n, ni=ni, no=no, sign=(+1 if kind == "c2r" else -1))
        assert all(k == "small" for k in real_kernel(spec, prec)), (kind, real_kernel(spec, prec))
        err = check(kind, n, prec, 0, ni=ni, no=no)
        assert err <= tol(prec), (kind, err)


@pytest.mark.parametrize("nf", [24, 60, 120, 240, 480])
def test_real_split_embed_trunc_shift(nf):
    L = 2 * nf
    for kind in ("r2c", "c2r"):
        for ni, no, flags in (([4, 3, L - 6], [4, 3, L - 10], 0),
                              ([4, 4, L - 4], [4, 4, L - 8], SI | SO),
                              ([4, 4, L], [4, 4, L], SI | SO | (TO if kind == "r2c" else TI))):
            if kind == "c2r":
                ni, no = no, ni
            n = [4, ni[1], L]
            err = check(kind, n, "f32", flags, ni=ni, no=no)
            assert err <= 2e-6, (kind, ni, no, flags, err)


@pytest.mark.parametrize("nf,prec", [(60, "f64"), (240, "f64"), (96, "f32")])
def test_real_split_adjoint_dot_product(nf, prec):
    n = [3, 4, 2 * nf]
    for kind, flags, sign in (("r2c", TO, -1), ("c2r", TI, +1)):
        spec = spec_of(kind, n, sign=sign, flags=flags)
        rng = np.random.default_rng(nf)
        x = rng.standard_normal(n_scalars(spec, "in"))
        y = rng.standard_normal(n_scalars(spec, "out"))
        if kind == "c2r":
            x = oracle_out(spec_of("r2c", n, flags=TO), rng.standard_normal(int(np.prod(n))))
        if prec == "f32":
            x = x.astype(np.float32).astype(np.float64)
            y = y.astype(np.float32).astype(np.float64)
        fx = run_gpu(spec, x, prec=prec, plan_flags_extra=EST)
        fty = run_gpu(spec, y, prec=prec, adjoint=True, plan_flags_extra=EST)
        lhs, rhs = float(fx @ y), float(x @ fty)
        t = 1e-10 if prec == "f64" else 1e-4
        assert abs(lhs - rhs) <= t * (abs(lhs) + abs(rhs)), (kind, lhs, rhs)


@pytest.mark.parametrize("kind,flags", [("r2c", P.PADDED_R2C), ("r2c", TO | P.PADDED_R2C),
                                        ("c2r", P.PADDED_C2R), ("c2r", TI | P.PADDED_C2R)])
@pytest.mark.parametrize("nf", [60, 240])
def test_real_split_padded_inplace(kind, flags, nf):
    spec = spec_of(kind, [6, 4, 2 * nf], sign=(+1 if kind == "c2r" else -1), flags=flags)
    x = random_input(spec, 11)
    got = run_gpu(spec, x, inplace=True, plan_flags_extra=EST)
    assert rel_l2(got, oracle_out(spec, x), out_mask(spec)) <= 1e-12


@pytest.mark.parametrize("kind,flags", [("r2c", TO), ("c2r", TI), ("c2c", 0), ("c2c", TO)])
def test_split_howmany(kind, flags):
    spec = spec_of(kind, [5, 4, 480], howmany=3, sign=(+1 if kind == "c2r" else -1),
                   flags=flags)
    x = random_input(spec, 12)
    got = run_gpu(spec, x, plan_flags_extra=EST)
    assert rel_l2(got, oracle_out(spec, x), out_mask(spec)) <= 1e-12
