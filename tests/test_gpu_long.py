"""GPU parity of dimensions longer than one workgroup's line (computed
length > 4096: two steps through a work buffer, n = n1 * n2) against the
oracle: c2c with power-of-two and mixed lengths, the long dimension
innermost, outermost and in the middle, transposed layouts, embed / trunc
and shifts on the long dimension, complex64, in-place, batch execution;
even-length r2c / c2r real lines (half-length complex line plus the
half-spectrum combine) in transposed, plain and padded layouts and both
directions; the adjoints; and the plan-time rejections."""
import numpy as np
import pytest

import pfft_amd as P
from gpu_util import (spec_of, random_input, run_gpu, oracle_out, rel_l2, n_scalars, comm_for,
                      out_mask)

pytestmark = pytest.mark.gpu

TO, TI = P.TRANSPOSED_OUT, P.TRANSPOSED_IN
SI, SO = P.SHIFTED_IN, P.SHIFTED_OUT


def check(n, ni=None, no=None, sign=-1, flags=0, rnk_pm=1, prec="f64", inplace=False,
          howmany=1):
    spec = spec_of("c2c", n, ni, no, howmany, sign, flags, rnk_pm)
    x = random_input(spec, 3)
    if prec == "f32":
        x = x.astype(np.float32).astype(np.float64)
    got = run_gpu(spec, x, prec=prec, inplace=inplace)
    err = rel_l2(got, oracle_out(spec, x))
    assert err <= (1e-12 if prec == "f64" else 2e-6), (n, ni, no, flags, err)


@pytest.mark.parametrize("N", [8192, 5000, 6000, 12288, 32768])
def test_long_innermost(N):
    check([2, N])


def test_long_outer_and_middle_dims():
    check([8192, 3, 2], rnk_pm=2)
    check([4, 6000, 3], rnk_pm=2)
    check([3, 10000, 2], rnk_pm=2, flags=TO)
    check([10000, 2, 3], rnk_pm=2, flags=TI, sign=+1)


def test_long_with_embed_trunc_and_shifts():
    check([2, 8192], ni=[2, 5000], no=[2, 6000])
    check([2, 8192], ni=[2, 5000], no=[2, 6000], flags=SI | SO)
    check([2, 6000], ni=[2, 6000], no=[2, 4100], flags=SO)
    check([2, 6000], ni=[2, 2000], flags=SI)
    check([4, 8192, 2], ni=[4, 7000, 2], no=[2, 3000, 2], flags=SI | SO, rnk_pm=2)


def test_long_fp32_inplace_howmany():
    check([2, 16384], prec="f32")
    check([2, 8192], inplace=True)
    check([2, 6000], howmany=2)


def test_long_adjoint_dot_product():
    n = [3, 8192]
    spec = spec_of("c2c", n, rnk_pm=1)
    rng = np.random.default_rng(4)
    nn = n_scalars(spec, "in")
    x = rng.standard_normal(nn)
    y = rng.standard_normal(nn)
    fx = run_gpu(spec, x)
    fty = run_gpu(spec, y, adjoint=True)
    lhs, rhs = float(fx @ y), float(x @ fty)
    assert abs(lhs - rhs) <= 1e-11 * (abs(lhs) + abs(rhs))


def test_long_in_batch_execution():
    rc, comm = P.create_procmesh([1])
    n = [2, 8192]
    spec = spec_of("c2c", n, rnk_pm=1)
    ns = n_scalars(spec, "in")
    bufs, plans, xs = [], [], []
    for s in range(3):
        a, b = P.DeviceBuffer(ns * 8), P.DeviceBuffer(ns * 8)
        x = random_input(spec, 10 + s)
        a.upload(x)
        pl = P.Plan("c2c", n, a.ptr, b.ptr, comm=comm)
        assert pl.ok
        bufs.append((a, b)); plans.append(pl); xs.append(x)
    P.execute_batch(plans)
    for (a, b), x in zip(bufs, xs):
        assert rel_l2(b.download(np.float64, ns), oracle_out(spec, x)) <= 1e-12
    for pl in plans:
        pl.destroy()
    for a, b in bufs:
        a.free(); b.free()


def check_real(kind, n, flags, sign=-1, rnk_pm=1, prec="f64"):
    spec = spec_of(kind, n, sign=sign, flags=flags, rnk_pm=rnk_pm)
    x = random_input(spec, 7)
    if prec == "f32":
        x = x.astype(np.float32).astype(np.float64)
    got = run_gpu(spec, x, prec=prec)
    want = oracle_out(spec, x)
    err = rel_l2(got, want, out_mask(spec))
    assert err <= (1e-12 if prec == "f64" else 2e-6), (kind, n, flags, err)


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_long_real_lines(prec):
    check_real("r2c", [2, 10000], TO, prec=prec)
    check_real("c2r", [2, 10000], TI, sign=+1, prec=prec)
    check_real("r2c", [3, 4, 9000], TO, rnk_pm=2, prec=prec)
    check_real("c2r", [3, 4, 9000], TI, sign=+1, rnk_pm=2, prec=prec)


def test_long_real_layouts():
    check_real("r2c", [2, 16384], 0)
    check_real("c2r", [2, 16384], 0, sign=+1)
    check_real("r2c", [2, 12000], P.PADDED_R2C)
    check_real("c2r", [2, 12000], P.PADDED_C2R, sign=+1)
    check_real("r2c", [2, 8194], 0, sign=+1)          # conjugated direction, odd half length
    check_real("c2r", [2, 8194], 0, sign=-1)


def test_long_real_adjoint_dot_product():
    for kind, flags, sign in (("r2c", TO, -1), ("c2r", TI, +1)):
        n = [2, 10000]
        spec = spec_of(kind, n, sign=sign, flags=flags, rnk_pm=1)
        rng = np.random.default_rng(8)
        x = rng.standard_normal(n_scalars(spec, "in"))
        y = rng.standard_normal(n_scalars(spec, "out"))
        if kind == "c2r":
            # a Hermitian-consistent input: the c2r of real data's spectrum
            x = oracle_out(spec_of("r2c", n, flags=TO, rnk_pm=1), rng.standard_normal(20000))
        fx = run_gpu(spec, x)
        fty = run_gpu(spec, y, adjoint=True)
        lhs, rhs = float(fx @ y), float(x @ fty)
        assert abs(lhs - rhs) <= 1e-10 * (abs(lhs) + abs(rhs)), (kind, lhs, rhs)


def test_long_rejections():
    c1 = comm_for(1)
    d = P.DeviceBuffer(16 * 2 * 12288)
    assert not P.Plan("c2c", [2, 4099], d.ptr, d.ptr, comm=c1).ok        # prime: no split
    assert not P.Plan("c2c", [2, 2 * 4099], d.ptr, d.ptr, comm=c1).ok    # no split <= 4096
    assert not P.Plan("r2c", [2, 10001], d.ptr, d.ptr, comm=c1).ok       # odd long real line
    assert not P.Plan("c2r", [2, 10001], d.ptr, d.ptr, comm=c1).ok
    assert not P.Plan("r2c", [2, 10000], d.ptr, d.ptr, ni=[2, 9000], comm=c1).ok   # embedded
    assert not P.Plan("r2c", [2, 10000], d.ptr, d.ptr, howmany=2, comm=c1).ok     # strided
    pl = P.Plan("c2c", [2, 4096 * 3], d.ptr, d.ptr, comm=c1)
    assert pl.ok and any(ln.endswith("kernel=long-two-step") for ln in pl.describe())
    pl.destroy()
    d.free()
