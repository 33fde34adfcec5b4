"""GPU adjoint checks.

* dot-product test <A x, y> = <x, A* y> (real inner product on the stored
  scalars) with A from the oracle and A* from the library's hand-written
  adjoint kernels, across transform kinds and flags;
* configuration 5 (oversampled, shifted r2c, ni=128^3 -> n=256^3):
  gradient of L = sum_k w_k |X_k|^2, w_k = exp(-|k|^2 / (2*40^2)), from
  the torch.autograd.Function over the plan, checked against central
  finite differences on 64 random directions and against a gradient
  computed independently with numpy's FP64 FFT.
"""
import time

import numpy as np
import pytest

import pfft_amd as P
from gpu_util import spec_of, random_input, run_gpu, oracle_out, n_scalars, out_mask

pytestmark = pytest.mark.gpu

TI, TO, SI, SO = P.TRANSPOSED_IN, P.TRANSPOSED_OUT, P.SHIFTED_IN, P.SHIFTED_OUT
PAD = P.PADDED_R2C


def in_mask(spec):
    m = np.ones(n_scalars(spec, "in"), dtype=bool)
    if spec.kind == "r2c" and (spec.tf & (1 << 13)):
        shp = spec.in_memory_shape()
        a = np.zeros(shp, dtype=bool)
        sl = [slice(None)] * len(shp)
        sl[-2] = slice(0, spec.ni[-1])
        a[tuple(sl)] = True
        m = a.reshape(-1)
    return m


def dot_test(kind, n, ni=None, no=None, flags=0, sign=-1, howmany=1, rnk_pm=2, skip=None,
             prec="f64"):
    spec = spec_of(kind, n, ni, no, howmany, sign, flags, rnk_pm, skip)
    x = random_input(spec, 11) * in_mask(spec)
    rng = np.random.default_rng(12)
    y = rng.uniform(-1, 1, n_scalars(spec, "out")) * out_mask(spec)
    Ax = oracle_out(spec, x)
    Aty = run_gpu(spec, y, prec=prec, adjoint=True)
    lhs = float(np.dot(Ax[out_mask(spec)], y[out_mask(spec)]))
    rhs = float(np.dot(x[in_mask(spec)], Aty[in_mask(spec)]))
    tol = 1e-12 if prec == "f64" else 2e-5
    assert abs(lhs - rhs) <= tol * max(abs(lhs), np.linalg.norm(Ax) * np.linalg.norm(y)), \
        (kind, n, flags, lhs, rhs)


@pytest.mark.parametrize("flags", [0, TO, TI, SI | SO, SI, SO])
@pytest.mark.parametrize("n", [[16, 12, 10], [64, 32, 128], [29, 27, 31]])
def test_adjoint_c2c(n, flags):
    if flags & (SI | SO) and any(v % 2 for v in n):
        pytest.skip("shifted plans need even sizes")
    dot_test("c2c", n, flags=flags)
    dot_test("c2c", n, flags=flags, sign=+1)


@pytest.mark.parametrize("kind,flags", [("r2c", 0), ("r2c", TO), ("r2c", PAD), ("r2c", SI | SO),
                                        ("c2r", 0), ("c2r", TI), ("c2r", PAD), ("c2r", SI | SO)])
@pytest.mark.parametrize("n", [[16, 12, 10], [32, 64, 256], [8, 6, 9]])
def test_adjoint_real(kind, flags, n):
    if flags & (SI | SO) and any(v % 2 for v in n):
        pytest.skip("shifted plans need even sizes")
    sign = -1 if kind == "r2c" else +1
    dot_test(kind, n, flags=flags, sign=sign)


@pytest.mark.parametrize("kind", ["c2c", "r2c", "c2r"])
def test_adjoint_ousam(kind):
    sign = +1 if kind == "c2r" else -1
    dot_test(kind, [29, 27, 31], ni=[16, 16, 16], no=[16, 16, 16], sign=sign)
    dot_test(kind, [64, 64, 64], ni=[32, 32, 32], no=[48, 40, 32], sign=sign, flags=SI | SO)


def test_adjoint_howmany_skip_fp32():
    dot_test("c2c", [8, 10, 12], howmany=2, flags=TO)
    dot_test("c2c", [8, 10, 12], skip=[0, 1, 0])
    dot_test("r2c", [64, 32, 128], flags=TO, prec="f32")
    dot_test("c2r", [64, 32, 128], flags=TI, sign=+1, prec="f32")


# ---------------------------------------------------------------- configuration 5
def _cfg5_setup(prec="f64", ni=128, n=256):
    import torch
    from pfft_amd.autograd import apply_plan
    rc, comm = P.create_procmesh([1, 1])
    flags = SI | SO
    rdt = torch.float64 if prec == "f64" else torch.float32
    dev = torch.device("cuda:0")
    x = torch.empty((ni, ni, ni), dtype=rdt, device=dev)
    X = torch.empty((n, n, n // 2 + 1), dtype=torch.complex128 if prec == "f64"
                    else torch.complex64, device=dev)
    plan = P.Plan("r2c", [n] * 3, x.data_ptr(), X.data_ptr(), ni=[ni] * 3, flags=flags,
                  comm=comm, prec=prec)
    assert plan.ok
    # centred wave numbers of the output: k0, k1 in [-n/2, n/2), k2 in [-n/2, 0]
    k01 = torch.arange(-n // 2, n // 2, device=dev, dtype=torch.float64)
    k2 = torch.arange(-n // 2, 1, device=dev, dtype=torch.float64)
    kk = k01[:, None, None] ** 2 + k01[None, :, None] ** 2 + k2[None, None, :] ** 2
    w = torch.exp(-kk / (2 * 40.0 ** 2)).to(rdt)

    def loss(xin):
        Xo = apply_plan(xin, plan, (n, n, n // 2 + 1), True)
        return (w * (Xo.real ** 2 + Xo.imag ** 2)).double().sum()
    return plan, loss, w


def _numpy_reference_gradient(x, ni, n, sigma=40.0):
    """Independent FP64 gradient of L(x) = sum_k w_k |X_k|^2 over the
    stored half spectrum k2 in [-n/2, 0], X = centred DFT of x embedded at
    [-ni/2, ni/2): grad_j = 2 Re sum_k w_k conj(X_k) exp(-2 pi i j.k/n)."""
    xe = np.zeros((n, n, n))
    o = (n - ni) // 2
    xe[o:o + ni, o:o + ni, o:o + ni] = x
    # centred DFT via fftshift conventions: index j_c = j - n/2
    Xf = np.fft.fftshift(np.fft.fftn(np.fft.ifftshift(xe)))
    k = np.arange(-n // 2, n // 2)
    kk = k[:, None, None] ** 2 + k[None, :, None] ** 2 + k[None, None, :] ** 2
    w = np.exp(-kk / (2 * sigma ** 2))
    half = (k <= 0)[None, None, :]        # stored modes k2 in [-n/2, 0]
    Y = np.where(half, w * Xf, 0)
    # sum_k conj(Y_k) e^{-2 pi i j k/n} = conj( sum_k Y_k e^{+2 pi i j k / n} )
    g = np.fft.fftshift(np.fft.ifftn(np.fft.ifftshift(Y))) * n ** 3
    g = 2 * np.real(np.conj(g))
    return g[o:o + ni, o:o + ni, o:o + ni]


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_config5_gradient(prec):
    import torch
    ni, n = 128, 256
    plan, loss, w = _cfg5_setup(prec, ni, n)
    rdt = torch.float64 if prec == "f64" else torch.float32
    g = torch.Generator(device="cpu").manual_seed(20261202)
    x0 = torch.randn((ni, ni, ni), generator=g, dtype=torch.float64).to(rdt).cuda()
    x = x0.clone().requires_grad_(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    L = loss(x)
    L.backward()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    grad = x.grad.detach()
    print("config5 %s: forward+backward %.2f ms, L = %.6e" % (prec, (t1 - t0) * 1e3, L.item()))
    # independent reference gradient (numpy FP64)
    gref = _numpy_reference_gradient(x0.double().cpu().numpy(), ni, n)
    gg = grad.double().cpu().numpy()
    rel = np.linalg.norm(gg - gref) / np.linalg.norm(gref)
    tol_ref = 1e-11 if prec == "f64" else 2e-5
    assert rel < tol_ref, rel
    # central finite differences on 64 random directions
    gen = torch.Generator(device="cpu").manual_seed(7)
    worst = 0.0
    with torch.no_grad():
        for i in range(64):
            d = torch.randn((ni, ni, ni), generator=gen, dtype=torch.float64)
            d = (d / d.norm()).to(rdt).cuda()
            # L is quadratic, so central differences are exact up to rounding
            # for any step; a large step keeps rounding small against the
            # difference
            eps = 1.0 if prec == "f64" else 30.0
            lp = loss(x0 + eps * d).double().item()
            lm = loss(x0 - eps * d).double().item()
            fd = (lp - lm) / (2 * eps)
            an = float((grad.double() * d.double()).sum().item())
            worst = max(worst, abs(fd - an) / max(abs(an), 1e-30 + abs(fd)))
    tol_fd = 1e-8 if prec == "f64" else 1e-2
    assert worst < tol_fd, worst
    plan.destroy()
