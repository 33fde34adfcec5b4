"""GPU parity: the HIP path (through the pfft.h C ABI) against the oracle
on the same seeded inputs.  FP64 bar: relative L2 <= 1e-12; FP32 bar:
relative L2 <= 2e-6 (single-precision Stockham, inputs of O(1))."""
import os

import numpy as np
import pytest

import pfft_amd as P
from oracle import layout as L
from oracle import semantics as S
from gpu_util import spec_of, random_input, run_gpu, oracle_out, out_mask, rel_l2

pytestmark = pytest.mark.gpu

TI, TO, SI, SO = P.TRANSPOSED_IN, P.TRANSPOSED_OUT, P.SHIFTED_IN, P.SHIFTED_OUT
PAD = P.PADDED_R2C

TOL64 = 1e-12
TOL32 = 2e-6


def check(kind, n, ni=None, no=None, howmany=1, sign=-1, flags=0, rnk_pm=2, skip=None,
          prec="f64", seed=1):
    spec = spec_of(kind, n, ni, no, howmany, sign, flags, rnk_pm, skip)
    x = random_input(spec, seed)
    if prec == "f32":
        x = x.astype(np.float32).astype(np.float64)
    got = run_gpu(spec, x, prec=prec)
    want = oracle_out(spec, x)
    err = rel_l2(got, want, out_mask(spec))
    tol = TOL64 if prec == "f64" else TOL32
    assert err <= tol, "rel L2 %.3e > %.1e for %s n=%s ni=%s no=%s flags=%#x" % (
        err, tol, kind, n, ni, no, flags)
    return err


# the configuration-1 shape: 64^3 c2c, default layout, both directions
@pytest.mark.parametrize("sign", [-1, +1])
def test_c2c_64cube_default(sign):
    check("c2c", [64, 64, 64], sign=sign)


@pytest.mark.parametrize("flags", [0, TO, TI])
@pytest.mark.parametrize("rnk_pm", [1, 2])
def test_c2c_odd_sizes_layouts(flags, rnk_pm):
    # the reference tests' deliberately uneven sizes (tests/simple_check_c2c.c)
    check("c2c", [29, 27, 31], flags=flags, rnk_pm=rnk_pm)


@pytest.mark.parametrize("kind,flags", [("r2c", 0), ("r2c", TO), ("r2c", PAD), ("r2c", TO | PAD),
                                        ("c2r", 0), ("c2r", TI), ("c2r", PAD), ("c2r", TI | PAD)])
@pytest.mark.parametrize("n", [[29, 27, 31], [16, 12, 20], [8, 6, 10]])
def test_real_transforms(kind, flags, n):
    sign = -1 if kind == "r2c" else +1
    check(kind, n, sign=sign, flags=flags)


@pytest.mark.parametrize("kind,sign", [("r2c", +1), ("c2r", -1)])
def test_conjugated_directions(kind, sign):
    # c2r FORWARD and r2c BACKWARD (kernel/partrafo.c:426-443)
    check(kind, [12, 10, 14], sign=sign)


@pytest.mark.parametrize("kind", ["c2c", "r2c", "c2r"])
@pytest.mark.parametrize("flags", [0, TO, TI])
def test_ousam_reference_sizes(kind, flags):
    # tests/simple_check_ousam_*.c: ni=16 -> n=(29,27,31) -> no=16
    if (kind == "r2c" and flags == TI) or (kind == "c2r" and flags == TO):
        pytest.skip("invalid combination (NULL plan)")
    sign = +1 if kind == "c2r" else -1
    check(kind, [29, 27, 31], ni=[16, 16, 16], no=[16, 16, 16], sign=sign, flags=flags)


@pytest.mark.parametrize("shift", [SI, SO, SI | SO])
@pytest.mark.parametrize("kind", ["c2c", "r2c", "c2r"])
def test_shifted_small(kind, shift):
    # tests/simple_check_c2r_c2c_shifted.c size 4x4x6
    sign = +1 if kind == "c2r" else -1
    check(kind, [4, 4, 6], sign=sign, flags=shift)


@pytest.mark.parametrize("shift", [SI, SO, SI | SO])
@pytest.mark.parametrize("kind", ["c2c", "c2r", "r2c"])
def test_shifted_ousam(kind, shift):
    # tests/simple_check_c2r_c2c_ousam_shifted.c: ni=4 -> n=6
    sign = +1 if kind == "c2r" else -1
    check(kind, [6, 6, 6], ni=[4, 4, 4], no=[4, 4, 4], sign=sign, flags=shift)


@pytest.mark.parametrize("flags", [SI | SO | TO, SI | SO])
def test_shifted_oversampled_r2c_config5_small(flags):
    # configuration-5 geometry at a reduced size: ni=16^3 -> n=32^3, both shifted
    check("r2c", [32, 32, 32], ni=[16, 16, 16], flags=flags)


@pytest.mark.parametrize("howmany", [2, 3])
@pytest.mark.parametrize("kind,flags", [("c2c", 0), ("c2c", TO), ("r2c", TO), ("c2r", TI)])
def test_howmany(kind, flags, howmany):
    sign = +1 if kind == "c2r" else -1
    check(kind, [10, 12, 8], howmany=howmany, sign=sign, flags=flags)


@pytest.mark.parametrize("flags", [0, TO, TI])
def test_4d(flags):
    # tests/simple_check_c2c_4d.c uses 13^4-like uneven sizes
    check("c2c", [13, 11, 9, 7], flags=flags)


def test_2d_on_1d_mesh():
    check("c2c", [24, 20], flags=TO, rnk_pm=1)
    check("r2c", [24, 20], flags=TO, rnk_pm=1)


@pytest.mark.parametrize("skip", [[1, 0, 0], [0, 1, 0], [0, 0, 1], [1, 1, 0]])
def test_skipped_c2c(skip):
    check("c2c", [8, 10, 12], skip=skip)
    check("c2c", [8, 10, 12], skip=skip, flags=TO)


@pytest.mark.parametrize("kind,flags", [("c2c", 0), ("c2c", TO), ("r2c", TO), ("c2r", TI)])
def test_fp32(kind, flags):
    sign = +1 if kind == "c2r" else -1
    check(kind, [32, 24, 40], sign=sign, flags=flags, prec="f32")


@pytest.mark.parametrize("n", [[128, 128, 128], [96, 80, 60], [256, 64, 32]])
def test_medium_r2c_c2r(n):
    check("r2c", n, flags=TO)
    check("c2r", n, sign=+1, flags=TI)


def test_roundtrip_512_r2c_c2r():
    """Configuration 2 at full size: r2c (TRANSPOSED_OUT) then c2r
    (TRANSPOSED_IN) returns N * input (size-independent property), two
    lines of the spectrum match direct DFTs of plane sums, and the whole
    forward spectrum matches the FP64 oracle to 1e-12 relative L2."""
    n = [512, 512, 512]
    N = 512 ** 3
    rc, comm = P.create_procmesh([1, 1])
    mem, lni, lis, lno, los = P.local_size("r2c", n, flags=TO, comm=comm)
    assert mem == 67371008 and lno == [512, 512, 257]
    rng = np.random.default_rng(20261202)
    x = rng.standard_normal(N)
    dx = P.DeviceBuffer(N * 8)
    dk = P.DeviceBuffer(mem * 16)
    dy = P.DeviceBuffer(2 * mem * 8)
    fw = P.Plan("r2c", n, dx.ptr, dk.ptr, flags=TO, comm=comm)
    bw = P.Plan("c2r", n, dk.ptr, dy.ptr, sign=+1, flags=TI, comm=comm)
    assert fw.ok and bw.ok
    dx.upload(x)
    fw.execute()
    k = dk.download(np.float64, 2 * mem)
    bw.execute()
    y = dy.download(np.float64, N)
    err = np.max(np.abs(y / N - x))
    assert err < 1e-12 * 50, err
    # spectrum layout [k1][k2][k0]: the k1=0,k2=0 line is the DFT along x0
    # of the plane sums over (x1, x2)
    line = k[: 2 * 512].reshape(512, 2)
    want = np.fft.fft(x.reshape(512, 512, 512).sum(axis=(1, 2)))
    got = line[:, 0] + 1j * line[:, 1]
    assert rel_l2(np.concatenate([got.real, got.imag]),
                  np.concatenate([want.real, want.imag])) < 1e-12
    # k2 = 256 (Nyquist) plane: DFT over x0 of alternating sums
    sgn = (-1.0) ** np.arange(512)
    want2 = np.fft.fft((x.reshape(512, 512, 512) * sgn).sum(axis=(1, 2)))
    off = (0 * 257 + 256) * 512
    got2 = k[2 * off: 2 * off + 1024].reshape(512, 2)
    got2 = got2[:, 0] + 1j * got2[:, 1]
    assert np.linalg.norm(got2 - want2) / np.linalg.norm(want2) < 1e-12
    # the whole forward spectrum against the oracle in FP64 (on this GPU's
    # share of the host threads)
    workers = S.WORKERS
    S.WORKERS = min(16, len(os.sched_getaffinity(0)))
    try:
        want_k = S.execute_flat(spec_of("r2c", n, flags=TO), x, precision=np.float64)
    finally:
        S.WORKERS = workers
    assert rel_l2(k, want_k) < 1e-12
    for o in (fw, bw):
        o.destroy()
    for b in (dx, dk, dy):
        b.free()


# ---- power-of-two shapes: these run the register-resident kernels ----
@pytest.mark.parametrize("flags", [0, TO, TI])
@pytest.mark.parametrize("rnk_pm", [1, 2])
@pytest.mark.parametrize("n", [[64, 128, 256], [512, 64, 128], [128, 1024, 64]])
def test_pow2_c2c_layouts(n, flags, rnk_pm):
    check("c2c", n, flags=flags, rnk_pm=rnk_pm)
    check("c2c", n, flags=flags, rnk_pm=rnk_pm, sign=+1)


@pytest.mark.parametrize("kind,flags", [("r2c", 0), ("r2c", TO), ("r2c", TO | PAD),
                                        ("c2r", 0), ("c2r", TI), ("c2r", TI | PAD)])
@pytest.mark.parametrize("n", [[64, 32, 256], [128, 128, 512], [32, 64, 2048]])
def test_pow2_real(kind, flags, n):
    sign = -1 if kind == "r2c" else +1
    check(kind, n, sign=sign, flags=flags)
    check(kind, n, sign=-sign, flags=flags)


@pytest.mark.parametrize("kind", ["c2c", "r2c", "c2r"])
@pytest.mark.parametrize("flags", [SI | SO, SI, SO, SI | SO | TO, SI | SO | TI])
def test_pow2_shift_ousam(kind, flags):
    # configuration-5 geometry scaled down: ni=64^3 embedded in n=128^3
    if (kind == "r2c" and flags & TI) or (kind == "c2r" and flags & TO):
        pytest.skip("invalid combination (NULL plan)")
    sign = +1 if kind == "c2r" else -1
    check(kind, [128, 128, 128], ni=[64, 64, 64], no=[128, 128, 128], sign=sign, flags=flags)
    check(kind, [128, 128, 128], ni=[128, 128, 128], no=[64, 64, 64], sign=sign, flags=flags)


@pytest.mark.parametrize("kind,flags", [("c2c", TO), ("r2c", TO), ("c2r", TI)])
def test_pow2_howmany_fp32(kind, flags):
    sign = +1 if kind == "c2r" else -1
    check(kind, [64, 128, 64], howmany=2, sign=sign, flags=flags)
    check(kind, [64, 128, 64], sign=sign, flags=flags, prec="f32")


# ---- long power-of-two lines: 16 elements per thread, radix-16 stages ----
@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("n,flags", [([1024, 8, 16], 0), ([16, 1024, 8], TO), ([8, 16, 2048], TI),
                                     ([4096, 4, 8], TO), ([4, 8, 4096], 0)])
def test_long_pow2_c2c(n, flags, prec):
    check("c2c", n, flags=flags, prec=prec)
    check("c2c", n, flags=flags, prec=prec, sign=+1)


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("n", [[8, 4, 2048], [4, 4, 8192], [16, 2, 4096]])
def test_long_pow2_real(n, prec):
    check("r2c", n, flags=TO, prec=prec)
    check("c2r", n, flags=TI, sign=+1, prec=prec)


# ---- complex64 MFMA arithmetic forms (radix-16 stages on the matrix cores) ----
@pytest.mark.parametrize("form,tol", [(P.AMD_MFMA_F32, 5e-6), (P.AMD_MFMA_BF16X2, 1e-5)])
@pytest.mark.parametrize("n,flags", [([512, 64, 256], 0), ([256, 512, 32], TO),
                                     ([32, 256, 1024], TI), ([16, 2048, 8], 0),
                                     ([4096, 4, 8], TO), ([40, 1024, 24], 0)])
def test_mfma_forms(n, flags, form, tol):
    for sign in (-1, +1):
        spec = spec_of("c2c", n, flags=flags, sign=sign)
        x = random_input(spec, 5).astype(np.float32).astype(np.float64)
        got = run_gpu(spec, x, prec="f32", plan_flags_extra=form)
        want = oracle_out(spec, x)
        err = rel_l2(got, want)
        assert err <= tol, (n, flags, form, sign, err)


@pytest.mark.parametrize("form,tol", [(P.AMD_MFMA_F32, 5e-6), (P.AMD_MFMA_BF16X2, 1e-5)])
@pytest.mark.parametrize("kind,flags", [("r2c", TO), ("c2r", TI)])
def test_mfma_forms_real_pair(kind, flags, form, tol):
    """the headline pair's shape (r2c with transposed output, c2r with
    transposed input) in the MFMA forms at a reduced size: the column c2c
    pass runs an MFMA kernel, the transposing c2c pass and the real pass the
    vector-ALU kernels; both directions against the oracle"""
    n = [256, 256, 128]
    spec = spec_of(kind, n, flags=flags, sign=(+1 if kind == "c2r" else -1))
    x = random_input(spec, 9).astype(np.float32).astype(np.float64)
    got = run_gpu(spec, x, prec="f32", plan_flags_extra=form)
    want = oracle_out(spec, x)
    assert rel_l2(got, want, out_mask(spec)) <= tol


@pytest.mark.parametrize("n", [[1024, 16, 512], [256, 32, 1024], [512, 8, 256]])
def test_mfma_forms_precision_input(n):
    """the precision configuration's input (complex64 U[-1,1)^2 plus a
    1e-3-RMS 1/|k|^2 component) at reduced sizes whose passes run the
    chained MFMA kernels (strided and contiguous 256/512/1024-point
    lines): forward and backward against the oracle (FP64) at the 1e-5
    relative-L2 bar, FP32-MFMA tighter."""
    import torch
    import bench_configs
    x = bench_configs.cfg3_input(n, 11, torch.device("cuda:0")).cpu().numpy()
    flat = np.empty(2 * x.size)
    flat[0::2], flat[1::2] = x.real.ravel(), x.imag.ravel()
    workers = S.WORKERS
    S.WORKERS = min(16, len(os.sched_getaffinity(0)))
    try:
        for sign in (-1, +1):
            spec = spec_of("c2c", n, sign=sign)
            want = S.execute_flat(spec, flat, precision=np.float64)
            for form, tol in ((P.AMD_MFMA_F32, 1e-6), (P.AMD_MFMA_BF16X2, 1e-5)):
                got = run_gpu(spec, flat, prec="f32", plan_flags_extra=form)
                assert rel_l2(got, want) <= tol, (n, sign, form, rel_l2(got, want))
    finally:
        S.WORKERS = workers


def test_mfma_form_rejected_for_fp64():
    import pfft_amd as PP
    rc, comm = PP.create_procmesh([1, 1])
    d = PP.DeviceBuffer(16 * 64)
    p = PP.Plan("c2c", [4, 4, 4], d.ptr, d.ptr, flags=PP.AMD_MFMA_F32, comm=comm, prec="f64")
    assert not p.ok
