"""In-place plans (one array for input and output) on the GPU, against the
oracle, and the reference's in-place round-trip tests restated at their
sizes: tests/simple_check_c2c_inplace.c and
tests/simple_check_c2c_transposed_inplace.c (n = 29 x 27 x 31, forward
with PFFT_DESTROY_INPUT, backward, scale by 1/N, pfft_check_output_*
below 1e-12 on arrays filled by pfft_init_input_*)."""
import ctypes

import numpy as np
import pytest

import pfft_amd as P
from gpu_util import spec_of, random_input, run_gpu, oracle_out, out_mask, rel_l2, n_scalars

pytestmark = pytest.mark.gpu

TI, TO, PAD = P.TRANSPOSED_IN, P.TRANSPOSED_OUT, P.PADDED_R2C
DESTROY = 1 << 9


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("kind,n,flags", [
    ("c2c", [29, 27, 31], 0), ("c2c", [29, 27, 31], TO), ("c2c", [29, 27, 31], TI),
    ("c2c", [64, 128, 256], TO), ("c2c", [256, 64, 128], TI), ("c2c", [16, 12, 10, 8], TO),
    ("r2c", [16, 12, 20], PAD), ("r2c", [16, 12, 20], TO | PAD), ("r2c", [128, 64, 256], TO | PAD),
    ("r2c", [16, 12, 20], 0), ("r2c", [64, 64, 512], TO),
    ("c2r", [16, 12, 20], PAD), ("c2r", [16, 12, 20], TI | PAD), ("c2r", [128, 64, 256], TI | PAD),
    ("c2r", [16, 12, 20], 0), ("c2r", [64, 64, 512], TI),
])
def test_inplace_parity(kind, n, flags, prec):
    sign = +1 if kind == "c2r" else -1
    spec = spec_of(kind, n, sign=sign, flags=flags | DESTROY)
    x = random_input(spec, 21)
    if prec == "f32":
        x = x.astype(np.float32).astype(np.float64)
    got = run_gpu(spec, x, prec=prec, inplace=True)
    want = oracle_out(spec, x)
    tol = 1e-12 if prec == "f64" else 2e-6
    assert rel_l2(got, want, out_mask(spec)) <= tol


@pytest.mark.parametrize("transposed", [False, True])
def test_reference_c2c_inplace_roundtrip(transposed):
    n = [29, 27, 31]
    L = P.lib()
    rc, comm = P.create_procmesh([1, 1])
    assert rc == 0
    fflag = TO if transposed else 0
    bflag = TI if transposed else 0
    alloc, lni, lis, lno, los = P.local_size("c2c", n, flags=fflag, comm=comm)
    buf = L.pfft_alloc_complex(alloc)
    assert buf
    fw = P.Plan("c2c", n, buf, buf, sign=P.FORWARD, flags=fflag | DESTROY, comm=comm)
    bw = P.Plan("c2c", n, buf, buf, sign=P.BACKWARD, flags=bflag | DESTROY, comm=comm)
    assert fw.ok and bw.ok
    L.pfft_init_input_complex(3, P.ivec(n), P.ivec(lni), P.ivec(lis), buf)
    fw.execute()
    bw.execute()
    N = int(np.prod(n))
    y = np.empty(2 * N)
    assert L.pfft_amd_memcpy_to_host(y.ctypes.data, buf, y.nbytes) == 0
    y /= N
    assert L.pfft_amd_memcpy_to_device(buf, y.ctypes.data, y.nbytes) == 0
    err = L.pfft_check_output_complex(3, P.ivec(n), P.ivec(lni), P.ivec(lis), buf, comm)
    assert err < 1e-12, err
    fw.destroy()
    bw.destroy()
    L.pfft_free(buf)


def test_new_array_execute_inplace():
    """execute_dft on arrays other than the planned ones, in place like the
    plan (tests/simple_check_*_newarray.c)."""
    n = [32, 16, 64]
    spec = spec_of("c2c", n, flags=TO | DESTROY)
    nsc = n_scalars(spec, "in")
    a = P.DeviceBuffer(nsc * 8)
    b = P.DeviceBuffer(nsc * 8)
    pl = P.Plan("c2c", n, a.ptr, a.ptr, flags=TO | DESTROY, comm=P.create_procmesh([1, 1])[1])
    assert pl.ok
    x = random_input(spec, 4)
    b.upload(x)
    pl.execute_on(b.ptr, b.ptr)
    got = b.download(np.float64, nsc)
    assert rel_l2(got, oracle_out(spec, x)) <= 1e-12
    pl.destroy()
    a.free()
    b.free()
