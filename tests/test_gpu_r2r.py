"""GPU parity of r2r plans (DHT, REDFT00..11, RODFT00..11) against the
oracle: every kind along every dimension, mixed kinds, transposed
layouts on 2-d and 1-d meshes, oversampled/truncated sizes, howmany,
skipped dimensions, complex64, in-place and new-array execution."""
import numpy as np
import pytest

import pfft_amd as P
from gpu_util import spec_of, random_input, run_gpu, oracle_out, rel_l2, n_scalars, comm_for

pytestmark = pytest.mark.gpu

TI, TO = P.TRANSPOSED_IN, P.TRANSPOSED_OUT
SI, SO = P.SHIFTED_IN, P.SHIFTED_OUT
KINDS = [P.DHT, P.REDFT00, P.REDFT01, P.REDFT10, P.REDFT11,
         P.RODFT00, P.RODFT01, P.RODFT10, P.RODFT11]


def check(n, kinds, ni=None, no=None, howmany=1, flags=0, rnk_pm=2, skip=None, prec="f64",
          inplace=False, seed=2):
    spec = spec_of("r2r", n, ni, no, howmany, 0, flags, rnk_pm, skip, kinds)
    x = random_input(spec, seed)
    if prec == "f32":
        x = x.astype(np.float32).astype(np.float64)
    got = run_gpu(spec, x, prec=prec, inplace=inplace)
    want = oracle_out(spec, x)
    tol = 1e-12 if prec == "f64" else 2e-6
    err = rel_l2(got, want)
    assert err <= tol, (n, kinds, flags, err)


@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("n", [[8, 6, 12], [5, 7, 9], [2, 3, 16]])
def test_single_kind_all_dims(kind, n):
    check(n, [kind] * 3)


@pytest.mark.parametrize("kind", KINDS)
def test_kind_along_each_dim(kind):
    for d in range(3):
        kinds = [P.REDFT10, P.RODFT10, P.DHT]
        kinds[d] = kind
        check([12, 10, 14], kinds)


@pytest.mark.parametrize("flags", [0, TO, TI])
@pytest.mark.parametrize("rnk_pm", [1, 2])
def test_layouts(flags, rnk_pm):
    check([10, 12, 8], [P.REDFT10, P.RODFT11, P.REDFT00], flags=flags, rnk_pm=rnk_pm)


def test_ousam_and_howmany():
    check([16, 12, 10], [P.REDFT10, P.RODFT00, P.DHT], ni=[8, 12, 6], no=[16, 6, 10])
    check([8, 6, 10], [P.REDFT01, P.RODFT10, P.REDFT11], howmany=3, flags=TO)


def test_skipped_and_4d():
    check([8, 6, 10], [P.REDFT10, P.RODFT10, P.DHT], skip=[0, 1, 0])
    check([6, 5, 4, 7], [P.REDFT00, P.RODFT01, P.DHT, P.REDFT11])


def test_long_lines_and_fp32():
    check([4, 1024, 3], [P.DHT, P.REDFT10, P.RODFT00])
    check([3, 2, 500], [P.RODFT11, P.DHT, P.REDFT11])
    check([16, 64, 32], [P.REDFT10, P.RODFT10, P.REDFT01], prec="f32")


def test_inplace_and_new_arrays():
    check([12, 10, 8], [P.REDFT10, P.RODFT00, P.DHT], inplace=True)
    n = [8, 16, 12]
    kinds = [P.RODFT10, P.REDFT01, P.DHT]
    spec = spec_of("r2r", n, sign=0, kinds=kinds)
    ns = n_scalars(spec, "in")
    a, b, c = P.DeviceBuffer(ns * 8), P.DeviceBuffer(ns * 8), P.DeviceBuffer(ns * 8)
    pl = P.Plan("r2r", n, a.ptr, b.ptr, comm=comm_for(2), kinds=kinds)
    assert pl.ok
    x = random_input(spec, 9)
    c.upload(x)
    pl.execute_on(c.ptr, a.ptr)
    assert rel_l2(a.download(np.float64, ns), oracle_out(spec, x)) <= 1e-12
    pl.destroy()
    for buf in (a, b, c):
        buf.free()


@pytest.mark.parametrize("flags", [P.SHIFTED_IN, P.SHIFTED_OUT, P.SHIFTED_IN | P.SHIFTED_OUT,
                                   P.SHIFTED_IN | P.SHIFTED_OUT | TO])
def test_shifted(flags):
    """The +-1 modulations and the centred embed/trunc of the reference
    (api/api-basic.c:1186-1285, kernel/ousample.c:241-322) on real lines."""
    check([8, 6, 10], [P.REDFT10, P.RODFT00, P.DHT], flags=flags)
    check([12, 10, 6], [P.RODFT11, P.REDFT01, P.REDFT00], ni=[6, 10, 4], no=[8, 4, 6],
          flags=flags)


def test_describe_names_r2r_passes():
    rc, comm = P.create_procmesh([1, 1])
    d = P.DeviceBuffer(8 * 4096)
    pl = P.Plan("r2r", [4, 8, 16], d.ptr, d.ptr, comm=comm, kinds=[P.REDFT10, P.DHT, P.RODFT00])
    assert pl.ok
    lines = pl.describe()
    assert len(lines) == 3 and all(": r2r n=" in ln for ln in lines), lines
    pl.destroy()
    d.free()


def test_rejections():
    rc, comm = P.create_procmesh([1, 1])
    d = P.DeviceBuffer(8 * 1024)
    assert not P.Plan("r2r", [4, 4, 4], d.ptr, d.ptr, comm=comm, kinds=[0, 3, 3]).ok   # R2HC
    assert not P.Plan("r2r", [4, 4, 5], d.ptr, d.ptr, comm=comm, kinds=[3, 3, 3],
                      flags=P.SHIFTED_IN | P.SHIFTED_OUT).ok                         # odd n
    assert not P.Plan("r2r", [4, 4, 4099], d.ptr, d.ptr, comm=comm, kinds=[3, 3, 2]).ok  # no split
    d.free()


@pytest.mark.parametrize("kind,n", [(P.REDFT10, 2048), (P.REDFT01, 1500), (P.REDFT11, 1200),
                                    (P.RODFT10, 1100), (P.RODFT01, 2048), (P.RODFT11, 640),
                                    (P.REDFT00, 5001), (P.RODFT00, 4999), (P.DHT, 6000)])
def test_long_extended_lines(kind, n):
    """Extended length above one workgroup's line: the extension is built
    in a work buffer and transformed in two steps."""
    check([2, n, 2], [P.DHT, kind, P.REDFT10], rnk_pm=1)


def test_long_extended_lines_layouts():
    check([3, 2048, 2], [P.REDFT10, P.REDFT10, P.RODFT00], ni=[3, 1500, 2], no=[2, 1800, 2])
    check([4, 1024, 2], [P.RODFT11, P.REDFT11, P.DHT], flags=SO | SI)
    check([2048, 3, 2], [P.REDFT10, P.DHT, P.RODFT10], flags=TO)
    check([2, 2, 1024], [P.DHT, P.REDFT01, P.REDFT11], prec="f32")


def test_shortest_lines():
    check([2, 1, 3], [P.REDFT00, P.RODFT00, P.DHT])
    check([1, 2, 1], [P.REDFT11, P.RODFT01, P.REDFT10])
    check([1, 1, 2], [P.RODFT10, P.REDFT01, P.RODFT11])
