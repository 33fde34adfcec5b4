"""GPU parity of ghost-cell plans (pfft_plan_many_rgc / _cgc, pfft_exchange,
pfft_reduce) against the oracle, bit-exact: exchange only copies, and
reduce adds in the reference's order.  Real and complex data, FP64 and
FP32, howmany, ghost widths beyond one period, transposed order on 2-d and
1-d meshes, padded r2c and complex r2c arrays, host arrays, NULL-plan rules
and the timers."""
import ctypes

import numpy as np
import pytest

import pfft_amd as P
from gpu_util import comm_for
from oracle import gcell as G

pytestmark = pytest.mark.gpu

SENTINEL = 5


def run(kind, n, gb, ga, howmany=1, flags=0, rnk_pm=2, prec="f64", seed=0, op="exchange"):
    p = G.plan_params(n, howmany, gb, ga, flags, rnk_pm, kind == "c")
    assert p is not None
    rdt = np.float64 if prec == "f64" else np.float32
    L = p["tuple"] * int(np.prod(p["loc_n"]))
    Gs = p["tuple"] * int(np.prod(p["ngc"]))
    rng = np.random.default_rng(seed)
    buf = np.full(Gs + SENTINEL, 9.0, dtype=rdt)
    if op == "exchange":
        buf[:L] = rng.standard_normal(L)
        want = G.exchange(buf.copy(), p)
    else:
        buf[:Gs] = rng.standard_normal(Gs) * 10 ** rng.uniform(-2, 2, Gs)
        want = G.reduce(buf.copy(), p)
    d = P.DeviceBuffer(buf.nbytes)
    d.upload(buf)
    plan = P.GcPlan(kind, n, d.ptr, gb, ga, howmany=howmany, flags=flags,
                    comm=comm_for(rnk_pm), prec=prec)
    assert plan.ok
    getattr(plan, op)()
    got = d.download(rdt, buf.size)
    plan.destroy()
    d.free()
    assert np.array_equal(got, want), (kind, n, gb, ga, flags, op,
                                       int(np.sum(got != want)))


SHAPES = [
    ([8, 6, 10], [1, 2, 0], [2, 0, 3]),
    ([5, 7, 3], [2, 2, 2], [2, 2, 2]),
    ([3, 2, 4], [7, 0, 9], [4, 5, 0]),        # ghost widths beyond one period
    ([1, 4, 6], [2, 1, 0], [3, 0, 1]),
    ([16, 9, 33], [0, 0, 4], [0, 0, 5]),
]


@pytest.mark.parametrize("n,gb,ga", SHAPES)
@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("op", ["exchange", "reduce"])
def test_real_data(n, gb, ga, prec, op):
    run("r", n, gb, ga, prec=prec, op=op)


@pytest.mark.parametrize("op", ["exchange", "reduce"])
def test_complex_howmany_and_transposed(op):
    run("c", [6, 5, 7], [1, 0, 2], [0, 3, 1], howmany=2, op=op)
    run("r", [6, 5, 7], [1, 2, 2], [3, 0, 1], howmany=3, op=op)
    for rnk_pm in (1, 2):
        run("c", [6, 5, 7], [1, 0, 2], [0, 3, 1], flags=P.GC_TRANSPOSED, rnk_pm=rnk_pm, op=op)
        run("r", [4, 9, 5], [2, 1, 0], [0, 4, 2], flags=P.GC_TRANSPOSED, rnk_pm=rnk_pm,
            prec="f32", op=op)


@pytest.mark.parametrize("op", ["exchange", "reduce"])
def test_r2c_layouts(op):
    # complex r2c arrays (n/2+1 along the last dimension) and padded real arrays
    run("c", [8, 6, 10], [1, 1, 1], [2, 0, 3], flags=P.GC_R2C, op=op)
    run("c", [8, 6, 9], [1, 1, 1], [2, 0, 3], flags=P.GC_PADDED_R2C, op=op)
    run("r", [8, 6, 9], [0, 2, 3], [1, 1, 1], flags=P.GC_PADDED_R2C, op=op)


def test_large_block():
    run("r", [128, 96, 130], [3, 2, 4], [2, 5, 1], op="exchange")
    run("r", [128, 96, 130], [3, 2, 4], [2, 5, 1], op="reduce", prec="f32")
    run("c", [64, 64, 64], [8, 8, 8], [8, 8, 8], flags=P.GC_R2C, op="reduce")


def test_exchange_then_reduce_counts_images():
    n, gb, ga = [5, 4, 6], [2, 5, 1], [3, 0, 7]
    p = G.plan_params(n, 1, gb, ga, 0, 2, False)
    L, Gs = int(np.prod(n)), int(np.prod(p["ngc"]))
    x = np.arange(1, L + 1, dtype=np.float64)
    d = P.DeviceBuffer(Gs * 8)
    d.upload(x)
    plan = P.GcPlan("r", n, d.ptr, gb, ga, comm=comm_for(2))
    plan.exchange()
    plan.reduce()
    got = d.download(np.float64, Gs)
    cnt = [np.bincount((np.arange(p["ngc"][t]) - gb[t]) % n[t], minlength=n[t])
           for t in range(3)]
    want = x.reshape(n) * cnt[0][:, None, None] * cnt[1][None, :, None] * cnt[2][None, None, :]
    assert np.array_equal(got[:L], want.reshape(-1))
    assert np.all(got[L:] == 0)
    ex, red = plan.timers()
    assert ex[0] == 1 and red[0] == 1 and ex[1] > 0 and red[1] > 0
    plan.destroy()
    d.free()


def test_host_array_is_staged():
    n, gb, ga = [6, 4, 5], [1, 2, 1], [2, 1, 0]
    p = G.plan_params(n, 1, gb, ga, 0, 2, False)
    L, Gs = int(np.prod(n)), int(np.prod(p["ngc"]))
    host = np.zeros(Gs, dtype=np.float64)
    host[:L] = np.random.default_rng(3).standard_normal(L)
    want = G.exchange(host.copy(), p)
    plan = P.GcPlan("r", n, host.ctypes.data, gb, ga, comm=comm_for(2))
    assert plan.ok
    plan.exchange()
    assert np.array_equal(host, want)
    want = G.reduce(host.copy(), p)
    plan.reduce()
    assert np.array_equal(host, want)
    plan.destroy()


def test_null_plans_and_timer_output(tmp_path):
    d = P.DeviceBuffer(8 * 4096)
    c2 = comm_for(2)
    assert not P.GcPlan("r", [4, 4, 4], d.ptr, None, None, comm=c2).ok          # no ghost cells
    assert not P.GcPlan("r", [4, 4, 4], d.ptr, [0, 0, 0], [0, 0, 0], comm=c2).ok
    assert not P.GcPlan("r", [4, 4], d.ptr, [1, 1], [1, 1], comm=c2).ok          # rank 2
    assert not P.GcPlan("r", [4, 4, 4], d.ptr, [1, 1, 1], [1, 1, 1], block=[2, 4],
                        comm=c2).ok                                              # partial block
    L = P.lib()
    L.pfft_exchange(None)          # NULL plans do nothing
    L.pfft_reduce(None)
    plan = P.GcPlan("r", [4, 5, 6], d.ptr, [1, 0, 2], [0, 3, 1], comm=c2)
    plan.exchange()
    plan.exchange()
    plan.reduce()
    ex, red = plan.timers()
    assert ex[0] == 2 and red[0] == 1
    f = tmp_path / "gc.m"
    L.pfft_write_average_gctimer_adv(plan.handle, str(f).encode(), c2)
    text = f.read_text()
    assert "n_gc(1, 1:3) = [4 5 6];" in text
    assert "gc_below(1, 1:3) = [1 0 2];" in text and "gc_above(1, 1:3) = [0 3 1];" in text
    assert "gcells_exg_gc_iter(1)      = 2;" in text
    assert "gcells_red_gc_iter(1)      = 1;" in text
    assert "gcells_exg_gc_exchange(1)" in text and "gcells_red_gc_pad_zeros(1)" in text
    L.pfft_reset_gctimers(plan.handle)
    assert plan.timers()[0][0] == 0
    t = L.pfft_get_gctimer_exg(plan.handle)
    t2 = L.pfft_copy_gctimer(t)
    v = (ctypes.c_double * 4)()
    L.pfft_convert_gctimer2vec(t2, v)
    assert list(v) == [0.0, 0.0, 0.0, 0.0]
    L.pfft_destroy_gctimer(t)
    L.pfft_destroy_gctimer(t2)
    plan.destroy()
    d.free()
