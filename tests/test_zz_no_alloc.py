"""pfft_amd_execute_batch_noalloc: the batch entry point that allocates
nothing (device work-list in the library's fixed device array), through
the C ABI, against pfft_execute_batch on the same plans (the same kernels,
grids and work-list, so the results agree bit for bit) and the oracle; two
calls in a row, the second with fewer plans, and the size limit."""
import numpy as np
import pytest

import pfft_amd as P
from gpu_util import spec_of, rel_l2

pytestmark = pytest.mark.gpu

SIDES = [24, 32, 48, 60, 64, 80, 96, 128]
MAX_PASSES = 2048


def make(count, seed, prec="f32", sides=SIDES):
    rng = np.random.default_rng(seed)
    rc, comm = P.create_procmesh([1, 1])
    assert rc == 0
    rdt = np.float32 if prec == "f32" else np.float64
    items = []
    for _ in range(count):
        n = [int(rng.choice(sides)) for _ in range(3)]
        N = int(np.prod(n))
        x = rng.uniform(-1, 1, 2 * N).astype(rdt)
        esz = np.dtype(rdt).itemsize
        din, dout = P.DeviceBuffer(2 * N * esz), P.DeviceBuffer(2 * N * esz)
        plan = P.Plan("c2c", n, din.ptr, dout.ptr, comm=comm, prec=prec)
        assert plan.ok
        din.upload(x)
        items.append((n, x, din, dout, plan))
    return items


def outs(items, rdt=np.float32):
    return [it[3].download(rdt, 2 * int(np.prod(it[0]))) for it in items]


def clear(items, rdt=np.float32):
    for it in items:
        it[3].upload(np.zeros(2 * int(np.prod(it[0])), dtype=rdt))


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_noalloc_matches_execute_batch(prec):
    from oracle import semantics as S
    rdt = np.float32 if prec == "f32" else np.float64
    sides = SIDES if prec == "f32" else [24, 32, 48, 64]
    tol = 2e-6 if prec == "f32" else 1e-12
    for count, seed in ((12, 1), (5, 2)):     # the second call with a smaller batch
        print(f"noalloc {prec} count={count} seed={seed}", flush=True)
        items = make(count, seed, prec, sides)
        plans = [it[4] for it in items]
        P.execute_batch(plans)
        want = outs(items, rdt)
        clear(items, rdt)
        assert P.execute_batch_noalloc(plans) == 0
        got = outs(items, rdt)
        for (n, x, _, _, _), g, w in zip(items, got, want):
            assert np.array_equal(g, w), n
            ref = S.execute_flat(spec_of("c2c", n), x.astype(np.float64), precision=np.float64)
            assert rel_l2(g.astype(np.float64), ref) < tol, n
        for it in items:
            it[4].destroy()


def test_noalloc_rejects_more_than_the_fixed_worklist(capfd):
    items = make(720, 3, sides=[2, 4])
    passes = [len(it[4].describe()) for it in items]
    assert sum(passes) > MAX_PASSES
    # the longest prefix that fits runs; one plan more is refused
    k = int(np.searchsorted(np.cumsum(passes), MAX_PASSES, side="right"))
    assert 0 < k < len(items)
    print(f"noalloc limit: {k} plans, {sum(passes[:k])} passes", flush=True)
    capfd.readouterr()
    assert P.execute_batch_noalloc([it[4] for it in items[:k + 1]]) == 1
    assert "fixed work-list of 2048" in capfd.readouterr().err
    fit = items[:k]
    assert P.execute_batch_noalloc([it[4] for it in fit]) == 0
    for (n, x, _, dout, _) in fit[:20]:
        from oracle import semantics as S
        g = dout.download(np.float32, x.size).astype(np.float64)
        ref = S.execute_flat(spec_of("c2c", n), x.astype(np.float64), precision=np.float64)
        assert rel_l2(g, ref) < 2e-6, n
