"""Batched execution (configuration 4 shape): many 3-d c2c complex64
transforms of uneven sizes (sides from {24,32,48,60,64,80,96,128}) run as
one persistent work-list launch, captured into a hipGraph and replayed."""
import os

import numpy as np
import pytest

import pfft_amd as P
from gpu_util import spec_of, oracle_out, rel_l2

pytestmark = pytest.mark.gpu

SIDES = [24, 32, 48, 60, 64, 80, 96, 128]


def make_batch(count, seed, prec="f32", max_side=128, flags=0):
    rng = np.random.default_rng(seed)
    rc, comm = P.create_procmesh([1, 1])
    sides = [s for s in SIDES if s <= max_side]
    items = []
    for _ in range(count):
        n = [int(rng.choice(sides)) for _ in range(3)]
        N = int(np.prod(n))
        rdt = np.float32 if prec == "f32" else np.float64
        x = rng.uniform(-1, 1, 2 * N).astype(rdt)
        esz = np.dtype(rdt).itemsize
        din = P.DeviceBuffer(2 * N * esz)
        dout = P.DeviceBuffer(2 * N * esz)
        plan = P.Plan("c2c", n, din.ptr, dout.ptr, flags=flags, comm=comm, prec=prec)
        assert plan.ok
        din.upload(x)
        items.append((n, x, din, dout, plan))
    return items


def outputs(items, prec="f32"):
    rdt = np.float32 if prec == "f32" else np.float64
    return [it[3].download(rdt, 2 * int(np.prod(it[0]))).astype(np.float64) for it in items]


def test_batch_matches_oracle_and_sequential():
    items = make_batch(12, 1, max_side=64)
    P.execute_batch([it[4] for it in items])
    got = outputs(items)
    for (n, x, _, _, _), g in zip(items, got):
        want = oracle_out(spec_of("c2c", n), x.astype(np.float64))
        assert rel_l2(g, want) < 2e-6, n
    for it in items:
        it[4].execute()
    seq = outputs(items)
    for a, b in zip(got, seq):
        assert rel_l2(a, b) < 2e-6


def oracle_f64(n, x):
    """the oracle in FP64 (a 128^3 plan in long double would take minutes)"""
    from oracle import semantics as S
    return S.execute_flat(spec_of("c2c", n), x.astype(np.float64), precision=np.float64)


def test_batch_40_plans_match_oracle():
    """40 plans with sides up to 128 (every side of the uneven-batch
    configuration) against the oracle, through the captured graph."""
    items = make_batch(40, 2)
    assert {s for it in items for s in it[0]} == set(SIDES)
    P.execute_batch([it[4] for it in items])
    for (n, x, _, _, _), g in zip(items, outputs(items)):
        assert rel_l2(g, oracle_f64(n, x)) < 2e-6, n


def test_batch_graph_replays_with_eager_work_between():
    items = make_batch(40, 2)
    plans = [it[4] for it in items]
    for it in items:
        it[4].execute()
    eager = outputs(items)
    other = make_batch(3, 99, max_side=48)
    rng = np.random.default_rng(7)
    for rep in range(4):
        P.execute_batch(plans)            # first call captures, later calls replay
        got = outputs(items)
        for a, b in zip(got, eager):
            assert rel_l2(a, b) < 1e-6, rep
        # unrelated eager work between replays
        for it in other:
            it[4].execute()
        # new inputs for a few plans: the next replay must read them
        # (eager executes give the new reference)
        for i in range(rep, 40, 9):
            n, _, din, dout, plan = items[i]
            x = rng.uniform(-1, 1, 2 * int(np.prod(n))).astype(np.float32)
            din.upload(x)
            plan.execute()
            eager[i] = dout.download(np.float32, x.size).astype(np.float64)
            items[i] = (n, x, din, dout, plan)
        # and cleared outputs: the replay must rewrite them
        for it in items[:5]:
            it[3].upload(np.zeros(2 * int(np.prod(it[0])), dtype=np.float32))


def test_batch_without_graph_and_fp64():
    os.environ["PFFT_AMD_BATCH_GRAPH"] = "0"
    try:
        items = make_batch(10, 3, max_side=60)
        P.execute_batch([it[4] for it in items])
        got = outputs(items)
        for (n, x, _, _, _), g in zip(items, got):
            want = oracle_out(spec_of("c2c", n), x.astype(np.float64))
            assert rel_l2(g, want) < 2e-6
    finally:
        os.environ.pop("PFFT_AMD_BATCH_GRAPH", None)
    items = make_batch(8, 4, prec="f64", max_side=48, flags=P.TRANSPOSED_OUT)
    P.execute_batch([it[4] for it in items])
    got = outputs(items, "f64")
    for (n, x, _, _, _), g in zip(items, got):
        want = oracle_out(spec_of("c2c", n, flags=P.TRANSPOSED_OUT), x)
        assert rel_l2(g, want) < 1e-12


def test_uneven_384_batch_workload():
    """The uneven-batch benchmark workload itself: 384 c2c complex64 plans
    with every side drawn from {24, ..., 128} (the bench's generator and
    seed), through the captured graph: the replay reproduces the first call
    bit for bit, every plan agrees with its own eager execution, and a
    sample of 24 plans (every largest side represented) matches the
    oracle."""
    import torch
    dev = torch.device("cuda:0")
    rng = np.random.default_rng(20261202)
    rc, comm = P.create_procmesh([1, 1])
    plans, bufs, shapes = [], [], []
    for _ in range(384):
        n = [int(rng.choice(SIDES)) for _ in range(3)]
        a = (torch.rand(n, device=dev, dtype=torch.complex64) * 2 - (1 + 1j)).to(torch.complex64)
        b = torch.empty_like(a)
        p = P.Plan("c2c", n, a.data_ptr(), b.data_ptr(), comm=comm, prec="f32")
        assert p.ok
        plans.append(p)
        bufs.append((a, b))
        shapes.append(n)
    print("uneven 384-plan batch: %d points" % sum(int(np.prod(n)) for n in shapes), flush=True)
    runs = []
    for rep in range(2):      # the first call captures, the second replays
        for _, b in bufs:
            b.zero_()
        P.execute_batch(plans)
        torch.cuda.synchronize()
        runs.append([b.clone() for _, b in bufs])
    for i in range(384):
        assert torch.equal(runs[0][i], runs[1][i]), (i, shapes[i])
    # eager execution may pick other kernels (power-of-two lengths), so
    # the batch agrees with it to FP32 accuracy rather than bit for bit
    for p in plans:
        p.execute()
    torch.cuda.synchronize()
    for i, (_, b) in enumerate(bufs):
        d = (runs[0][i] - b).abs().pow(2).sum().sqrt() / b.abs().pow(2).sum().sqrt()
        assert float(d) < 2e-6, (i, shapes[i], float(d))
    # sample: the first plan of each side class plus a seeded draw
    pick, seen = [], set()
    for i, n in enumerate(shapes):
        if max(n) not in seen:
            seen.add(max(n))
            pick.append(i)
    pick += [int(i) for i in np.random.default_rng(5).choice(384, 24 - len(pick), replace=False)
             if int(i) not in pick]
    for i in pick:
        x = bufs[i][0].cpu().numpy().astype(np.complex128).reshape(-1)
        flat = np.empty(2 * x.size)
        flat[0::2], flat[1::2] = x.real, x.imag
        y = runs[0][i].cpu().numpy().astype(np.complex128).reshape(-1)
        g = np.empty(2 * y.size)
        g[0::2], g[1::2] = y.real, y.imag
        assert rel_l2(g, oracle_f64(shapes[i], flat)) < 2e-6, shapes[i]
    for p in plans:
        p.destroy()
