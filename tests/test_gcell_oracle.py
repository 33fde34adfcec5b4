"""The oracle's ghost-cell exchange/reduce (tests only; no device needed),
and the library's host-side ghost-cell data distribution.

The oracle restates the reference's chunked send/recv loops; these tests
pin it with what the mathematics fixes on one process:
  * exchange is the periodic extension of the local block
    (doc/reference.tex:570-586: "the data is wrapped periodically"),
  * reduce adds every box element into its periodic owner
    (doc/reference.tex:602-605: "adds the sum of all the duplicates of a
    local data element to the original data element"), brute force,
  * reduce is the adjoint of exchange: <E x, y> = <x, R y>,
and planted mistakes (below/above swapped, no truncation) fail them."""
import itertools

import numpy as np
import pytest

from oracle import gcell as G

CASES = [
    # n (periods),  gb,        ga,        tuple
    ([4, 5, 6], [1, 2, 0], [0, 1, 3], 1),
    ([3, 2, 4], [2, 0, 1], [1, 3, 0], 2),
    ([2, 3, 2], [5, 1, 0], [0, 7, 4], 1),      # ghost cells beyond one period
    ([5, 1, 3], [0, 2, 6], [4, 1, 0], 3),
    ([1, 4, 2], [3, 0, 2], [2, 2, 5], 2),
]


def params(n, gb, ga, t):
    return {"n": n, "loc_n": n, "gb": gb, "ga": ga,
            "ngc": [gb[i] + n[i] + ga[i] for i in range(3)], "tuple": t}


def periodic(x, p):
    """closed form: box[k] = x[(k - gb) mod n]"""
    idx = [(np.arange(p["ngc"][t]) - p["gb"][t]) % p["n"][t] for t in range(3)]
    return x[np.ix_(idx[0], idx[1], idx[2])]


def brute_reduce(y, p):
    out = np.zeros(p["n"] + [p["tuple"]])
    for k in itertools.product(*[range(v) for v in p["ngc"]]):
        o = tuple((k[t] - p["gb"][t]) % p["n"][t] for t in range(3))
        out[o] += y[k]
    return out


@pytest.mark.parametrize("n,gb,ga,t", CASES)
def test_exchange_is_periodic_extension(n, gb, ga, t):
    p = params(n, gb, ga, t)
    L = t * int(np.prod(n))
    Gs = t * int(np.prod(p["ngc"]))
    rng = np.random.default_rng(1)
    buf = np.full(Gs + 5, 7.25)
    x = rng.standard_normal(L)
    buf[:L] = x
    out = G.exchange(buf, p)
    want = periodic(x.reshape(n + [t]), p).reshape(-1)
    assert np.array_equal(out[:Gs], want)
    assert np.array_equal(out[Gs:], buf[Gs:])    # nothing beyond the box is touched


@pytest.mark.parametrize("n,gb,ga,t", CASES)
def test_reduce_sums_all_images(n, gb, ga, t):
    p = params(n, gb, ga, t)
    L = t * int(np.prod(n))
    Gs = t * int(np.prod(p["ngc"]))
    rng = np.random.default_rng(2)
    y = rng.integers(-1000, 1000, Gs).astype(np.float64)   # exact sums in any order
    buf = np.concatenate([y, [3.5, 4.5]])
    out = G.reduce(buf, p)
    want = brute_reduce(y.reshape(p["ngc"] + [t]), p).reshape(-1)
    assert np.array_equal(out[:L], want)
    assert np.all(out[L:Gs] == 0)
    assert np.array_equal(out[Gs:], [3.5, 4.5])


@pytest.mark.parametrize("n,gb,ga,t", CASES)
def test_reduce_is_adjoint_of_exchange(n, gb, ga, t):
    p = params(n, gb, ga, t)
    L = t * int(np.prod(n))
    Gs = t * int(np.prod(p["ngc"]))
    rng = np.random.default_rng(3)
    x = rng.standard_normal(L)
    y = rng.standard_normal(Gs)
    bx = np.zeros(Gs)
    bx[:L] = x
    Ex = G.exchange(bx, p)
    Ry = G.reduce(y.copy(), p)
    lhs, rhs = float(Ex @ y), float(x @ Ry[:L])
    assert abs(lhs - rhs) <= 1e-12 * (abs(lhs) + 1)


def test_reduce_of_exchange_counts_images():
    """R(E(x)) multiplies each element by its number of images."""
    n, gb, ga = [3, 4, 2], [4, 1, 0], [2, 5, 3]
    p = params(n, gb, ga, 1)
    L = int(np.prod(n))
    Gs = int(np.prod(p["ngc"]))
    x = np.arange(1, L + 1, dtype=np.float64)
    b = np.zeros(Gs)
    b[:L] = x
    r = G.reduce(G.exchange(b, p), p)[:L].reshape(n)
    cnt = [np.bincount((np.arange(p["ngc"][t]) - gb[t]) % n[t], minlength=n[t])
           for t in range(3)]
    want = x.reshape(n) * cnt[0][:, None, None] * cnt[1][None, :, None] * cnt[2][None, None, :]
    assert np.array_equal(r, want)


def test_planted_below_above_swap_is_detected():
    n, gb, ga = [4, 5, 6], [1, 2, 0], [0, 1, 3]
    p = params(n, gb, ga, 1)
    wrong = params(n, ga, gb, 1)          # the two ghost widths mixed up
    x = np.random.default_rng(4).standard_normal(int(np.prod(n)))
    b = np.zeros(int(np.prod(p["ngc"])))
    b[:x.size] = x
    out = G.exchange(b, wrong)
    assert not np.array_equal(out[:b.size], periodic(x.reshape(n + [1]), p).reshape(-1))


def test_planted_missing_truncation_is_detected(monkeypatch):
    n, gb, ga = [3, 3, 3], [1, 0, 1], [1, 1, 0]
    p = params(n, gb, ga, 1)
    L = int(np.prod(n))
    y = np.random.default_rng(6).integers(-50, 50, int(np.prod(p["ngc"]))).astype(float)
    want = brute_reduce(y.reshape(p["ngc"] + [1]), p).reshape(-1)
    assert np.array_equal(G.reduce(y, p)[:L], want)
    monkeypatch.setattr(G, "truncate_outside", lambda buf, *a: buf)   # plant: no compaction
    assert not np.array_equal(G.reduce(y, p)[:L], want)


def test_planted_single_chunk_is_detected(monkeypatch):
    """Ghost widths beyond one period need the chunk loop; a single
    min(gc, n) copy leaves the far ghost cells zero."""
    n, gb, ga = [2, 3, 2], [5, 1, 0], [0, 7, 4]
    p = params(n, gb, ga, 1)
    x = np.random.default_rng(7).standard_normal(int(np.prod(n)))
    b = np.zeros(int(np.prod(p["ngc"])))
    b[:x.size] = x
    want = periodic(x.reshape(n + [1]), p).reshape(-1)
    assert np.array_equal(G.exchange(b, p), want)
    orig = G._copy_slices
    seen = set()

    def first_chunk_only(box, dim, send, recv, count):
        if (dim, send < recv) in seen:
            return
        seen.add((dim, send < recv))
        orig(box, dim, send, recv, count)
    monkeypatch.setattr(G, "_copy_slices", first_chunk_only)
    assert not np.array_equal(G.exchange(b, p), want)


def test_plan_params_rules():
    # zero ghost cells and rank != 3: NULL plan (gcells_plan.c:97-113)
    assert G.plan_params([4, 4, 4], 1, None, None, 0, 2, False) is None
    assert G.plan_params([4, 4], 1, [1, 1], [1, 1], 0, 2, False) is None
    # transposed order on a 2-d mesh swaps the first two dimensions
    p = G.plan_params([4, 5, 6], 1, [1, 2, 3], [0, 1, 0], G.GC_TRANSPOSED, 2, False)
    assert p["n"] == [5, 4, 6] and p["gb"] == [2, 1, 3] and p["ga"] == [1, 0, 0]
    # ... and changes nothing on a 1-d mesh
    p = G.plan_params([4, 5, 6], 1, [1, 2, 3], [0, 1, 0], G.GC_TRANSPOSED, 1, False)
    assert p["n"] == [4, 5, 6]
    # padded real arrays: 2*(n/2+1) along the last dimension
    p = G.plan_params([4, 5, 7], 2, [1, 0, 0], None, G.GC_PADDED_R2C, 2, False)
    assert p["n"] == [4, 5, 8] and p["tuple"] == 2
    # complex data: tuple 2*howmany; r2c arrays n/2+1; padding ignored
    p = G.plan_params([4, 5, 7], 3, [1, 0, 0], None, G.GC_PADDED_R2C, 2, True)
    assert p["n"] == [4, 5, 4] and p["tuple"] == 6
    p = G.plan_params([4, 5, 7], 1, [1, 0, 0], None, G.GC_PADDED, 2, True)
    assert p["n"] == [4, 5, 7]


def test_local_size_gc_matches_documented_formula():
    lmem, lngc, lgcs = G.local_size_gc([4, 5, 6], [0, -2, 3], 2, [1, 2, 0], [3, 0, 1])
    assert lngc == [8, 7, 7] and lgcs == [-1, -4, 3] and lmem == 2 * 8 * 7 * 7
    assert G.local_size_gc([4, 5, 6], [0, 0, 0], 1, None, None) == (120, [4, 5, 6], [0, 0, 0])


try:
    import pfft_amd as P
    P.lib()
    HAVE_LIB = True
except Exception:   # library not built
    HAVE_LIB = False


@pytest.mark.skipif(not HAVE_LIB, reason="libpfft.so not built")
@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_library_local_size_gc_matches_oracle(prec):
    rng = np.random.default_rng(5)
    for _ in range(50):
        rnk = int(rng.integers(1, 5))
        ln = [int(v) for v in rng.integers(1, 9, rnk)]
        ls = [int(v) for v in rng.integers(-4, 5, rnk)]
        gb = [int(v) for v in rng.integers(0, 12, rnk)]
        ga = [int(v) for v in rng.integers(0, 12, rnk)]
        hm = int(rng.integers(1, 4))
        want = G.local_size_gc(ln, ls, hm, gb, ga)
        got = P.local_size_gc(ln, ls, gb, ga, howmany=hm, prec=prec)
        assert got == want
    assert P.local_size_gc([3, 4, 5], [0, 0, 0], None, None) == (60, [3, 4, 5], [0, 0, 0])
