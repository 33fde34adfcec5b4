"""Layout contract (CPU): local sizes, starts and allocation sizes.

Three independent sources are compared:
  * the product's C planner through the pfft.h C ABI (libpfft.so);
  * the oracle's literal Python restatement (oracle/layout.py);
  * closed forms fixed by the reference's documentation and the values
    derived from its source for the benchmark configurations.
"""
import itertools

import pytest

import pfft_amd as P
from oracle import layout as L

TI, TO, SI, SO = P.TRANSPOSED_IN, P.TRANSPOSED_OUT, P.SHIFTED_IN, P.SHIFTED_OUT
PAD = P.PADDED_R2C
TF = {"c2c": L.T_C2C, "r2c": L.T_R2C, "c2r": L.T_C2R, "r2r": L.T_R2R}


def comms():
    rc2, c2 = P.create_procmesh([1, 1])
    rc1, c1 = P.create_procmesh([1])
    assert rc2 == 0 and rc1 == 0
    return {2: c2, 1: c1, 0: P.COMM_WORLD}


SIZES = [
    ([64, 64, 64], None, None),
    ([512, 512, 512], None, None),
    ([29, 27, 31], None, None),
    ([29, 27, 31], [16, 16, 16], [16, 16, 16]),   # tests/simple_check_ousam_*.c
    ([256, 256, 256], [128, 128, 128], None),     # oversampled (configuration 5)
    ([4, 4, 6], None, None),                      # tests/simple_check_c2r_c2c_shifted.c
    ([6, 6, 6], [4, 4, 4], [4, 4, 4]),            # tests/simple_check_c2r_c2c_ousam_shifted.c
    ([13, 13, 13, 13], None, None),               # 4-d tests
    ([8, 10], None, None),
    ([1024, 1024, 512], None, None),
]


def cases():
    for kind in ("c2c", "r2c", "c2r", "r2r"):
        for n, ni, no in SIZES:
            for flags in range(16):
                if flags & 3 == 3:
                    continue
                for pad in ((0, PAD) if kind in ("r2c", "c2r") else (0,)):
                    yield kind, n, ni or n, no or n, flags | pad


@pytest.mark.parametrize("howmany", [1, 3])
@pytest.mark.parametrize("mesh", [2, 1, 0])
def test_product_matches_oracle(mesh, howmany):
    cm = comms()
    rnk_pm = 1 if mesh == 0 else mesh
    bad = []
    n_cases = 0
    for kind, n, ni, no, flags in cases():
        if len(n) < rnk_pm or (len(n) == rnk_pm and len(n) != 3):
            continue
        tf = TF[kind] | (L.T_PADDED if flags & PAD else 0)
        want = L.local_size(n, ni, no, howmany, None, None, rnk_pm, tf, flags)
        got = P.local_size(kind, n, ni, no, howmany, comm=cm[mesh], flags=flags)
        n_cases += 1
        if tuple(want) != tuple(got):
            bad.append((kind, n, ni, no, flags, want, got))
    assert n_cases > 300
    assert not bad, bad[:3]


def test_user_blocks_match_oracle():
    cm = comms()
    for n, blk in (([16, 12, 10], [16, 12]), ([16, 12, 10], [20, 30]), ([16, 12, 10], [8, 6])):
        for flags in (0, TO, TI):
            want = L.local_size(n, n, n, 1, blk, blk, 2, L.T_C2C, flags)
            got = P.local_size("c2c", n, iblock=blk, oblock=blk, comm=cm[2], flags=flags)
            assert tuple(want) == tuple(got), (n, blk, flags)


def test_local_block_matches_local_size_starts():
    cm = comms()
    for kind, n, ni, no, flags in cases():
        if len(n) != 3 or ni != n or no != n:
            continue
        tf = TF[kind] | (L.T_PADDED if flags & PAD else 0)
        want = L.local_block(n, n, None, None, 2, 0, tf, flags)
        got = P.local_block(kind, n, comm=cm[2], flags=flags)
        assert tuple(want) == tuple(got), (kind, n, flags)


# ---- closed forms ------------------------------------------------------
def test_configuration_values_derived_from_reference():
    """Values derived by reading kernel/partrafo*.c for the benchmark
    configurations on a 1x1 mesh."""
    cm = comms()
    # configuration 1: c2c 64^3 default layout
    assert P.local_size("c2c", [64] * 3, comm=cm[2]) == (
        262144, [64, 64, 64], [0, 0, 0], [64, 64, 64], [0, 0, 0])
    # configuration 2: r2c 512^3 TRANSPOSED_OUT, c2r TRANSPOSED_IN
    assert P.local_size("r2c", [512] * 3, flags=TO, comm=cm[2]) == (
        67371008, [512, 512, 512], [0, 0, 0], [512, 512, 257], [0, 0, 0])
    assert P.local_size("c2r", [512] * 3, flags=TI, comm=cm[2]) == (
        67371008, [512, 512, 257], [0, 0, 0], [512, 512, 512], [0, 0, 0])
    # configuration 3
    assert P.local_size("c2c", [1024, 1024, 512], comm=cm[2])[0] == 536870912
    # configuration 5: ni=128^3 -> n=256^3 shifted in and out
    assert P.local_size("r2c", [256] * 3, ni=[128] * 3, flags=SI | SO, comm=cm[2]) == (
        8454144, [128, 128, 128], [-64, -64, -64], [256, 256, 129], [-128, -128, -128])


def test_shifted_starts_are_minus_half():
    # kernel/partrafo.c:287-299
    cm = comms()
    for n in ([8, 6, 4], [16, 16, 10]):
        m, lni, lis, lno, los = P.local_size("c2c", n, flags=SI | SO, comm=cm[2])
        assert lis == [-x // 2 for x in n] and los == [-x // 2 for x in n]


def test_padded_r2c_real_counts():
    # partrafo-transposed.c:67-85: padded r2c input reports 2*(n/2+1) reals
    cm = comms()
    m, lni, _, lno, _ = P.local_size("r2c", [8, 6, 10], flags=PAD, comm=cm[2])
    assert lni == [8, 6, 12] and lno == [8, 6, 6]
    m, lni, _, lno, _ = P.local_size("r2c", [8, 6, 10], comm=cm[2])
    assert lni == [8, 6, 10]
    m, lni, _, lno, _ = P.local_size("c2r", [8, 6, 10], flags=PAD, comm=cm[2])
    assert lni == [8, 6, 6] and lno == [8, 6, 12]


def test_alloc_is_max_of_embedded_sizes():
    """The reference embeds and truncates one dimension per step
    (kernel/outrafo.c:118-168), so with ni = no = 16 < n = 32 the largest
    intermediate holds one embedded dimension: 16*16*32 complex for c2c;
    for r2c the later steps see (16/2+1)*16*32."""
    cm = comms()
    m = P.local_size("c2c", [32, 32, 32], ni=[16, 16, 16], no=[16, 16, 16], comm=cm[2])[0]
    assert m == 16 * 16 * 32
    m = P.local_size("r2c", [32, 32, 32], ni=[16, 16, 16], no=[16, 16, 16], comm=cm[2])[0]
    assert m == 9 * 16 * 32
    # no truncation: the fully embedded array is the largest
    m = P.local_size("c2c", [32, 32, 32], ni=[16, 16, 16], comm=cm[2])[0]
    assert m == 32 * 32 * 32


def test_transposed_orders():
    # doc/tutorial.tex:319-330: n1 x .. x n_r x n0 x n_{r+1} x ..
    assert L.transposed_order(3, 2) == [1, 2, 0]
    assert L.transposed_order(3, 1) == [1, 0, 2]
    assert L.transposed_order(4, 2) == [1, 2, 0, 3]
    assert L.transposed_order(2, 1) == [1, 0]


# ---- process meshes ------------------------------------------------------
def test_procmesh_rules():
    # kernel/procmesh.c:47-48: 1 if prod(np) != comm size
    assert P.create_procmesh([2, 1])[0] == 1
    assert P.create_procmesh([1, 1])[0] == 0
    assert P.create_procmesh([1])[0] == 0
    assert P.create_procmesh([1, 1, 1])[0] == 0


def test_null_plan_rules_match_oracle():
    """PX(plan_partrafo) returns NULL for these combinations
    (kernel/partrafo.c:337-373); checked on the oracle, which the GPU
    tests then compare with the library's plan creation."""
    assert L.plan_is_null([5, 4, 4], 2, L.T_C2C, SI)
    assert L.plan_is_null([4, 4, 4], 2, L.T_C2C, TI | TO)
    assert L.plan_is_null([4, 4, 4], 2, L.T_R2C, TI)
    assert L.plan_is_null([4, 4, 4], 2, L.T_C2R, TO)
    assert L.plan_is_null([4, 4], 2, L.T_C2C, 0)
    assert not L.plan_is_null([4, 4, 4], 2, L.T_C2C, TO)
    assert not L.plan_is_null([4, 4, 4], 1, L.T_R2C, TO)
