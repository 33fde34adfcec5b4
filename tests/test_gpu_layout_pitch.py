"""Pitches of the intermediate arrays a plan chooses (pfft_amd/csrc/plan.c):
rows rounded up to whole 128-byte segments and then to an odd number of
segments, planes one segment longer than their rows, and rows shorter than
16 segments (2 KiB) left unpadded.  Read from the plan description: the
line stride of a pass along dimension 1 is the row pitch, along dimension
0 the plane pitch.  Results of these layouts are checked against the oracle
by the parity tests; this pins the layout rule itself."""
import re

import pytest

import pfft_amd as P
from gpu_util import comm_for

pytestmark = pytest.mark.gpu

TO = P.TRANSPOSED_OUT


def _strides(kind, n, flags, prec):
    c = comm_for(2)
    mem = P.local_size(kind, n, flags=flags, comm=c)[0]
    esz = 16 if prec == "f64" else 8
    a, b = P.DeviceBuffer(mem * esz), P.DeviceBuffer(mem * esz)
    pl = P.Plan(kind, n, a.ptr, b.ptr, flags=flags | P.ESTIMATE, comm=c, prec=prec)
    assert pl.ok
    lines = pl.describe()
    pl.destroy()
    a.free()
    b.free()
    out = []
    for ln in lines:
        m = re.search(r"sin=(\d+) sout=(\d+)", ln)
        assert m, ln
        out.append((int(m.group(1)), int(m.group(2))))
    return out, lines


@pytest.mark.parametrize("prec,row", [("f64", 264), ("f32", 272)])
def test_r2c_rows_of_an_odd_segment_count_stay(prec, row):
    # 257 complex per row: 264 complex128 = 33 segments, 272 complex64 = 17,
    # both odd, so no extra segment; planes get one (8 or 16 elements)
    s, lines = _strides("r2c", [64, 64, 512], TO, prec)
    seg = 8 if prec == "f64" else 16
    assert s[0][1] == 1, lines                  # pass 0 writes whole rows
    assert s[1] == (row, row), lines            # pass 1: lines along dim 1
    assert s[2][0] == 64 * row + seg, lines     # pass 2 reads along dim 0


def test_even_segment_rows_get_one_more():
    # 512 complex128 per row = 64 segments -> 65 (520 elements); the first
    # pass (dimension 1) reads the user array and writes the intermediate
    s, lines = _strides("c2c", [64, 64, 512], 0, "f64")
    assert s[0] == (512, 520), lines
    assert s[1] == (64 * 520 + 8, 64 * 520 + 8), lines


@pytest.mark.parametrize("kind,n,prec,row", [("c2c", [64, 64, 100], "f64", 100),
                                             ("r2c", [64, 64, 128], "f32", 65)])
def test_short_rows_stay_unpadded(kind, n, prec, row):
    s, lines = _strides(kind, n, TO if kind == "r2c" else 0, prec)
    if kind == "c2c":
        assert s[0] == (row, row), lines
        assert s[1] == (64 * row, 64 * row), lines
    else:
        assert s[1] == (row, row), lines
        assert s[2][0] == 64 * row, lines
