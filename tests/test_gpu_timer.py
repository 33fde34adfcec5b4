"""Plan timers on the GPU: every execute counts, each pass's device time
lands in the trafo slot of the reference step that transforms its
dimension (T_OUT pipeline: trafo[0] last dims, trafo[r-d] dim d; T_IN
pipeline: trafo[r+1+d], trafo[2r+1]), and the Matlab-style output of
write_average_timer_adv (kernel/timer.c:114-220, 402-455)."""
import ctypes

import numpy as np
import pytest

import pfft_amd as P

pytestmark = pytest.mark.gpu
D = ctypes.POINTER(ctypes.c_double)


def timer_vec(plan):
    L = P.lib()
    L.pfft_get_timer.argtypes = [ctypes.c_void_p]
    L.pfft_get_timer.restype = ctypes.c_void_p
    L.pfft_convert_timer2vec.argtypes = [ctypes.c_void_p]
    L.pfft_convert_timer2vec.restype = D
    L.pfft_destroy_timer.argtypes = [ctypes.c_void_p]
    L.free.argtypes = [ctypes.c_void_p]
    t = L.pfft_get_timer(plan.handle)
    p = L.pfft_convert_timer2vec(t)
    r = int(p[0])
    n = int(p[1]) + int(p[2]) + 9
    out = [p[i] for i in range(n)]
    L.free(ctypes.cast(p, ctypes.c_void_p))
    L.pfft_destroy_timer(t)
    return r, out


@pytest.mark.parametrize("kind,flags,busy", [
    ("c2c", 0, {0, 1, 2}),
    ("c2c", P.TRANSPOSED_IN, {3, 4, 5}),
    ("r2c", P.TRANSPOSED_OUT, {0, 1, 2}),
    ("c2r", P.TRANSPOSED_IN, {3, 4, 5}),
])
def test_timer_slots(kind, flags, busy, tmp_path):
    rc, comm = P.create_procmesh([1, 1])
    n = [64, 48, 40]
    d_in = P.DeviceBuffer(16 * 64 * 48 * 41 * 2)
    d_out = P.DeviceBuffer(16 * 64 * 48 * 41 * 2)
    sign = +1 if kind == "c2r" else -1
    plan = P.Plan(kind, n, d_in.ptr, d_out.ptr, sign=sign, flags=flags, comm=comm)
    assert plan.ok
    for _ in range(3):
        plan.execute()
    r, v = timer_vec(plan)
    assert r == 2 and v[3] == 3 and v[4] > 0
    trafo = v[5:5 + 2 * r + 2]
    remap = v[5 + 2 * r + 2:5 + 4 * r + 2]
    assert {i for i, x in enumerate(trafo) if x > 0} == busy, trafo
    assert all(x == 0 for x in remap)
    assert sum(trafo) <= v[4]
    L = P.lib()
    L.pfft_write_average_timer_adv.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int]
    f = tmp_path / "t.m"
    L.pfft_write_average_timer_adv(plan.handle, str(f).encode(), comm)
    text = f.read_text()
    pre = "pfft_back" if sign > 0 else "pfft_forw"
    assert "n_pfft(1, 1:3) = [64 48 40 ];" in text
    assert "%s_iter(1)    = 3;" % pre in text
    for k in range(1, 7):
        assert "%s_trafo%d(1, 2)" % (pre, k) in text
    L.pfft_reset_timer.argtypes = [ctypes.c_void_p]
    L.pfft_reset_timer(plan.handle)
    assert timer_vec(plan)[1][3] == 0
    plan.destroy()
    d_in.free()
    d_out.free()
