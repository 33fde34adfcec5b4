"""Plan timers, host side only (no device calls): vector conversion,
copy, sum and average of pfft_timer objects (kernel/timer.c:222-340)."""
import ctypes

import pytest

import pfft_amd as P

try:
    P.lib()
    HAVE_LIB = True
except Exception:   # library not built
    HAVE_LIB = False

pytestmark = pytest.mark.skipif(not HAVE_LIB, reason="libpfft.so not built")
D = ctypes.POINTER(ctypes.c_double)


def setup_lib():
    L = P.lib()
    for pre in ("pfft_", "pfftf_"):
        getattr(L, pre + "convert_vec2timer").argtypes = [D]
        getattr(L, pre + "convert_vec2timer").restype = ctypes.c_void_p
        getattr(L, pre + "convert_timer2vec").argtypes = [ctypes.c_void_p]
        getattr(L, pre + "convert_timer2vec").restype = D
        for nm in ("copy_timer",):
            getattr(L, pre + nm).argtypes = [ctypes.c_void_p]
            getattr(L, pre + nm).restype = ctypes.c_void_p
        getattr(L, pre + "add_timers").argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        getattr(L, pre + "add_timers").restype = ctypes.c_void_p
        for nm in ("average_timer", "destroy_timer"):
            getattr(L, pre + nm).argtypes = [ctypes.c_void_p]
            getattr(L, pre + nm).restype = None
    L.free.argtypes = [ctypes.c_void_p]
    return L


def vec(L, pre, t, n):
    p = getattr(L, pre + "convert_timer2vec")(t)
    out = [p[i] for i in range(n)]
    L.free(ctypes.cast(p, ctypes.c_void_p))
    return out


@pytest.mark.parametrize("pre", ["pfft_", "pfftf_"])
def test_timer_vectors_roundtrip_and_arithmetic(pre):
    L = setup_lib()
    r = 2
    ntr, nre = 2 * r + 2, 2 * r
    n = ntr + nre + 9
    v = [float(r), float(ntr), float(nre), 4.0, 8.0] + [float(i + 1) for i in range(ntr)] + \
        [0.5 * (i + 1) for i in range(nre)] + [0.25, 0.75, 1.5, 2.5]
    assert len(v) == n
    t = getattr(L, pre + "convert_vec2timer")((ctypes.c_double * n)(*v))
    assert vec(L, pre, t, n) == v
    c = getattr(L, pre + "copy_timer")(t)
    s = getattr(L, pre + "add_timers")(t, c)
    sv = vec(L, pre, s, n)
    assert sv[:3] == v[:3] and sv[3:] == [2 * x for x in v[3:]]
    getattr(L, pre + "average_timer")(s)
    av = vec(L, pre, s, n)
    assert av[3] == 1.0 and av[4:] == [x / 4.0 for x in v[4:]]
    for x in (t, c, s):
        getattr(L, pre + "destroy_timer")(x)
