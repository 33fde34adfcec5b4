"""GPU checks of the autotuner: every launch variant the planner can choose
computes the transform (parity with the oracle), forced through wisdom
files; timing-based selection fills the wisdom cache and a second plan of
the same geometry reuses it without timing; PFFT_ESTIMATE plans time
nothing."""
import os

import numpy as np
import pytest

import pfft_amd as P
from gpu_util import spec_of, random_input, run_gpu, oracle_out, out_mask, rel_l2, comm_for

pytestmark = pytest.mark.gpu

TI, TO = P.TRANSPOSED_IN, P.TRANSPOSED_OUT

CASES = [
    ("c2c", [64, 128, 256], TO, "f64"),
    ("r2c", [128, 64, 512], TO, "f64"),
    ("c2r", [128, 64, 512], TI, "f64"),
    ("c2c", [1024, 8, 16], 0, "f32"),
    ("r2c", [32, 16, 2048], 0, "f32"),
    # complex64 transposing passes (line-major variants stage the
    # tile-adjacent side through LDS) and real passes with a contiguous
    # complex side (line-major real kernels)
    ("r2c", [64, 32, 512], TO, "f32"),
    ("c2r", [64, 32, 512], TI, "f32"),
    ("c2c", [128, 64, 256], TO, "f32"),
]


@pytest.fixture
def tune_all(monkeypatch):
    """Tuning for every pass size, fresh cache, no wisdom file."""
    monkeypatch.setenv("PFFT_AMD_TUNE", "2")
    monkeypatch.delenv("PFFT_AMD_WISDOM", raising=False)
    monkeypatch.delenv("PFFT_AMD_XCD_REMAP", raising=False)
    P.forget_wisdom()
    yield
    P.forget_wisdom()


def _plan_lines(kind, n, flags, prec):
    spec = spec_of(kind, n, sign=(+1 if kind == "c2r" else -1), flags=flags)
    esz = 8 if prec == "f64" else 4
    from gpu_util import n_scalars
    a = P.DeviceBuffer(n_scalars(spec, "in") * esz)
    b = P.DeviceBuffer(n_scalars(spec, "out") * esz)
    pl = P.Plan(kind, n, a.ptr, b.ptr, sign=spec.sign, flags=flags, comm=comm_for(2), prec=prec)
    assert pl.ok
    lines = pl.describe()
    pl.destroy()
    a.free()
    b.free()
    return spec, lines


def _rewrite(path, variant, remap):
    lines = open(path).read().splitlines()
    out = [lines[0]]
    for ln in lines[1:]:
        key, v, r, ms = ln.split()
        out.append("%s %d %d %s" % (key, variant, remap, ms))
    open(path, "w").write("\n".join(out) + "\n")


SMALL = {2, 3, 4, 5, 6, 8, 10, 12, 15, 16, 20, 24, 30, 32, 40, 48, 60, 64, 80, 96, 120, 128}


def _expected(line, prec, variant, remap):
    """Kernel the describe line must name for a forced (variant, remap)."""
    import re
    m = re.search(r"pass \d+: (\w+) n=(\d+) nf=(\d+)", line)
    kind, n, nf = m.group(1), int(m.group(2)), int(m.group(3))
    pow2 = 64 <= nf <= 4096 and nf & (nf - 1) == 0
    small = kind == "c2c" and nf in SMALL
    if variant == 1:
        return "general"
    if variant == 3 and small:
        return "small"
    if pow2:
        name = {0: "fast", 2: "fast-narrow", 3: "fast", 4: "fast-grid", 6: "fast-narrow-grid"}
        full = {"c2c": (nf, nf), "r2c": (n, nf + 1), "c2r": (nf + 1, n)}.get(kind)
        plain = full is not None and "nin=%d@0 nout=%d@0" % full in line
        lm = (variant & 8) and kind != "r2r" and (prec == "f32" or plain)
        return name[variant & 7] + ("-lm" if lm else "") + ("+xcd" if remap else "")
    return "small" if small else "general"


@pytest.mark.parametrize("variant,remap", [(0, 0), (0, 1), (2, 0), (2, 1), (4, 0), (4, 1),
                                           (6, 0), (6, 1), (8, 0), (12, 1), (10, 0), (3, 0),
                                           (1, 0)])
@pytest.mark.parametrize("kind,n,flags,prec", CASES)
def test_forced_variant_parity(tune_all, tmp_path, kind, n, flags, prec, variant, remap):
    spec, lines = _plan_lines(kind, n, flags, prec)     # tunes, fills the cache
    assert P.wisdom_count() >= 1, lines
    w = tmp_path / "w.txt"
    assert P.export_wisdom(str(w)) == 0
    _rewrite(str(w), variant, remap)
    P.forget_wisdom()
    assert P.import_wisdom(str(w)) == 0
    t0 = P.tune_timings()
    _, forced = _plan_lines(kind, n, flags, prec)
    assert P.tune_timings() == t0, "wisdom hit must not time"
    _, est = _plan_lines(kind, n, flags | P.ESTIMATE, prec)
    for ln, le in zip(forced, est):
        tuned = not le.endswith("kernel=general")   # passes with a choice were tuned
        want = _expected(ln, prec, variant, remap) if tuned else "general"
        assert ln.endswith("kernel=" + want), (ln, want)
    x = random_input(spec, 5)
    y = run_gpu(spec, x, prec=prec)
    ref = oracle_out(spec, x)
    tol = 1e-12 if prec == "f64" else 2e-6
    assert rel_l2(y, ref, out_mask(spec)) < tol


def test_measure_fills_cache_and_reuses(tune_all):
    t0 = P.tune_timings()
    _plan_lines("c2c", [64, 128, 256], TO, "f64")
    n1, t1 = P.wisdom_count(), P.tune_timings()
    # three passes; candidates per pass: eight power-of-two variants, the
    # short-length kernel where the length has one (64 and 128), the general
    # one, and four line-major variants on the passes whose lines are
    # contiguous on a side (the 256-point line pass and the transposing
    # 64-point pass; not the 128-point column pass)
    assert n1 == 3 and t1 - t0 == (10 + 4) + 10 + (9 + 4)
    _plan_lines("c2c", [64, 128, 256], TO, "f64")
    assert P.wisdom_count() == n1 and P.tune_timings() == t1
    # different geometry -> new selections
    _plan_lines("c2c", [64, 128, 256], 0, "f64")
    assert P.wisdom_count() > n1 and P.tune_timings() > t1


def test_estimate_times_nothing(tune_all):
    t0 = P.tune_timings()
    _, lines = _plan_lines("c2c", [64, 128, 256], TO | P.ESTIMATE, "f64")
    assert P.tune_timings() == t0 and P.wisdom_count() == 0
    assert all(ln.endswith("kernel=fast") for ln in lines)


def test_wisdom_env_file(tune_all, tmp_path, monkeypatch):
    w = tmp_path / "env_wisdom.txt"
    monkeypatch.setenv("PFFT_AMD_WISDOM", str(w))
    _plan_lines("r2c", [128, 64, 512], TO, "f64")
    assert w.exists()
    assert open(w).read().startswith("pfft-amd-wisdom 1 3")


def test_tuned_large_plan_roundtrip(monkeypatch):
    """Default effort (MEASURE) on a plan whose passes exceed the timing
    threshold, then a round trip; the tuned plan must not have touched the
    arrays it was planned on."""
    monkeypatch.delenv("PFFT_AMD_TUNE", raising=False)
    monkeypatch.delenv("PFFT_AMD_WISDOM", raising=False)
    P.forget_wisdom()
    n = [256, 256, 256]
    rc, comm = P.create_procmesh([1, 1])
    mem = P.local_size("r2c", n, flags=TO, comm=comm)[0]
    x = np.random.default_rng(3).standard_normal(int(np.prod(n)))
    dx = P.DeviceBuffer(x.size * 8)
    dk = P.DeviceBuffer(mem * 16)
    dy = P.DeviceBuffer(x.size * 8)
    dx.upload(x)
    t0 = P.tune_timings()
    fw = P.Plan("r2c", n, dx.ptr, dk.ptr, flags=TO, comm=comm)
    bw = P.Plan("c2r", n, dk.ptr, dy.ptr, sign=+1, flags=TI, comm=comm)
    assert fw.ok and bw.ok
    assert P.tune_timings() > t0
    assert np.array_equal(dx.download(np.float64, x.size), x)
    fw.execute()
    bw.execute()
    y = dy.download(np.float64, x.size)
    assert np.max(np.abs(y / x.size - x)) < 1e-12
    for o in (fw, bw):
        o.destroy()
    for b in (dx, dk, dy):
        b.free()
    P.forget_wisdom()


@pytest.mark.parametrize("variant,remap", [(8, 0), (12, 1), (4, 0)])
@pytest.mark.parametrize("form,tol,name", [(P.AMD_MFMA_F32, 5e-6, "mfma-f32"),
                                           (P.AMD_MFMA_BF16X2, 1e-5, "mfma-bf16x2")])
@pytest.mark.parametrize("n,flags", [([32, 256, 1024], TI), ([40, 512, 16], 0),
                                     ([24, 16, 512], 0), ([8, 40, 256], 0)])
def test_forced_mfma_variants(tune_all, tmp_path, n, flags, form, tol, name, variant, remap):
    """Every launch variant of the MFMA register kernel (line-major lane
    order, per-tile grid, XCD order), forced through a wisdom file."""
    spec, _ = _plan_lines("c2c", n, flags | form, "f32")
    w = tmp_path / "w.txt"
    assert P.export_wisdom(str(w)) == 0
    _rewrite(str(w), variant, remap)
    P.forget_wisdom()
    assert P.import_wisdom(str(w)) == 0
    _, lines = _plan_lines("c2c", n, flags | form, "f32")
    for ln in lines:
        if "kernel=mfma" in ln:
            assert ln.endswith("kernel=" + name + ("-lm" if variant & 8 else "")), ln
    for sign in (-1, +1):
        spec = spec_of("c2c", n, flags=flags, sign=sign)
        x = random_input(spec, 5).astype(np.float32).astype(np.float64)
        y = run_gpu(spec, x, prec="f32", plan_flags_extra=form)
        assert rel_l2(y, oracle_out(spec, x)) <= tol
