"""Measurements of the secondary configurations (bench.py --workload ...).

cfg3  precision study: 3-d c2c 1024x1024x512, complex64 input U[-1,1)^2
      plus a 1e-3-RMS 1/|k|^2 spectral component; forward+backward in FP64
      (plain), FP32 (plain Stockham) and the two MFMA forms; relative L2 /
      max error of each forward result vs an FP64 CPU reference (also part
      of the default bench.py line, as "precision_study").
cfg4  384 3-d c2c complex64 transforms, sides drawn from
      {24,32,48,60,64,80,96,128}: batched (persistent launches per pass
      level and kernel, replayed hipGraph) vs a per-transform execute loop.
cfg5  oversampled shifted r2c ni=128^3 -> n=256^3, loss sum w_k |X_k|^2:
      forward + adjoint (gradient) time, FP64 and FP32.
gcell ghost-cell exchange and reduce of a 512^3 block with 4 ghost cells on
      every side (FP64 real, FP64 complex r2c array, FP32 real).
mixed 3-d c2c forward transforms of mixed-radix sides (480^3, 384^3, 240^3
      and 960x960x64) and r2c forward transforms (480^3, 384^3) in FP64 and
      complex64: per-pass times and kernels, and the relative L2 difference
      from torch.fft.fftn / rfftn of the same input.
pair  the 512^3 r2c+c2r pair in FP64 and in the complex64 forms (FP32,
      FP32-MFMA, split-BF16) without the CPU legs: the default line's
      "mixed" measurement alone, for profiler runs.
Each prints one JSON line.  Inputs are generated on the GPU with torch
(test-data plumbing) and resident in HBM before timing.
"""
import ctypes
import json
import math
import time

import numpy as np

import pfft_amd as P


def _sync():
    import torch
    torch.cuda.synchronize()


def _time(fn, reps, warm=1):
    for _ in range(warm):
        fn()
    _sync()
    t0 = time.perf_counter()
    for _ in range(reps):
        fn()
    _sync()
    return (time.perf_counter() - t0) * 1e3 / reps


def _pass_times(plan, in_ptr, out_ptr, iters=5):
    """Per-pass mean duration (HIP events on the library stream), bytes
    moved, bandwidth and the kernel that runs each pass."""
    L = P.lib()
    L.pfft_amd_time_passes.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                       ctypes.c_int, ctypes.POINTER(ctypes.c_double),
                                       ctypes.POINTER(ctypes.c_double)]
    L.pfft_amd_plan_npasses.argtypes = [ctypes.c_void_p]
    npass = L.pfft_amd_plan_npasses(plan.handle)
    ms = (ctypes.c_double * npass)()
    by = (ctypes.c_double * npass)()
    if L.pfft_amd_time_passes(plan.handle, in_ptr, out_ptr, iters, ms, by):
        return None
    lines = plan.describe()
    return [{"ms": round(ms[k], 4), "GBps": round(by[k] / (ms[k] * 1e-3) / 1e9, 1),
             "kernel": lines[k].split("kernel=")[-1],
             "n": int(lines[k].split(" n=")[1].split()[0])} for k in range(npass)]


def cfg3_input(n, seed, dev):
    """complex64 input U[-1,1)^2 plus a 1/|k|^2 spectral component with
    random phase whose RMS is 1e-3 of the base RMS (exposes dynamic-range
    loss), generated on the GPU from `seed`."""
    import torch
    g = torch.Generator(device=dev).manual_seed(seed)
    x = (torch.rand(n, generator=g, device=dev) * 2 - 1) + \
        1j * (torch.rand(n, generator=g, device=dev) * 2 - 1)
    x = x.to(torch.complex64)
    k0 = torch.fft.fftfreq(n[0], d=1.0 / n[0], device=dev)
    k1 = torch.fft.fftfreq(n[1], d=1.0 / n[1], device=dev)
    k2 = torch.fft.fftfreq(n[2], d=1.0 / n[2], device=dev)
    kk = k0[:, None, None] ** 2 + k1[None, :, None] ** 2 + k2[None, None, :] ** 2
    amp = torch.where(kk > 0, 1.0 / kk.clamp(min=1.0), torch.zeros_like(kk))
    ph = torch.rand(n, generator=g, device=dev) * (2 * math.pi)
    spec = (amp * torch.exp(1j * ph)).to(torch.complex64)
    del amp, ph, kk
    comp = torch.fft.ifftn(spec)
    del spec
    comp *= (1e-3 * x.abs().pow(2).mean().sqrt() / comp.abs().pow(2).mean().sqrt())
    x = (x + comp).contiguous()
    del comp
    return x


def _err_vs(ref_np, got, chunk=1 << 24):
    """relative L2 and max error (relative to max |ref|) of device array
    `got` against the host complex128 array `ref_np`, in chunks"""
    flat = got.reshape(-1)
    num = 0.0
    mx = 0.0
    ref = ref_np.reshape(-1)
    for i in range(0, ref.size, chunk):
        g = flat[i:i + chunk].cpu().numpy().astype(np.complex128)
        d = g - ref[i:i + chunk]
        num += float(np.vdot(d, d).real)
        mx = max(mx, float(np.abs(d).max()))
    den = float(np.vdot(ref, ref).real)
    return math.sqrt(num / den), mx / float(np.abs(ref).max())


def cfg3_forms(seed=20261202, steps=5, n=(1024, 1024, 512), reference=True, cores=16):
    """The precision study: 3-d c2c of `n` (default 1024x1024x512) forward
    then backward in FP64 plain, FP32 plain, FP32 with MFMA radix-16
    stages and split-BF16 MFMA.  Each form is timed over `steps` fwd+bwd
    pairs (after one warm-up pair, inputs resident in HBM); its forward
    output is compared with an FP64 reference of the same input:
    scipy.fft.fftn in double precision on `cores` host threads (pocketfft;
    for this unshifted, untransposed plan the oracle computes exactly this
    transform, which a CPU test pins)."""
    import torch
    dev = torch.device("cuda:0")
    n = list(n)
    N = int(np.prod(n))
    x = cfg3_input(n, seed, dev)
    rc, comm = P.create_procmesh([1, 1])
    res = {"workload": "c2c %s forward+backward, complex64 input U[-1,1)^2 + 1e-3-RMS 1/|k|^2 "
                       "component (seed %d)" % ("x".join(map(str, n)), seed),
           "n": n, "steps": steps, "forms": []}
    ref = None
    if reference:
        import scipy.fft as sfft
        t0 = time.perf_counter()
        ref = sfft.fftn(x.cpu().numpy().astype(np.complex128), workers=cores)
        res["reference"] = ("scipy.fft.fftn FP64 (pocketfft) of the same input on %d host "
                            "threads, %.1f s" % (cores, time.perf_counter() - t0))
    flops = 2 * 5.0 * N * math.log2(N)
    forms = (("fp64 plain", "f64", 0), ("fp32 plain", "f32", 0),
             ("fp32 mfma radix-16", "f32", P.AMD_MFMA_F32),
             ("split-bf16 mfma", "f32", P.AMD_MFMA_BF16X2))
    ms64 = None
    for form, prec, extra in forms:
        xi = x.to(torch.complex128) if prec == "f64" else x
        y = torch.empty_like(xi)
        z = torch.empty_like(xi)
        fw = P.Plan("c2c", n, xi.data_ptr(), y.data_ptr(), comm=comm, prec=prec, flags=extra)
        bw = P.Plan("c2c", n, y.data_ptr(), z.data_ptr(), sign=+1, comm=comm, prec=prec,
                    flags=extra)
        if not (fw.ok and bw.ok):
            res["forms"].append({"form": form, "status": "plan not available"})
            continue
        ms = _time(lambda: (fw.execute(), bw.execute()), steps)
        passes = _pass_times(fw, xi.data_ptr(), y.data_ptr())
        fw.execute()
        bw.execute()
        _sync()
        ent = {"form": form, "ms_fwd_bwd": round(ms, 3),
               "GFLOPs": round(flops / (ms * 1e-3) / 1e9, 1),
               "GBps_alg": round(2 * 3 * 2 * N * xi.element_size() / (ms * 1e-3) / 1e9, 1),
               "passes_fwd": passes}
        if prec == "f64":
            ms64 = ms
        elif ms64:
            ent["speedup_vs_fp64"] = round(ms64 / ms, 3)
        if ref is not None:
            e2, emax = _err_vs(ref, y)
            ent["rel_l2_fwd"] = e2
            ent["max_err_fwd"] = emax
        zz = z.reshape(-1)
        xx = xi.reshape(-1)
        num = den = 0.0
        for i in range(0, N, 1 << 24):
            d = zz[i:i + (1 << 24)].to(torch.complex128) / N - xx[i:i + (1 << 24)]
            num += float(d.abs().pow(2).sum())
            den += float(xx[i:i + (1 << 24)].to(torch.complex128).abs().pow(2).sum())
        ent["roundtrip_rel_l2"] = math.sqrt(num / den)
        res["forms"].append(ent)
        fw.destroy()
        bw.destroy()
        del y, z, xi
    return res


def cfg3(args):
    """--steps fwd+bwd pairs per form; --no-cpu skips the host FP64
    reference (for profiler runs)"""
    print(json.dumps(cfg3_forms(args.seed, steps=args.steps, reference=not args.no_cpu)))


def cfg4(args):
    import torch
    dev = torch.device("cuda:0")
    sides = [24, 32, 48, 60, 64, 80, 96, 128]
    rng = np.random.default_rng(args.seed)
    rc, comm = P.create_procmesh([1, 1])
    plans, bufs = [], []
    pts = 0
    for _ in range(384):
        n = [int(rng.choice(sides)) for _ in range(3)]
        pts += int(np.prod(n))
        a = torch.rand(n, device=dev, dtype=torch.complex64)
        b = torch.empty_like(a)
        p = P.Plan("c2c", n, a.data_ptr(), b.data_ptr(), comm=comm, prec="f32")
        assert p.ok
        plans.append(p)
        bufs.append((a, b))
    ms_batch = _time(lambda: P.execute_batch(plans), 10, warm=2)
    ms_loop = _time(lambda: [p.execute() for p in plans], 3)
    by = 16.0 * pts                  # one read + one write of every point
    kernels = {}
    for p in plans:
        for ln in p.describe():
            k = ln.split("kernel=")[-1]
            kernels[k] = kernels.get(k, 0) + 1
    print(json.dumps({"workload": "cfg4 384 uneven c2c complex64 (seed %d)" % args.seed,
                      "points": pts, "compulsory_bytes": by, "alg_bytes_3_sweeps": 3 * by,
                      "ms_batched_graph": round(ms_batch, 3),
                      "ms_per_transform_loop": round(ms_loop, 3),
                      "speedup": round(ms_loop / ms_batch, 2),
                      "GBps_batched_compulsory": round(by / (ms_batch * 1e-3) / 1e9, 1),
                      "GBps_batched_alg": round(3 * by / (ms_batch * 1e-3) / 1e9, 1),
                      "pass_kernels": kernels}))


def cfg5(args):
    import torch
    from pfft_amd.autograd import apply_plan
    dev = torch.device("cuda:0")
    out = {"workload": "cfg5 ni=128^3 -> n=256^3 r2c SHIFTED_IN|SHIFTED_OUT, forward+adjoint"}
    rc, comm = P.create_procmesh([1, 1])
    ni, n = 128, 256
    for prec in ("f64", "f32"):
        rdt = torch.float64 if prec == "f64" else torch.float32
        cdt = torch.complex128 if prec == "f64" else torch.complex64
        x = torch.randn((ni, ni, ni), device=dev, dtype=rdt)
        X = torch.empty((n, n, n // 2 + 1), device=dev, dtype=cdt)
        plan = P.Plan("r2c", [n] * 3, x.data_ptr(), X.data_ptr(), ni=[ni] * 3,
                      flags=P.SHIFTED_IN | P.SHIFTED_OUT, comm=comm, prec=prec)
        k01 = torch.arange(-n // 2, n // 2, device=dev, dtype=torch.float64)
        k2 = torch.arange(-n // 2, 1, device=dev, dtype=torch.float64)
        w = torch.exp(-(k01[:, None, None] ** 2 + k01[None, :, None] ** 2 +
                        k2[None, None, :] ** 2) / (2 * 40.0 ** 2)).to(rdt)
        gX = torch.empty_like(X)
        gx = torch.empty_like(x)

        def fwd_bwd():
            plan.execute_on(x.data_ptr(), X.data_ptr())
            torch.mul(X, 2 * w, out=gX)
            plan.execute_adjoint(gX.data_ptr(), gx.data_ptr())
        ms = _time(fwd_bwd, 10, warm=2)
        msf = _time(lambda: plan.execute_on(x.data_ptr(), X.data_ptr()), 10, warm=2)
        out[prec] = {"ms_forward": round(msf, 3), "ms_forward_plus_adjoint": round(ms, 3)}
        plan.destroy()
    print(json.dumps(out))


def gcell(args):
    """pfft_exchange / pfft_reduce on a 512^3 block with 4 ghost cells per
    side.  Minimal HBM bytes: exchange reads the block and writes the box
    (L + G reals), reduce reads the box and writes the block plus the zeroed
    rest of the box (2 G reals)."""
    import torch
    dev = torch.device("cuda:0")
    rc, comm = P.create_procmesh([1, 1])
    n, g = [512, 512, 512], 4
    out = {"workload": "ghost cells 512^3 block, 4 cells each side", "cases": []}
    for kind, prec, flags in (("r", "f64", 0), ("c", "f64", P.GC_R2C), ("r", "f32", 0)):
        rdt = torch.float64 if prec == "f64" else torch.float32
        pn = list(n)
        tup = 1
        if kind == "c":
            pn[2] = n[2] // 2 + 1
            tup = 2
        L = tup * int(np.prod(pn))
        G = tup * int(np.prod([v + 2 * g for v in pn]))
        buf = torch.randn(G, device=dev, dtype=rdt)
        plan = P.GcPlan(kind, n, buf.data_ptr(), [g] * 3, [g] * 3, flags=flags, comm=comm,
                        prec=prec)
        assert plan.ok
        ms_x = _time(plan.exchange, 10, warm=2)
        ms_r = _time(plan.reduce, 10, warm=2)
        getattr(P.lib(), plan.P + "reset_gctimers")(plan.handle)
        for _ in range(5):
            plan.exchange()
            plan.reduce()
        (xi, xw, xp, xe), (ri, rw, rp, re_) = plan.timers()
        esz = 8 if prec == "f64" else 4
        out["cases"].append({
            "data": "%s %s%s" % ("real" if kind == "r" else "complex", prec,
                                  " r2c" if flags else ""),
            "block_reals": L, "box_reals": G,
            "ms_exchange": round(ms_x, 3), "ms_reduce": round(ms_r, 3),
            "GBps_exchange_min_bytes": round(esz * (L + G) / (ms_x * 1e-3) / 1e9, 1),
            "GBps_reduce_min_bytes": round(esz * 2 * G / (ms_r * 1e-3) / 1e9, 1),
            "device_ms_exchange": {"block_copy": round(1e3 * xp / xi, 3),
                                   "fill": round(1e3 * xe / xi, 3)},
            "device_ms_reduce": {"sum": round(1e3 * re_ / ri, 3),
                                 "compact": round(1e3 * rp / ri, 3)}})
        plan.destroy()
        del buf
    print(json.dumps(out))


def mixed(args):
    import torch
    dev = torch.device("cuda:0")
    rc, comm = P.create_procmesh([1, 1])
    out = {"workload": "mixed-radix 3-d c2c and r2c forward", "cases": []}
    for n in ([480, 480, 480], [384, 384, 384], [240, 240, 240], [960, 960, 64]):
        for prec in ("f64", "f32"):
            cdt = torch.complex128 if prec == "f64" else torch.complex64
            g = torch.Generator(device=dev).manual_seed(args.seed)
            x = torch.rand(n, device=dev, dtype=cdt, generator=g) * 2 - (1 + 1j)
            y = torch.empty_like(x)
            p = P.Plan("c2c", n, x.data_ptr(), y.data_ptr(), comm=comm, prec=prec)
            assert p.ok
            ms = _time(p.execute, 5, warm=2)
            passes = _pass_times(p, x.data_ptr(), y.data_ptr())
            ref = torch.fft.fftn(x)
            err = (torch.linalg.vector_norm(y - ref) / torch.linalg.vector_norm(ref)).item()
            pts = int(np.prod(n))
            out["cases"].append({
                "kind": "c2c", "n": n, "prec": prec, "ms": round(ms, 3),
                "GFLOPs": round(5 * pts * math.log2(pts) / (ms * 1e-3) / 1e9, 1),
                "GBps_alg": round(3 * 2 * pts * x.element_size() / (ms * 1e-3) / 1e9, 1),
                "rel_l2_vs_torch_fft": float("%.3g" % err), "passes": passes})
            p.destroy()
            del x, y, ref
    for n in ([480, 480, 480], [384, 384, 384]):
        for prec in ("f64", "f32"):
            rdt = torch.float64 if prec == "f64" else torch.float32
            cdt = torch.complex128 if prec == "f64" else torch.complex64
            g = torch.Generator(device=dev).manual_seed(args.seed)
            x = torch.rand(n, device=dev, dtype=rdt, generator=g) * 2 - 1
            y = torch.empty(n[:-1] + [n[-1] // 2 + 1], device=dev, dtype=cdt)
            p = P.Plan("r2c", n, x.data_ptr(), y.data_ptr(), comm=comm, prec=prec)
            assert p.ok
            ms = _time(p.execute, 5, warm=2)
            passes = _pass_times(p, x.data_ptr(), y.data_ptr())
            ref = torch.fft.rfftn(x)
            err = (torch.linalg.vector_norm(y - ref) / torch.linalg.vector_norm(ref)).item()
            pts = int(np.prod(n))
            out["cases"].append({
                "kind": "r2c", "n": n, "prec": prec, "ms": round(ms, 3),
                "GFLOPs": round(2.5 * pts * math.log2(pts) / (ms * 1e-3) / 1e9, 1),
                "rel_l2_vs_torch_fft": float("%.3g" % err), "passes": passes})
            p.destroy()
            del x, y, ref
    print(json.dumps(out))


def pair(args):
    """The 512^3 pair (r2c TRANSPOSED_OUT + c2r TRANSPOSED_IN) in FP64 and
    the three complex64 forms, without the CPU legs (for profiler runs):
    ms per pair and per pass, errors against the FP64 path's forward."""
    import ctypes
    import bench
    L = P.lib()
    L.pfft_amd_time_passes.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                       ctypes.c_int, ctypes.POINTER(ctypes.c_double),
                                       ctypes.POINTER(ctypes.c_double)]
    L.pfft_amd_plan_npasses.argtypes = [ctypes.c_void_p]
    n = args.n
    N = n ** 3
    rc, comm = P.create_procmesh([1, 1])
    mem = P.local_size("r2c", [n] * 3, flags=P.TRANSPOSED_OUT, comm=comm)[0]
    x = np.random.default_rng(args.seed).standard_normal(N)
    dx, dk, dy = P.DeviceBuffer(N * 8), P.DeviceBuffer(mem * 16), P.DeviceBuffer(2 * mem * 8)
    fw = P.Plan("r2c", [n] * 3, dx.ptr, dk.ptr, flags=P.TRANSPOSED_OUT, comm=comm)
    bw = P.Plan("c2r", [n] * 3, dk.ptr, dy.ptr, sign=+1, flags=P.TRANSPOSED_IN, comm=comm)
    dx.upload(x)
    ms64 = _time(lambda: (fw.execute(), bw.execute()), args.steps, warm=args.warmup)
    fw.execute()
    k64 = dk.download(np.float64, 2 * mem)
    for o in (fw, bw):
        o.destroy()
    for b in (dx, dk, dy):
        b.free()
    forms = bench.pair_forms(L, n, x, args.steps, args.warmup, ms64, k64)
    out = bench.pair_form_errors(forms, k64, "the FP64 path's forward output")
    out["fp64_ms_per_pair"] = round(ms64, 4)
    print(json.dumps(out))


def run(args):
    {"cfg3": cfg3, "cfg4": cfg4, "cfg5": cfg5, "gcell": gcell, "mixed": mixed,
     "pair": pair}[args.workload](args)
