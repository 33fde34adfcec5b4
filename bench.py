"""Benchmark of the PFFT transform pipeline on one MI355X.

Workload (the headline configuration): 512^3 FP64 real-to-complex forward
transform with PFFT_TRANSPOSED_OUT followed by the complex-to-real
backward transform with PFFT_TRANSPOSED_IN, out-of-place, on a 1x1
process mesh, input i.i.d. standard normal (seeded), already resident in
HBM.  One step = one r2c + c2r pair.

Prints one JSON line (rank 0, N=1): throughput in GFLOP/s counted as
5*N*log2(N)*0.5 per r2c or c2r transform, the per-pair time, the roofline
of the dominant kernel (algorithmic bytes / measured duration vs the
8 TB/s HBM peak, HIP events on the library stream) beside the attainable
bandwidth a streaming probe measures on the same GPU, the precision
comparison ("mixed": the 1024x1024x512 c2c study in FP64, FP32, FP32-MFMA
and split-BF16 with errors against an FP64 reference) and a CPU baseline
(labelled stand-ins timed on this GPU's share of the host cores).
"""
import argparse
import ctypes
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("512³ r2c+c2r ms & %HBM peak; FP64 plain vs FP32-MFMA at stated rel-L2 error")
HBM_PEAK_GBS = 8000.0  # MI355X published HBM3E bandwidth


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-mixed", action="store_true",
                    help="skip the precision comparison (FP64 vs FP32-MFMA / split-BF16)")
    ap.add_argument("--cpu-cores", type=int, default=0)
    ap.add_argument("--seed", type=int, default=20261202)
    ap.add_argument("--workload", default="cfg2",
                    choices=["cfg2", "cfg3", "cfg4", "cfg5", "gcell", "mixed", "pair"],
                    help="cfg2 (default): 512^3 FP64 r2c+c2r; cfg3: precision study "
                         "1024x1024x512 c2c; cfg4: 384 uneven c2c batch; cfg5: oversampled "
                         "shifted r2c forward+adjoint; gcell: ghost cells; mixed: "
                         "mixed-radix c2c sides 240..960; pair: the 512^3 pair in FP64 and "
                         "the complex64 forms (no CPU legs)")
    return ap.parse_args()


TRAFFIC_SUMMARY = os.path.join(ROOT, "profiles", "r2c_c2r_512_fp64_hbm_traffic.md")


def traffic_from_profiles(plan, npass):
    """HBM bytes per launch of pass `npass` of plan `plan` ("r2c"/"c2r")
    from the committed rocprofv3 counter summary (FETCH_SIZE and
    WRITE_SIZE per dispatch, scaled by the same counters on a copy kernel
    of known size), or None when the summary has no such row."""
    if not os.path.exists(TRAFFIC_SUMMARY):
        return None
    with open(TRAFFIC_SUMMARY) as f:
        for line in f:
            cells = [c.strip() for c in line.strip().strip("|").split("|")]
            if len(cells) >= 7 and cells[0] == plan and cells[1] == str(npass):
                try:
                    return int(float(cells[6]))
                except ValueError:
                    return None
    return None


def hbm_probe(L, bytes_per_buffer=1 << 30, iters=3):
    """Attainable HBM bandwidth on this GPU: a streaming kernel swept over
    read-only / write-only / copy, 16 / 32 / 64 bytes per lane and access,
    cached or nontemporal accesses, and five grids (persistent grids of 2, 4,
    8 and 16 workgroups per CU, and one workgroup per chunk).  Returns the
    best configuration of each mix (GB/s of bytes read plus written)."""
    L.pfft_amd_bw_probe.restype = ctypes.c_double
    L.pfft_amd_bw_probe.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_longlong,
                                    ctypes.c_size_t, ctypes.c_int]
    import torch
    ncu = torch.cuda.get_device_properties(0).multi_processor_count
    out = {}
    for mode, name in ((0, "read"), (1, "write"), (2, "copy")):
        best = None
        for width in (16, 32, 64):
            for nt in (0, 1):
                for grid in (2 * ncu, 4 * ncu, 8 * ncu, 16 * ncu, 0):
                    g = L.pfft_amd_bw_probe(mode, width, nt, grid, bytes_per_buffer, iters)
                    if g > 0 and (best is None or g > best["GBps"]):
                        best = {"GBps": round(g, 1), "bytes_per_lane": width,
                                "nontemporal": bool(nt),
                                "grid": "one pass" if grid == 0 else "%d per CU" % (grid // ncu)}
        out[name] = best
    return out


def host_cpu():
    """(threads to use, description): this GPU's share of the host CPU
    (OMP_NUM_THREADS when set, else the CPUs this process may run on) and
    the CPU model."""
    avail = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    share = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or avail
    cores = max(1, min(share, avail))
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return cores, "%s; %d logical CPUs available, %d used (this GPU's share)" % (
        model, avail, cores)


def cpu_baseline(n, cores, cpu_desc, seed):
    """The reference PFFT/FFTW-MPI is not installed on the GPU box (and
    cannot be fetched), so two labelled stand-ins run on the same host
    threads: the oracle (FP64 restatement of the reference's pipeline,
    numpy + scipy.fft pocketfft) and plain scipy.fft rfftn/irfftn."""
    import scipy.fft as sfft
    from oracle import semantics as S
    S.WORKERS = cores
    N = n ** 3
    flops = 5.0 * N * math.log2(N)
    fw = S.PlanSpec("r2c", [n] * 3, flags=2)
    bw = S.PlanSpec("c2r", [n] * 3, sign=+1, flags=1)
    x = np.random.default_rng(seed).standard_normal(N)
    t0 = time.perf_counter()
    k = S.execute_flat(fw, x, precision=np.float64)
    y = S.execute_flat(bw, k, precision=np.float64)
    t1 = time.perf_counter()
    global _ORACLE_FORWARD
    _ORACLE_FORWARD = k
    err = float(np.max(np.abs(y / N - x)))
    xr = x.reshape(n, n, n)
    t2 = time.perf_counter()
    kk = sfft.rfftn(xr, workers=cores)
    yy = sfft.irfftn(kk, s=xr.shape, workers=cores)
    t3 = time.perf_counter()
    err2 = float(np.max(np.abs(yy - xr)))
    return {
        "value": flops / (t1 - t0) / 1e9,
        "unit": "GFLOP/s",
        "cores": cores,
        "kind": "port",
        "cpu": cpu_desc,
        "sample": ("one full %d^3 FP64 r2c(TRANSPOSED_OUT)+c2r(TRANSPOSED_IN) pair through the "
                   "oracle (the reference pipeline restated in numpy + scipy.fft pocketfft, %d "
                   "threads); %.2f s; roundtrip max err %.1e.  The reference PFFT/FFTW-MPI is "
                   "not installed on the box and there is no network, so this is a labelled "
                   "stand-in" % (n, cores, t1 - t0, err)),
        "scipy_fft": {"value": flops / (t3 - t2) / 1e9, "unit": "GFLOP/s", "cores": cores,
                      "sample": "scipy.fft.rfftn + irfftn of the same %d^3 input (untransposed "
                                "layout), %d workers; %.2f s; roundtrip max err %.1e; second "
                                "labelled stand-in" % (n, cores, t3 - t2, err2)},
    }


_ORACLE_FORWARD = None   # the oracle's forward output of the CPU leg (flat scalars)


def pair_forms(L, n, x, steps, warmup, ms64, k64):
    """The headline pair (n^3 r2c TRANSPOSED_OUT + c2r TRANSPOSED_IN) in the
    complex64 forms: plain FP32, FP32 with MFMA radix-16 stages, split-BF16
    MFMA, on the same input rounded to FP32.  Returns per form the ms per
    pair over `steps` pairs, the speed-up over the FP64 pair (`ms64`), the
    per-pass times and kernels, the forward output (host, flat) and the
    roundtrip error.  k64: the FP64 path's forward output (flat)."""
    import pfft_amd as P
    N = n ** 3
    rc, comm = P.create_procmesh([1, 1])
    mem = P.local_size("r2c", [n] * 3, flags=P.TRANSPOSED_OUT, comm=comm)[0]
    x32 = x.astype(np.float32)
    forms = []
    for name, extra in (("fp32 plain", 0), ("fp32 mfma radix-16", P.AMD_MFMA_F32),
                        ("split-bf16 mfma", P.AMD_MFMA_BF16X2)):
        dx = P.DeviceBuffer(N * 4)
        dk = P.DeviceBuffer(mem * 8)
        dy = P.DeviceBuffer(2 * mem * 4)
        fw = P.Plan("r2c", [n] * 3, dx.ptr, dk.ptr, flags=P.TRANSPOSED_OUT | extra, comm=comm,
                    prec="f32")
        bw = P.Plan("c2r", [n] * 3, dk.ptr, dy.ptr, sign=+1, flags=P.TRANSPOSED_IN | extra,
                    comm=comm, prec="f32")
        assert fw.ok and bw.ok
        dx.upload(x32)
        for _ in range(warmup):
            fw.execute()
            bw.execute()
        t0 = time.perf_counter()
        for _ in range(steps):
            fw.execute()
            bw.execute()
        ms = (time.perf_counter() - t0) * 1e3 / steps
        fw.execute()
        kf = dk.download(np.float32, 2 * mem)
        bw.execute()
        y = dy.download(np.float32, N).astype(np.float64)
        passes = []
        for pname, pl, a, b in (("r2c", fw, dx, dk), ("c2r", bw, dk, dy)):
            npass = L.pfft_amd_plan_npasses(pl.handle)
            msv = (ctypes.c_double * npass)()
            byv = (ctypes.c_double * npass)()
            assert L.pfft_amd_time_passes(pl.handle, a.ptr, b.ptr, 10, msv, byv) == 0
            lines = pl.describe()
            for k in range(npass):
                passes.append({"plan": pname, "pass": k, "ms": round(msv[k], 4),
                               "GBps": round(byv[k] / (msv[k] * 1e-3) / 1e9, 1),
                               "kernel": lines[k].split("kernel=")[-1]})
        forms.append({"form": name, "ms_per_pair": round(ms, 4),
                      "speedup_vs_fp64": round(ms64 / ms, 3),
                      "roundtrip_max_err": float(np.max(np.abs(y / N - x))),
                      "passes": passes, "_forward": kf})
        for o in (fw, bw):
            o.destroy()
        for b in (dx, dk, dy):
            b.free()
    return forms


def pair_form_errors(forms, ref, ref_name):
    """rel-L2 and max error (relative to the largest magnitude) of each
    form's forward output against the FP64 reference `ref` (flat)."""
    den = float(np.sqrt(np.dot(ref, ref)))
    mx = float(np.max(np.abs(ref)))
    for f in forms:
        k = f.pop("_forward").astype(np.float64)
        d = k - ref
        f["rel_l2_fwd"] = float(np.sqrt(np.dot(d, d)) / den)
        f["max_err_fwd"] = float(np.max(np.abs(d)) / mx)
    target = [f for f in forms if f["form"] == "fp32 mfma radix-16"][0]
    return {"workload": "the same pair in complex64 (input rounded to FP32): %s" % (
                "plain, FP32 MFMA radix-16 stages, split-BF16 MFMA"),
            "reference": ref_name, "forms": forms,
            "target": {"form": target["form"], "speedup_vs_fp64": target["speedup_vs_fp64"],
                       "rel_l2_fwd": target["rel_l2_fwd"], "goal": ">= 2x at <= 1e-5",
                       "met": target["speedup_vs_fp64"] >= 2.0 and target["rel_l2_fwd"] <= 1e-5}}


def mixed_summary(r):
    """compact form of the precision comparison for the bench line"""
    forms = []
    for f in r["forms"]:
        e = {k: f[k] for k in ("form", "ms_fwd_bwd", "speedup_vs_fp64", "rel_l2_fwd",
                               "max_err_fwd", "roundtrip_rel_l2", "GBps_alg") if k in f}
        if f.get("passes_fwd"):
            e["passes_fwd_ms"] = [p["ms"] for p in f["passes_fwd"]]
            e["kernels"] = [p["kernel"] for p in f["passes_fwd"]]
        forms.append(e)
    out = {"workload": r["workload"], "reference": r.get("reference"), "steps": r["steps"],
           "forms": forms}
    mf = [f for f in forms if f["form"] == "fp32 mfma radix-16"]
    if mf and "speedup_vs_fp64" in mf[0] and "rel_l2_fwd" in mf[0]:
        out["target"] = {"form": "fp32 mfma radix-16", "speedup_vs_fp64": mf[0]["speedup_vs_fp64"],
                         "rel_l2_fwd": mf[0]["rel_l2_fwd"], "goal": ">= 2x at <= 1e-5",
                         "met": mf[0]["speedup_vs_fp64"] >= 2.0 and mf[0]["rel_l2_fwd"] <= 1e-5}
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only (no sharding, no collectives); "
              "use --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.workload != "cfg2":
        import bench_configs
        bench_configs.run(args)
        return
    import pfft_amd as P
    L = P.lib()
    L.pfft_amd_time_passes.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                       ctypes.c_int, ctypes.POINTER(ctypes.c_double),
                                       ctypes.POINTER(ctypes.c_double)]
    L.pfft_amd_plan_npasses.argtypes = [ctypes.c_void_p]
    L.pfft_amd_describe_plan.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_size_t]

    n = args.n
    N = n ** 3
    rc, comm = P.create_procmesh([1, 1])
    mem = P.local_size("r2c", [n] * 3, flags=P.TRANSPOSED_OUT, comm=comm)[0]
    x = np.random.default_rng(args.seed).standard_normal(N)
    dx = P.DeviceBuffer(N * 8)
    dk = P.DeviceBuffer(mem * 16)
    dy = P.DeviceBuffer(2 * mem * 8)
    fw = P.Plan("r2c", [n] * 3, dx.ptr, dk.ptr, flags=P.TRANSPOSED_OUT, comm=comm)
    bw = P.Plan("c2r", [n] * 3, dk.ptr, dy.ptr, sign=+1, flags=P.TRANSPOSED_IN, comm=comm)
    assert fw.ok and bw.ok
    dx.upload(x)

    for _ in range(args.warmup):
        fw.execute()
        bw.execute()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        fw.execute()
        bw.execute()
    t1 = time.perf_counter()
    ms = (t1 - t0) * 1e3 / args.steps
    flops_pair = 5.0 * N * math.log2(N)
    gflops = flops_pair / (ms * 1e-3) / 1e9

    y = dy.download(np.float64, N)
    rt_err = float(np.max(np.abs(y / N - x)))
    fw.execute()
    k64 = dk.download(np.float64, 2 * mem)

    # per-pass durations (HIP events on the library stream)
    passes = []
    for name, pl, a, b in (("r2c", fw, dx, dk), ("c2r", bw, dk, dy)):
        npass = L.pfft_amd_plan_npasses(pl.handle)
        msv = (ctypes.c_double * npass)()
        byv = (ctypes.c_double * npass)()
        assert L.pfft_amd_time_passes(pl.handle, a.ptr, b.ptr, 10, msv, byv) == 0
        desc = ctypes.create_string_buffer(4096)
        L.pfft_amd_describe_plan(pl.handle, desc, 4096)
        lines = desc.value.decode().strip().split("\n")
        for k in range(npass):
            passes.append({"plan": name, "pass": k, "ms": msv[k], "bytes": byv[k],
                           "GBps": byv[k] / (msv[k] * 1e-3) / 1e9, "desc": lines[k]})
    dom = max(passes, key=lambda p: p["ms"])
    probe = hbm_probe(L)
    copy_gbs = probe["copy"]["GBps"]
    sum_ms = sum(p["ms"] for p in passes)
    bytes_pair = sum(p["bytes"] for p in passes)

    traffic = traffic_from_profiles(dom["plan"], dom["pass"])
    roofline = {
        "bound": "hbm",
        "achieved": round(dom["GBps"], 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(dom["GBps"] / HBM_PEAK_GBS, 4),
        "traffic": traffic,
        "kernel": "%s pass %d (%s)" % (dom["plan"], dom["pass"], dom["desc"]),
        "algorithmic_bytes_per_launch": dom["bytes"],
        "avg_launch_ms": round(dom["ms"], 4),
    }
    out = {
        "metric": METRIC,
        "value": round(gflops, 2),
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: i.i.d. N(0,1) real input, seed %d" % args.seed,
        "config": {"workload": "%d^3 r2c(PFFT_TRANSPOSED_OUT) + c2r(PFFT_TRANSPOSED_IN), FP64, "
                               "out-of-place, procmesh 1x1" % n,
                   "n": [n, n, n], "flops_per_step": flops_pair,
                   "input": "PFFT_PRESERVE_INPUT (default; plans never write their input, so "
                            "PFFT_DESTROY_INPUT runs the same kernels)",
                   "alg_bytes_per_step": bytes_pair},
        "roofline": roofline,
        "pipeline": {
            "sum_of_pass_ms": round(sum_ms, 4),
            "pipeline_GBps_alg": round(bytes_pair / (sum_ms * 1e-3) / 1e9, 1),
            "frac_of_copy_bw": round(bytes_pair / (sum_ms * 1e-3) / 1e9 / copy_gbs, 4),
            "measured_copy_GBps": round(copy_gbs, 1),
            "hbm_probe": probe,
            "passes": [{k: (round(v, 4) if isinstance(v, float) else v) for k, v in p.items()}
                       for p in passes],
        },
        "roundtrip_max_err": rt_err,
    }
    for o in (fw, bw):
        o.destroy()
    for b in (dx, dk, dy):
        b.free()
    cores, cpu_desc = host_cpu()
    if args.cpu_cores:
        cores = args.cpu_cores
    forms = None
    if not args.no_mixed:
        forms = pair_forms(L, n, x, args.steps, args.warmup, ms, k64)
        import bench_configs
        out["precision_study"] = mixed_summary(
            bench_configs.cfg3_forms(args.seed, steps=5, cores=cores))
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(n, cores, cpu_desc, args.seed)
    else:
        out["cpu_baseline"] = None
    if forms is not None:
        if _ORACLE_FORWARD is not None:
            ref, ref_name = _ORACLE_FORWARD, "the FP64 oracle's forward output (CPU leg)"
        else:
            ref, ref_name = k64, ("the FP64 path's forward output (within 1e-12 of the oracle "
                                  "at this size, tests/test_gpu_parity.py)")
        out["mixed"] = pair_form_errors(forms, ref, ref_name)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
