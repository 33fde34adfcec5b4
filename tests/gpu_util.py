"""Helpers for the GPU parity tests: run a plan through the C ABI and
compare with the oracle."""
import numpy as np

import pfft_amd as P
from oracle import layout as L
from oracle import semantics as S


def spec_of(kind, n, ni=None, no=None, howmany=1, sign=-1, flags=0, rnk_pm=2, skip=None,
            kinds=None):
    return S.PlanSpec(kind, n, ni, no, howmany, sign, flags, rnk_pm, skip, kinds)


def comm_for(rnk_pm):
    if rnk_pm == 0:
        return P.COMM_WORLD
    rc, c = P.create_procmesh([1] * rnk_pm)
    assert rc == 0
    return c


def n_scalars(spec, side):
    shp = spec.in_memory_shape() if side == "in" else spec.out_memory_shape()
    real = spec.in_is_real() if side == "in" else spec.out_is_real()
    return int(np.prod(shp)) * (1 if real else 2)


def random_input(spec, seed=0):
    rng = np.random.default_rng(seed)
    return rng.uniform(-1.0, 1.0, n_scalars(spec, "in"))


def run_gpu(spec, x, prec="f64", comm=None, adjoint=False, plan_flags_extra=0, inplace=False):
    """Execute `spec` on the GPU; x is the flat input (float64 scalars).
    inplace: one array serves as input and output (sized for the larger)."""
    rdt = np.float64 if prec == "f64" else np.float32
    esz = np.dtype(rdt).itemsize
    nin, nout = n_scalars(spec, "in"), n_scalars(spec, "out")
    if adjoint:
        nin, nout = nout, nin
    if inplace:
        din = P.DeviceBuffer(max(nin, nout, 1) * esz)
        dout = din
    else:
        din = P.DeviceBuffer(max(nin, 1) * esz)
        dout = P.DeviceBuffer(max(nout, 1) * esz)
    if comm is None:
        comm = comm_for(spec.rnk_pm)
    skip = spec.skip if any(spec.skip) else None
    plan = P.Plan(spec.kind, spec.n, din.ptr, dout.ptr, spec.ni, spec.no, spec.howmany,
                  spec.sign, spec.flags | plan_flags_extra, comm=comm, skip=skip, prec=prec,
                  kinds=spec.kinds)
    assert plan.ok, "plan creation failed"
    din.upload(np.asarray(x, dtype=rdt))
    if adjoint:
        plan.execute_adjoint(din.ptr, dout.ptr)
    else:
        plan.execute()
    out = dout.download(rdt, nout).astype(np.float64)
    plan.destroy()
    din.free()
    if not inplace:
        dout.free()
    return out


def oracle_out(spec, x):
    return S.execute_flat(spec, x)


def out_mask(spec):
    """Scalars of the output array that hold data (excludes c2r padding)."""
    m = np.ones(n_scalars(spec, "out"), dtype=bool)
    if spec.kind == "c2r" and (spec.tf & L.T_PADDED):
        shp = spec.out_memory_shape()
        a = np.zeros(shp, dtype=bool)
        sl = [slice(None)] * len(shp)
        sl[-2] = slice(0, spec.no[-1])
        a[tuple(sl)] = True
        m = a.reshape(-1)
    return m


def rel_l2(a, b, mask=None):
    if mask is not None:
        a, b = a[mask], b[mask]
    den = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (den if den > 0 else 1.0))
