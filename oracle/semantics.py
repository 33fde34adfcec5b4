"""Oracle (test infrastructure only): what a PFFT plan computes.

A numpy restatement of the reference's execution semantics on a
one-process mesh, step by step in the reference's order, in long double
(x87, 64-bit mantissa) unless asked otherwise.  Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline use it.

Followed (reference file:line):
  execute_full                       api/api-basic.c:1044-1107
  conjugation / sign swap            kernel/partrafo.c:426-443, api/api-basic.c:279-292
  twiddle_input  (SHIFTED_OUT)       api/api-basic.c:1186-1238 (incl. the temp<n/2 guard)
  twiddle_output (SHIFTED_IN)        api/api-basic.c:1242-1285
  malloc_and_transpose_INT           api/api-basic.c:1167-1183
  step order of the T_OUT / T_IN pipelines
                                     kernel/partrafo-transposed.c:185-291, api-basic.c:1288-1311
  embed / trunc parameters           kernel/ousample.c:86-151 (order of dims), 208-339 (Zl,D,Zr,Pi,Po)
  embed / trunc execution            kernel/ousample.c:391-506
  serial DFTs                        FFTW 3.3.4 guru64 dft/r2c/c2r (kernel/sertrafo.c:489-554):
                                     unnormalised, exp(sign*2*pi*i*j*k/n); r2c forward only,
                                     c2r backward only, c2r ignores Im of DC/Nyquist
  serial r2r                         FFTW guru64 r2r (kernel/sertrafo.c:531-537), one kind per
                                     dimension (api/api-adv.c:141-155): REDFT00..11 and
                                     RODFT00..11 are scipy's unnormalised DCT/DST types I-IV,
                                     DHT = Re F - Im F of the forward DFT F
  memory layouts                     doc/tutorial.tex:304-330 (transposed decomposition)
The DFTs use scipy.fft (pocketfft) in long double; pocketfft is pinned
against brute-force DFT sums in tests/test_oracle_values.py.
"""
import numpy as np
import scipy.fft as sfft

from . import layout as L

FORWARD = -1
BACKWARD = +1

# host threads the DFTs may use (scipy.fft workers); the CPU baseline
# sets this to the core count it reports
WORKERS = 1


class PlanSpec:
    """Parameters of a plan: what PX(plan_many_dft*) receives."""

    def __init__(self, kind, n, ni=None, no=None, howmany=1, sign=FORWARD, flags=0,
                 rnk_pm=2, skip=None, kinds=None):
        assert kind in ("c2c", "r2c", "c2r", "r2r")
        assert (kind == "r2r") == (kinds is not None)
        self.kind = kind
        self.kinds = list(kinds) if kinds is not None else None
        self.n = list(n)
        self.ni = list(ni if ni is not None else n)
        self.no = list(no if no is not None else n)
        self.howmany = howmany
        self.sign = sign
        self.flags = flags
        self.rnk_pm = rnk_pm
        self.skip = list(skip) if skip is not None else [0] * len(self.n)
        tf = {"c2c": L.T_C2C, "r2c": L.T_R2C, "c2r": L.T_C2R, "r2r": L.T_R2R}[kind]
        if kind != "c2c" and (flags & L.PADDED_R2C):
            tf |= L.T_PADDED
        self.tf = tf
        self.mem, self.lni, self.lis, self.lno, self.los = L.local_size(
            self.n, self.ni, self.no, howmany, None, None, rnk_pm, tf, flags)

    @property
    def rnk(self):
        return len(self.n)

    @property
    def r(self):
        return L.mesh_rank(self.rnk, self.rnk_pm)

    def in_order(self):
        if self.flags & L.TRANSPOSED_IN:
            return L.transposed_order(self.rnk, self.r)
        return list(range(self.rnk))

    def out_order(self):
        if self.flags & L.TRANSPOSED_OUT:
            return L.transposed_order(self.rnk, self.r)
        return list(range(self.rnk))

    def in_is_real(self):
        return self.kind in ("r2c", "r2r")

    def out_is_real(self):
        return self.kind in ("c2r", "r2r")

    def in_logical_shape(self):
        s = list(self.ni)
        if self.kind == "c2r":
            s[-1] = self.ni[-1] // 2 + 1
        return s

    def out_logical_shape(self):
        s = list(self.no)
        if self.kind == "r2c":
            s[-1] = self.no[-1] // 2 + 1
        return s

    def in_memory_shape(self):
        """Extents in memory order (with r2c padding), then howmany."""
        phys = list(self.in_logical_shape())
        if self.kind == "r2c" and (self.tf & L.T_PADDED):
            phys[-1] = 2 * (self.ni[-1] // 2 + 1)
        return [phys[d] for d in self.in_order()] + [self.howmany]

    def out_memory_shape(self):
        phys = list(self.out_logical_shape())
        if self.kind == "c2r" and (self.tf & L.T_PADDED):
            phys[-1] = 2 * (self.no[-1] // 2 + 1)
        return [phys[d] for d in self.out_order()] + [self.howmany]


# ---------------------------------------------------------------- layout io
def memory_to_logical(spec, flat, side):
    """User array (flat numpy, scalar type of that side) -> logical ndarray
    [d0..d_{R-1}, h] (complex values as complex dtype)."""
    if side == "in":
        mshape, order, real = spec.in_memory_shape(), spec.in_order(), spec.in_is_real()
        lshape = spec.in_logical_shape()
    else:
        mshape, order, real = spec.out_memory_shape(), spec.out_order(), spec.out_is_real()
        lshape = spec.out_logical_shape()
    a = np.asarray(flat)
    if not real and not np.iscomplexobj(a):
        a = a[0::2] + 1j * a[1::2] if a.dtype != np.longdouble else _ld_pairs(a)
    a = a[: int(np.prod(mshape))].reshape(mshape)
    inv = np.argsort(order + [len(order)])
    a = np.transpose(a, inv)
    sl = tuple(slice(0, s) for s in lshape) + (slice(None),)
    return a[sl]


def _ld_pairs(a):
    return a[0::2].astype(np.clongdouble) + 1j * a[1::2].astype(np.clongdouble)


def logical_to_memory(spec, arr, side, dtype=np.float64):
    """Logical ndarray -> flat user array of `dtype` (real scalars; complex
    stored interleaved).  Padding elements are zero."""
    if side == "in":
        mshape, order, real = spec.in_memory_shape(), spec.in_order(), spec.in_is_real()
    else:
        mshape, order, real = spec.out_memory_shape(), spec.out_order(), spec.out_is_real()
    full_logical = [mshape[order.index(d)] for d in range(len(order))] + [mshape[-1]]
    buf = np.zeros(full_logical, dtype=(arr.dtype if not real else np.real(arr).dtype))
    sl = tuple(slice(0, s) for s in arr.shape)
    buf[sl] = arr if not real else np.real(arr)
    buf = np.transpose(buf, order + [len(order)])
    buf = np.ascontiguousarray(buf)
    if real:
        return buf.reshape(-1).astype(dtype)
    out = np.empty(buf.size * 2, dtype=dtype)
    out[0::2] = buf.real.reshape(-1)
    out[1::2] = buf.imag.reshape(-1)
    return out


# ---------------------------------------------------------------- twiddles
def _transpose_list(v, rnk, r, transp):
    """api/api-basic.c:1149-1183"""
    if not transp:
        return list(v)
    nt = [0] * rnk
    for t in range(r):
        nt[t] = v[t + 1]
    nt[r] = v[0]
    for t in range(r + 1, rnk):
        nt[t] = v[t]
    return nt


def twiddle_input_factors(spec):
    """Per-element +-1 factors of twiddle_input over the input array in
    memory order (without the howmany axis)."""
    transp = bool(spec.flags & L.TRANSPOSED_IN)
    R, r = spec.rnk, spec.r
    n = _transpose_list(spec.n, R, r, transp)
    ni = _transpose_list(spec.ni, R, r, transp)
    lni = _transpose_list(spec.lni, R, r, transp)
    lis = _transpose_list(spec.lis, R, r, transp)
    skip = _transpose_list(spec.skip, R, r, transp)
    both = bool(spec.flags & L.SHIFTED_IN)
    facs = []
    for t in range(R):
        f = np.ones(lni[t])
        for kt in range(lni[t]):
            temp = kt + lis[t]
            if temp < L.cdiv(ni[t], 2) and not skip[t]:
                v = 1.0
                if L.cmod(temp, 2):
                    v *= -1.0
                if both and L.cmod(L.cdiv(n[t], 2), 2):
                    v *= -1.0
                f[kt] = v
        facs.append(f)
    return facs


def twiddle_output_factors(spec):
    transp = bool(spec.flags & L.TRANSPOSED_OUT)
    R, r = spec.rnk, spec.r
    no = _transpose_list(spec.no, R, r, transp)
    lno = _transpose_list(spec.lno, R, r, transp)
    los = _transpose_list(spec.los, R, r, transp)
    skip = _transpose_list(spec.skip, R, r, transp)
    facs = []
    for t in range(R):
        f = np.ones(lno[t])
        for kt in range(lno[t]):
            temp = kt + los[t]
            if temp < L.cdiv(no[t], 2) and not skip[t]:
                if L.cmod(temp, 2):
                    f[kt] = -1.0
        facs.append(f)
    return facs


def _apply_memory_factors(spec, arr, facs, side):
    """Multiply logical array by factors given in that side's memory order."""
    order = spec.in_order() if side == "in" else spec.out_order()
    for k, d in enumerate(order):
        f = facs[k][: arr.shape[d]]
        shape = [1] * arr.ndim
        shape[d] = f.size
        g = np.ones(arr.shape[d])
        g[: f.size] = f
        shape[d] = arr.shape[d]
        arr = arr * g.reshape(shape).astype(np.longdouble)
    return arr


# ---------------------------------------------------------------- ousam
def ousam_1d_params(n1i, n1o, tf, si_flag, embed):
    """plan_ousam_1d (kernel/ousample.c:208-339) without the howmany
    scaling: returns (Zl, D, Zr) in element units of the logical line, or
    None where the reference returns an empty plan."""
    pni = L.physical_dft_size_1d(n1i, tf)
    pno = L.physical_dft_size_1d(n1o, tf)
    if n1i == n1o and not (tf & L.T_R2C and embed) and not (tf & L.T_C2R and not embed):
        return None
    if embed:
        Zl, D, Zr = 0, n1i, n1o - n1i
        if si_flag & L.SHIFTED_IN:
            Zl = Zr = Zr // 2
    else:
        Zl, D, Zr = 0, n1o, n1i - n1o
        if si_flag & L.SHIFTED_OUT:
            Zl = Zr = Zr // 2
    if tf & L.T_R2C:
        if not embed:
            Zl, D, Zr = pni - pno, pno, 0
        else:
            Zl, D, Zr = 0, n1i, n1o - n1i
            if si_flag & L.SHIFTED_IN:
                Zl = Zr = Zr // 2
    if tf & L.T_C2R:
        if embed:
            Zl, D, Zr = pno - pni, pni, 0
        else:
            Zl, D, Zr = 0, n1o, n1i - n1o
            if si_flag & L.SHIFTED_OUT:
                Zl = Zr = Zr // 2
    return Zl, D, Zr


def embed_axis(arr, axis, params):
    if params is None:
        return arr
    Zl, D, Zr = params
    if Zl == 0 and Zr == 0 and D == arr.shape[axis]:
        return arr
    a = np.take(arr, np.arange(D), axis=axis)
    pad = [(0, 0)] * arr.ndim
    pad[axis] = (Zl, Zr)
    return np.pad(a, pad)


def trunc_axis(arr, axis, params):
    if params is None:
        return arr
    Zl, D, Zr = params
    if Zl == 0 and D == arr.shape[axis]:
        return arr
    return np.take(arr, np.arange(Zl, Zl + D), axis=axis)


# ---------------------------------------------------------------- DFTs
def dft_c2c(arr, axis, sign):
    if sign == FORWARD:
        return sfft.fft(arr, axis=axis, workers=WORKERS)
    return sfft.ifft(arr, axis=axis, workers=WORKERS) * arr.shape[axis]


def dft_r2c(arr, axis, n):
    """FFTW r2c: real input of length n, forward sign, n//2+1 outputs."""
    return sfft.rfft(np.real(arr), n=n, axis=axis, workers=WORKERS)


def dft_c2r(arr, axis, n):
    """FFTW c2r: n//2+1 complex inputs, backward sign, n real outputs."""
    return sfft.irfft(arr, n=n, axis=axis, workers=WORKERS) * n


# FFTW r2r kind -> (scipy transform, type)
R2R_KINDS = {3: ("dct", 1), 4: ("dct", 3), 5: ("dct", 2), 6: ("dct", 4),
             7: ("dst", 1), 8: ("dst", 3), 9: ("dst", 2), 10: ("dst", 4)}
DHT = 2


def dft_r2r(arr, axis, kind):
    """FFTW's unnormalised r2r of `kind` along `axis` (real in, real out)."""
    a = np.real(arr)
    if kind == DHT:
        f = sfft.fft(a, axis=axis, workers=WORKERS)
        return np.real(f) - np.imag(f)
    fn, typ = R2R_KINDS[kind]
    if fn == "dct":
        return sfft.dct(a, type=typ, axis=axis, workers=WORKERS)
    return sfft.dst(a, type=typ, axis=axis, workers=WORKERS)


# ---------------------------------------------------------------- execute
def execute(spec, x_in, precision=np.longdouble):
    """Run the plan on the user input `x_in` (flat array in the input's
    memory layout, interleaved complex) and return the user output as a
    flat float64 array in the output's memory layout."""
    cdt = np.clongdouble if precision == np.longdouble else np.complex128
    rdt = np.longdouble if precision == np.longdouble else np.float64
    x = np.asarray(x_in)
    if spec.in_is_real():
        a = memory_to_logical(spec, x.astype(rdt), "in").astype(rdt)
    else:
        a = memory_to_logical(spec, x.astype(rdt), "in").astype(cdt)
    R, r = spec.rnk, spec.r
    flags = spec.flags
    sign = spec.sign
    conj_in = conj_out = False
    # kernel/partrafo.c:426-443
    if sign == FORWARD and spec.kind == "c2r":
        conj_in, sign = True, BACKWARD
    elif sign == BACKWARD and spec.kind == "r2c":
        conj_out, sign = True, FORWARD
    if conj_in:
        a = np.conj(a)
    if flags & L.SHIFTED_OUT:
        a = _apply_memory_factors(spec, a, twiddle_input_factors(spec), "in")
    si = flags & (L.SHIFTED_IN | L.SHIFTED_OUT)

    # The dimensions each executed outrafo covers and the flags it uses.
    transp = flags & (L.TRANSPOSED_IN | L.TRANSPOSED_OUT)
    to, ti = L.init_param_size_and_trafo_flags(spec.n, spec.ni, spec.no, r, spec.tf, transp,
                                               spec.skip if any(spec.skip) else None)
    steps = []
    if not transp & L.TRANSPOSED_IN:
        n_, ni_, no_, tfs = to
        steps.append((list(range(r, R)), n_, ni_, no_, tfs[r]))
        for t in range(r):
            d = r - 1 - t
            steps.append(([d], n_, ni_, no_, tfs[d]))
    if not transp & L.TRANSPOSED_OUT:
        n_, ni_, no_, tfs = ti
        for t in range(r):
            d = t
            steps.append(([d], n_, ni_, no_, tfs[d]))
        steps.append((list(range(r, R)), n_, ni_, no_, tfs[r]))

    for dims, n_, ni_, no_, tf in steps:
        if tf & L.T_PHANTOM:
            continue
        # embed: dims from last to first (kernel/ousample.c:112-145)
        for d in reversed(dims):
            tfd = tf if (d == R - 1 or not (tf & L.T_RDFT)) else L.T_C2C
            a = embed_axis(a, d, ousam_1d_params(ni_[d], n_[d], tfd, si, True))
        if not (tf & L.T_SKIP):
            # multi-dim serial trafo: r2c transforms the last dim first,
            # c2r last (FFTW rdft2 semantics)
            for d in (reversed(dims) if (tf & L.T_R2C) else dims):
                last = d == R - 1
                if tf & L.T_R2R:
                    a = dft_r2r(a, d, spec.kinds[d])
                elif (tf & L.T_R2C) and last:
                    a = dft_r2c(a, d, n_[d])
                elif (tf & L.T_C2R) and last:
                    a = dft_c2r(a, d, n_[d])
                else:
                    a = dft_c2c(a.astype(cdt), d, sign)
        for d in dims:
            tfd = tf if (d == R - 1 or not (tf & L.T_RDFT)) else L.T_C2C
            a = trunc_axis(a, d, ousam_1d_params(n_[d], no_[d], tfd, si, False))
    if flags & L.SHIFTED_IN:
        a = _apply_memory_factors(spec, a, twiddle_output_factors(spec), "out")
    if conj_out:
        a = np.conj(a)
    return a


def execute_flat(spec, x_in, precision=np.longdouble):
    """execute() returning the user output as a flat float64 array."""
    return logical_to_memory(spec, execute(spec, x_in, precision), "out")
