"""pfft_amd -- Python host mirror of the pfft.h C ABI (ctypes).

The product is the native library pfft_amd/lib/libpfft.so (C host code +
HIP kernels for gfx950).  This module only binds it: every call goes
through the exported pfft_* / pfftf_* symbols declared in include/pfft.h.
If the library is missing this module raises at import; there is no
fallback implementation.

Flag values and entry-point names are the reference's (api/pfft.h).
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("PFFT_AMD_LIB") or os.path.join(HERE, "lib", "libpfft.so")

# api/pfft.h:528-545
TRANSPOSED_NONE = 0
TRANSPOSED_IN = 1 << 0
TRANSPOSED_OUT = 1 << 1
SHIFTED_NONE = 0
SHIFTED_IN = 1 << 2
SHIFTED_OUT = 1 << 3
MEASURE = 0
ESTIMATE = 1 << 4
# r2r kinds (FFTW numbering)
DHT, REDFT00, REDFT01, REDFT10, REDFT11 = 2, 3, 4, 5, 6
RODFT00, RODFT01, RODFT10, RODFT11 = 7, 8, 9, 10
PATIENT = 1 << 5
EXHAUSTIVE = 1 << 6
TUNE = 1 << 7
PRESERVE_INPUT = 1 << 8
DESTROY_INPUT = 1 << 9
BUFFERED_INPLACE = 1 << 10
PADDED_R2C = 1 << 11
PADDED_C2R = PADDED_R2C
AMD_MFMA_F32 = 1 << 12
AMD_MFMA_BF16X2 = 1 << 13
FORWARD = -1
BACKWARD = +1
# ghost-cell flags: api/pfft.h:563-572
GC_TRANSPOSED_NONE = 0
GC_TRANSPOSED = 1 << 0
GC_SENDRECV = 1 << 1
GC_RMA = 1 << 2
GC_R2C = 1 << 3
GC_PADDED = 1 << 4
GC_C2R = GC_R2C
GC_PADDED_R2C = GC_R2C | GC_PADDED
GC_PADDED_C2R = GC_PADDED_R2C
DEFAULT_BLOCK = 0
COMM_WORLD = 0x44000000

_lib = None
INT = ctypes.c_ssize_t
INTP = ctypes.POINTER(INT)
VP = ctypes.c_void_p


class NativeLibraryMissing(RuntimeError):
    pass


def lib():
    """Load (once) and return the native library."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise NativeLibraryMissing(
                "pfft_amd native library not built: %s (run python -m pfft_amd.build)" % LIB_PATH)
        _lib = ctypes.CDLL(LIB_PATH, mode=ctypes.RTLD_GLOBAL)
        _declare(_lib)
    return _lib


def _declare(L):
    for P in ("pfft_", "pfftf_"):
        for nm in ("local_size_many_dft", "local_size_many_dft_r2c", "local_size_many_dft_c2r",
                   "local_size_many_r2r"):
            f = getattr(L, P + nm)
            f.restype = INT
            f.argtypes = [ctypes.c_int, INTP, INTP, INTP, INT, INTP, INTP, ctypes.c_int,
                          ctypes.c_uint, INTP, INTP, INTP, INTP]
        for nm in ("local_size_dft_3d", "local_size_dft_r2c_3d", "local_size_dft_c2r_3d"):
            f = getattr(L, P + nm)
            f.restype = INT
            f.argtypes = [INTP, ctypes.c_int, ctypes.c_uint, INTP, INTP, INTP, INTP]
        for nm in ("local_block_many_dft", "local_block_many_dft_r2c",
                   "local_block_many_dft_c2r", "local_block_many_r2r"):
            f = getattr(L, P + nm)
            f.restype = None
            f.argtypes = [ctypes.c_int, INTP, INTP, INTP, INTP, ctypes.c_int, ctypes.c_int,
                          ctypes.c_uint, INTP, INTP, INTP, INTP]
        for nm in ("plan_many_dft", "plan_many_dft_r2c", "plan_many_dft_c2r"):
            f = getattr(L, P + nm)
            f.restype = VP
            f.argtypes = [ctypes.c_int, INTP, INTP, INTP, INT, INTP, INTP, VP, VP,
                          ctypes.c_int, ctypes.c_int, ctypes.c_uint]
        for nm in ("plan_many_dft_skipped", "plan_many_dft_r2c_skipped",
                   "plan_many_dft_c2r_skipped"):
            f = getattr(L, P + nm)
            f.restype = VP
            f.argtypes = [ctypes.c_int, INTP, INTP, INTP, INT, INTP, INTP,
                          ctypes.POINTER(ctypes.c_int), VP, VP, ctypes.c_int, ctypes.c_int,
                          ctypes.c_uint]
        f = getattr(L, P + "plan_many_r2r")
        f.restype = VP
        f.argtypes = [ctypes.c_int, INTP, INTP, INTP, INT, INTP, INTP, VP, VP, ctypes.c_int,
                      ctypes.POINTER(ctypes.c_int), ctypes.c_uint]
        f = getattr(L, P + "plan_many_r2r_skipped")
        f.restype = VP
        f.argtypes = [ctypes.c_int, INTP, INTP, INTP, INT, INTP, INTP,
                      ctypes.POINTER(ctypes.c_int), VP, VP, ctypes.c_int,
                      ctypes.POINTER(ctypes.c_int), ctypes.c_uint]
        for nm in ("execute",):
            getattr(L, P + nm).argtypes = [VP]
            getattr(L, P + nm).restype = None
        for nm in ("execute_dft", "execute_dft_r2c", "execute_dft_c2r", "execute_r2r",
                   "execute_adjoint"):
            getattr(L, P + nm).argtypes = [VP, VP, VP]
            getattr(L, P + nm).restype = None
        getattr(L, P + "execute_batch").argtypes = [ctypes.POINTER(VP), ctypes.c_int]
        getattr(L, P + "execute_batch").restype = None
        getattr(L, P + "destroy_plan").argtypes = [VP]
        getattr(L, P + "destroy_plan").restype = None
        getattr(L, P + "malloc").argtypes = [ctypes.c_size_t]
        getattr(L, P + "malloc").restype = VP
        for nm in ("alloc_real", "alloc_complex"):
            getattr(L, P + nm).argtypes = [INT]
            getattr(L, P + nm).restype = VP
        getattr(L, P + "free").argtypes = [VP]
        getattr(L, P + "free").restype = None
        getattr(L, P + "init").restype = None
        for nm, rt in (("create_procmesh", ctypes.c_int),):
            f = getattr(L, P + nm)
            f.restype = rt
            f.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                          ctypes.POINTER(ctypes.c_int)]
        f = getattr(L, P + "create_procmesh_2d")
        f.restype = ctypes.c_int
        f.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
        f = getattr(L, P + "create_procmesh_1d")
        f.restype = ctypes.c_int
        f.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
        for nm in ("init_input_complex", "init_input_complex_hermitian", "init_input_real",
                   "clear_input_complex", "clear_input_complex_hermitian", "clear_input_real"):
            f = getattr(L, P + nm)
            f.restype = None
            f.argtypes = [ctypes.c_int, INTP, INTP, INTP, VP]
        for nm in ("check_output_complex", "check_output_complex_hermitian",
                   "check_output_real"):
            f = getattr(L, P + nm)
            f.restype = ctypes.c_double if P == "pfft_" else ctypes.c_float
            f.argtypes = [ctypes.c_int, INTP, INTP, INTP, VP, ctypes.c_int]
        f = getattr(L, P + "local_size_many_gc")
        f.restype = INT
        f.argtypes = [ctypes.c_int, INTP, INTP, INT, INTP, INTP, INTP, INTP]
        for nm in ("plan_many_rgc", "plan_many_cgc"):
            f = getattr(L, P + nm)
            f.restype = VP
            f.argtypes = [ctypes.c_int, INTP, INT, INTP, INTP, INTP, VP, ctypes.c_int,
                          ctypes.c_uint]
        for nm in ("exchange", "reduce", "destroy_gcplan", "reset_gctimers", "destroy_gctimer",
                   "average_gctimer"):
            getattr(L, P + nm).argtypes = [VP]
            getattr(L, P + nm).restype = None
        for nm in ("get_gctimer_exg", "get_gctimer_red", "copy_gctimer"):
            getattr(L, P + nm).argtypes = [VP]
            getattr(L, P + nm).restype = VP
        getattr(L, P + "convert_gctimer2vec").argtypes = [VP, ctypes.POINTER(ctypes.c_double)]
        getattr(L, P + "convert_gctimer2vec").restype = None
        for nm in ("print_average_gctimer", "print_average_gctimer_adv"):
            getattr(L, P + nm).argtypes = [VP, ctypes.c_int]
            getattr(L, P + nm).restype = None
        for nm in ("write_average_gctimer", "write_average_gctimer_adv"):
            getattr(L, P + nm).argtypes = [VP, ctypes.c_char_p, ctypes.c_int]
            getattr(L, P + nm).restype = None
    L.pfft_amd_memcpy_to_device.argtypes = [VP, VP, ctypes.c_size_t]
    L.pfft_amd_memcpy_to_device.restype = ctypes.c_int
    L.pfft_amd_memcpy_to_host.argtypes = [VP, VP, ctypes.c_size_t]
    L.pfft_amd_memcpy_to_host.restype = ctypes.c_int
    L.pfft_amd_version.restype = ctypes.c_char_p
    L.pfft_amd_last_error.restype = ctypes.c_char_p
    for nm in ("export_wisdom", "import_wisdom"):
        getattr(L, "pfft_amd_" + nm).argtypes = [ctypes.c_char_p]
        getattr(L, "pfft_amd_" + nm).restype = ctypes.c_int
    L.pfft_amd_forget_wisdom.restype = None
    L.pfft_amd_wisdom_count.restype = ctypes.c_int
    L.pfft_amd_tune_timings.restype = ctypes.c_int
    L.pfft_amd_describe_plan.argtypes = [VP, ctypes.c_char_p, ctypes.c_size_t]
    L.pfft_amd_describe_plan.restype = ctypes.c_int


def ivec(v):
    if v is None:
        return None
    a = (INT * len(v))(*[int(x) for x in v])
    return a


def _prefix(prec):
    return "pfft_" if prec in ("f64", "double", np.float64) else "pfftf_"


# ---------------------------------------------------------------- meshes
def create_procmesh(np_dims, comm=COMM_WORLD):
    """pfft_create_procmesh: returns (rc, cart_comm_handle)."""
    arr = (ctypes.c_int * len(np_dims))(*np_dims)
    out = ctypes.c_int(0)
    rc = lib().pfft_create_procmesh(len(np_dims), comm, arr, ctypes.byref(out))
    return rc, out.value


# ---------------------------------------------------------------- layout
def local_size(kind, n, ni=None, no=None, howmany=1, iblock=None, oblock=None,
               comm=COMM_WORLD, flags=0, prec="f64"):
    """pfft_local_size_many_dft{,_r2c,_c2r} / pfft_local_size_many_r2r: returns
    (alloc_local, local_ni, local_i_start, local_no, local_o_start)."""
    rnk = len(n)
    ni = n if ni is None else ni
    no = n if no is None else no
    lni, lis, lno, los = (INT * rnk)(), (INT * rnk)(), (INT * rnk)(), (INT * rnk)()
    name = {"c2c": "local_size_many_dft", "r2c": "local_size_many_dft_r2c",
            "c2r": "local_size_many_dft_c2r", "r2r": "local_size_many_r2r"}[kind]
    f = getattr(lib(), _prefix(prec) + name)
    mem = f(rnk, ivec(n), ivec(ni), ivec(no), howmany, ivec(iblock), ivec(oblock), comm, flags,
            lni, lis, lno, los)
    return int(mem), list(lni), list(lis), list(lno), list(los)


def local_block(kind, ni, no=None, iblock=None, oblock=None, comm=COMM_WORLD, pid=0, flags=0,
                prec="f64"):
    rnk = len(ni)
    no = ni if no is None else no
    lni, lis, lno, los = (INT * rnk)(), (INT * rnk)(), (INT * rnk)(), (INT * rnk)()
    name = {"c2c": "local_block_many_dft", "r2c": "local_block_many_dft_r2c",
            "c2r": "local_block_many_dft_c2r", "r2r": "local_block_many_r2r"}[kind]
    f = getattr(lib(), _prefix(prec) + name)
    f(rnk, ivec(ni), ivec(no), ivec(iblock), ivec(oblock), comm, pid, flags, lni, lis, lno, los)
    return list(lni), list(lis), list(lno), list(los)


# ---------------------------------------------------------------- device memory
class DeviceBuffer:
    """Device memory from pfft_malloc (hipMalloc inside the library)."""

    def __init__(self, nbytes, prec="f64"):
        self.nbytes = int(nbytes)
        self.prec = prec
        self.ptr = lib().pfft_malloc(max(self.nbytes, 1))
        if not self.ptr:
            raise MemoryError("pfft_malloc(%d) failed: %s" % (nbytes, last_error()))

    def upload(self, arr):
        a = np.ascontiguousarray(arr)
        assert a.nbytes <= self.nbytes, (a.nbytes, self.nbytes)
        rc = lib().pfft_amd_memcpy_to_device(self.ptr, a.ctypes.data, a.nbytes)
        if rc:
            raise RuntimeError("upload failed: " + last_error())

    def download(self, dtype, count=None):
        dt = np.dtype(dtype)
        count = self.nbytes // dt.itemsize if count is None else count
        out = np.empty(count, dtype=dt)
        rc = lib().pfft_amd_memcpy_to_host(out.ctypes.data, self.ptr, out.nbytes)
        if rc:
            raise RuntimeError("download failed: " + last_error())
        return out

    def free(self):
        if self.ptr:
            lib().pfft_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def last_error():
    return lib().pfft_amd_last_error().decode()


# ---------------------------------------------------------------- wisdom
def export_wisdom(path):
    """Writes the planner's launch-variant selections to `path` (0 = ok)."""
    return lib().pfft_amd_export_wisdom(os.fsencode(path))


def import_wisdom(path):
    """Merges a wisdom file into the process-wide cache (0 = ok)."""
    return lib().pfft_amd_import_wisdom(os.fsencode(path))


def forget_wisdom():
    lib().pfft_amd_forget_wisdom()


def wisdom_count():
    return lib().pfft_amd_wisdom_count()


def tune_timings():
    """Number of candidate timings the planner has performed."""
    return lib().pfft_amd_tune_timings()


# ---------------------------------------------------------------- plans
class Plan:
    """A pfft plan handle (pfft_plan_many_dft* / pfftf_*)."""

    def __init__(self, kind, n, in_ptr, out_ptr, ni=None, no=None, howmany=1, sign=FORWARD,
                 flags=0, comm=COMM_WORLD, iblock=None, oblock=None, skip=None, prec="f64",
                 kinds=None):
        self.kind = kind
        self.prec = prec
        self.P = _prefix(prec)
        rnk = len(n)
        ni = n if ni is None else ni
        no = n if no is None else no
        L = lib()
        if kind == "r2r":
            kv = (ctypes.c_int * rnk)(*kinds)
            if skip is None:
                self.handle = getattr(L, self.P + "plan_many_r2r")(
                    rnk, ivec(n), ivec(ni), ivec(no), howmany, ivec(iblock), ivec(oblock),
                    in_ptr, out_ptr, comm, kv, flags)
            else:
                sk = (ctypes.c_int * rnk)(*skip)
                self.handle = getattr(L, self.P + "plan_many_r2r_skipped")(
                    rnk, ivec(n), ivec(ni), ivec(no), howmany, ivec(iblock), ivec(oblock), sk,
                    in_ptr, out_ptr, comm, kv, flags)
            return
        suffix = {"c2c": "", "r2c": "_r2c", "c2r": "_c2r"}[kind]
        if skip is None:
            f = getattr(L, self.P + "plan_many_dft" + suffix)
            self.handle = f(rnk, ivec(n), ivec(ni), ivec(no), howmany, ivec(iblock), ivec(oblock),
                            in_ptr, out_ptr, comm, sign, flags)
        else:
            f = getattr(L, self.P + "plan_many_dft" + suffix + "_skipped")
            sk = (ctypes.c_int * rnk)(*skip)
            self.handle = f(rnk, ivec(n), ivec(ni), ivec(no), howmany, ivec(iblock), ivec(oblock),
                            sk, in_ptr, out_ptr, comm, sign, flags)

    @property
    def ok(self):
        return bool(self.handle)

    def execute(self):
        getattr(lib(), self.P + "execute")(self.handle)

    def describe(self):
        """One line per device pass: geometry and the kernel that runs it."""
        buf = ctypes.create_string_buffer(8192)
        lib().pfft_amd_describe_plan(self.handle, buf, 8192)
        return buf.value.decode().strip().split("\n")

    def execute_on(self, in_ptr, out_ptr):
        suffix = {"c2c": "dft", "r2c": "dft_r2c", "c2r": "dft_c2r", "r2r": "r2r"}[self.kind]
        getattr(lib(), self.P + "execute_" + suffix)(self.handle, in_ptr, out_ptr)

    def execute_adjoint(self, in_ptr, out_ptr):
        getattr(lib(), self.P + "execute_adjoint")(self.handle, in_ptr, out_ptr)

    def destroy(self):
        if self.handle:
            getattr(lib(), self.P + "destroy_plan")(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


def exported_symbols():
    """Names of the functions declared in include/pfft.h (parsed)."""
    import re
    hdr = open(os.path.join(os.path.dirname(HERE), "include", "pfft.h")).read()
    names = set()
    m = re.search(r"#define PFFT_AMD_DECLARE_API\(P, R, C, INT\)(.*?)\n\n", hdr, re.S)
    body = m.group(1)
    for fn in re.findall(r"P##(\w+)\(", body):
        if fn.endswith("plan_s"):
            continue
        names.add("pfft_" + fn)
        names.add("pfftf_" + fn)
    rest = hdr[m.end():]
    for fn in re.findall(r"\b(pfftf?_\w+|pfft_amd_\w+)\s*\(", rest):
        names.add(fn)
    return sorted(names)


def execute_batch(plans):
    """pfft_execute_batch / pfftf_execute_batch over Plan objects of one
    precision (persistent launches per pass level, replayed from a captured
    graph)."""
    if not plans:
        return
    arr = (VP * len(plans))(*[p.handle for p in plans])
    getattr(lib(), plans[0].P + "execute_batch")(arr, len(plans))


def execute_batch_noalloc(plans):
    """pfft_amd_execute_batch_noalloc: the batch with no allocation (fixed
    device work-list); returns the C return code (0 on success)."""
    L = lib()
    L.pfft_amd_execute_batch_noalloc.restype = ctypes.c_int
    L.pfft_amd_execute_batch_noalloc.argtypes = [ctypes.c_void_p, ctypes.c_int]
    arr = (VP * max(len(plans), 1))(*[p.handle for p in plans])
    return L.pfft_amd_execute_batch_noalloc(arr, len(plans))


# ---------------------------------------------------------------- ghost cells
def local_size_gc(local_n, local_start, gc_below=None, gc_above=None, howmany=1, prec="f64"):
    """pfft_local_size_many_gc: returns (mem, local_ngc, local_gc_start)."""
    rnk = len(local_n)
    lngc, lgcs = (INT * rnk)(), (INT * rnk)()
    mem = getattr(lib(), _prefix(prec) + "local_size_many_gc")(
        rnk, ivec(local_n), ivec(local_start), howmany, ivec(gc_below), ivec(gc_above),
        lngc, lgcs)
    return int(mem), list(lngc), list(lgcs)


class GcPlan:
    """A ghost-cell plan (pfft_plan_many_rgc / pfft_plan_many_cgc) on
    `data` (real data: kind "r", complex data: kind "c")."""

    def __init__(self, kind, n, data_ptr, gc_below=None, gc_above=None, howmany=1, block=None,
                 flags=0, comm=COMM_WORLD, prec="f64"):
        self.P = _prefix(prec)
        name = {"r": "plan_many_rgc", "c": "plan_many_cgc"}[kind]
        self.handle = getattr(lib(), self.P + name)(
            len(n), ivec(n), howmany, ivec(block), ivec(gc_below), ivec(gc_above), data_ptr,
            comm, flags)

    @property
    def ok(self):
        return bool(self.handle)

    def exchange(self):
        getattr(lib(), self.P + "exchange")(self.handle)

    def reduce(self):
        getattr(lib(), self.P + "reduce")(self.handle)

    def timers(self):
        """(exchange, reduce) timers as [iter, whole, pad_zeros, exchange] lists."""
        out = []
        for nm in ("get_gctimer_exg", "get_gctimer_red"):
            t = getattr(lib(), self.P + nm)(self.handle)
            v = (ctypes.c_double * 4)()
            getattr(lib(), self.P + "convert_gctimer2vec")(t, v)
            getattr(lib(), self.P + "destroy_gctimer")(t)
            out.append(list(v))
        return out

    def destroy(self):
        if self.handle:
            getattr(lib(), self.P + "destroy_gcplan")(self.handle)
            self.handle = None
