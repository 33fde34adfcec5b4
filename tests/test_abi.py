"""The C-ABI library loads and exports every function include/pfft.h
declares (CPU only: no compute calls)."""
import ctypes
import os
import subprocess

import pytest

import pfft_amd as P

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_loads():
    assert os.path.exists(P.LIB_PATH)
    L = P.lib()
    assert L.pfft_amd_version().decode().startswith("pfft-amd")


def test_every_declared_symbol_is_exported():
    names = P.exported_symbols()
    assert len(names) > 140
    L = P.lib()
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_reference_entry_points_present():
    # the pfft.h entry points the reference's tests and users bind
    L = P.lib()
    for pre in ("pfft_", "pfftf_"):
        for nm in ("init", "cleanup", "create_procmesh", "create_procmesh_1d",
                   "create_procmesh_2d", "local_size_dft_3d", "local_size_dft_r2c_3d",
                   "local_size_dft_c2r_3d", "local_size_many_dft", "plan_dft_3d",
                   "plan_dft_r2c_3d", "plan_dft_c2r_3d", "plan_many_dft",
                   "plan_many_dft_r2c", "plan_many_dft_c2r", "plan_many_dft_skipped",
                   "execute", "execute_dft", "execute_dft_r2c", "execute_dft_c2r",
                   "destroy_plan", "alloc_real", "alloc_complex", "malloc", "free",
                   "init_input_complex_3d", "check_output_complex_3d",
                   "init_input_real_3d", "check_output_real_3d", "local_block_dft_3d"):
            assert hasattr(L, pre + nm), pre + nm


def test_header_compiles_as_c_and_links(tmp_path):
    """A C program written against include/pfft.h (reference API style)
    compiles and links against libpfft.so."""
    src = r'''
#include <stdio.h>
#include "pfft.h"
int main(void) {
  ptrdiff_t n[3] = {29, 27, 31}, lni[3], lis[3], lno[3], los[3];
  MPI_Comm cart;
  if (pfft_create_procmesh_2d(MPI_COMM_WORLD, 1, 1, &cart)) return 1;
  ptrdiff_t m = pfft_local_size_dft_3d(n, cart, PFFT_TRANSPOSED_NONE, lni, lis, lno, los);
  printf("%td %td %td %td\n", m, lni[0], lno[2], los[1]);
  return m == 29 * 27 * 31 ? 0 : 2;
}
'''
    c = str(tmp_path / "abi_probe.c")
    exe = str(tmp_path / "abi_probe")
    with open(c, "w") as f:
        f.write(src)
    libdir = os.path.dirname(P.LIB_PATH)
    r = subprocess.run(["gcc", "-std=c99", "-I" + os.path.join(ROOT, "include"), c, "-o", exe,
                        "-L" + libdir, "-lpfft", "-Wl,-rpath," + libdir],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([exe], capture_output=True, text=True)
    assert r.returncode == 0, (r.stdout, r.stderr)
    assert r.stdout.split() == ["24273", "29", "31", "0"]


def test_get_args_and_array_printing(tmp_path):
    """pfft_get_args (util/getargs.c) and pfft_apr_real_3d /
    pfft_apr_complex_permuted_3d (api/api-basic.c:584-710) from a C
    program, on host arrays."""
    src = r'''
#include <stdio.h>
#include "pfft.h"
int main(int argc, char **argv) {
  ptrdiff_t n[3] = {2, 2, 2};
  double tol = 0;
  int sw = 0, miss = 7;
  pfft_get_args(argc, argv, "-pfft_n", 3, PFFT_PTRDIFF_T, n);
  pfft_get_args(argc, argv, "-tol", 1, PFFT_DOUBLE, &tol);
  pfft_get_args(argc, argv, "-verbose", 0, PFFT_SWITCH, &sw);
  pfft_get_args(argc, argv, "-absent", 1, PFFT_INT, &miss);
  printf("%td %td %td %.3f %d %d\n", n[0], n[1], n[2], tol, sw, miss);
  double r[4] = {1, 2, 3, 4};
  ptrdiff_t ln[3] = {1, 2, 2}, ls[3] = {0, 1, 0};
  pfft_apr_real_3d(r, ln, ls, "r", MPI_COMM_WORLD);
  double c[4] = {1, -1, 2, -2};
  ptrdiff_t lc[3] = {2, 1, 1}, lsc[3] = {3, 0, 0};
  pfft_apr_complex_permuted_3d((const pfft_complex *)c, lc, lsc, 1, 2, 0, "c", MPI_COMM_WORLD);
  return 0;
}
'''
    c = str(tmp_path / "args_probe.c")
    exe = str(tmp_path / "args_probe")
    with open(c, "w") as f:
        f.write(src)
    libdir = os.path.dirname(P.LIB_PATH)
    r = subprocess.run(["gcc", "-std=c99", "-I" + os.path.join(ROOT, "include"), c, "-o", exe,
                        "-L" + libdir, "-lpfft", "-Wl,-rpath," + libdir],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([exe, "-pfft_n", "8", "16", "32", "-tol", "0.25", "-verbose"],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    lines = r.stdout.split("\n")
    assert lines[0] == "8 16 32 0.250 1 7"
    assert lines[1] == "Rank 0:"
    assert lines[2] == "r(0,1:2,0:1):"
    assert lines[3].split() == ["1.00e+00,", "2.00e+00,"]
    assert lines[4].split() == ["3.00e+00,", "4.00e+00,"]
    # permuted: sizes (1, 1, 2), starts (0, 0, 3)
    assert "c(0,0:0,3:4):" in r.stdout
    assert "1.00e+00 + -1.00e+00i," in r.stdout and "2.00e+00 + -2.00e+00i," in r.stdout


def test_launch_chunks_fit_32bit_grids():
    """Passes with more tiles than one launch can hold run in chunks of
    pfk_max_grid(threads) workgroups: each chunk's work-item count fits in
    32 bits and the chunks cover every tile exactly once."""
    import ctypes
    L = P.lib()
    L.pfk_max_grid.restype = ctypes.c_int64
    L.pfk_max_grid.argtypes = [ctypes.c_int]
    for threads in (64, 128, 256, 320, 512, 768, 1024):
        g = L.pfk_max_grid(threads)
        assert g * threads <= 0xFFFFFFFF < (g + 1) * threads
        for ntiles in (1, g - 1, g, g + 1, 3 * g + 7):
            covered, base = 0, 0
            while base < ntiles:
                nblk = min(ntiles - base, g)
                assert 0 < nblk * threads <= 0xFFFFFFFF
                covered += nblk
                base += g
            assert covered == ntiles
