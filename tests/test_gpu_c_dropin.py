"""A C program written the way the reference's own tests are
(tests/simple_check_c2c.c:36-67, tests/simple_check_r2c.c:30-70):
allocate with pfft_alloc_*, plan, fill with pfft_init_input_*, execute,
clear, execute the backward plan, scale with a host loop, check with
pfft_check_output_*.  It is compiled against include/pfft.h and linked
with libpfft.so, and runs on the GPU with PFFT_AMD_HOST_ALLOC=1 (pfft_alloc_*
then return pinned host memory that host loops can touch; plans stage
through device buffers).  The only change from the reference's sources is
that a one-process program needs no MPI_Init / MPI_Barrier /
MPI_Finalize."""
import os
import subprocess

import pytest

import pfft_amd as P

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

SRC = r'''
#include <complex.h>
#include <stdio.h>
#include <pfft.h>

static int c2c(const ptrdiff_t *n, MPI_Comm comm, double *err) {
  ptrdiff_t alloc_local, local_ni[3], local_i_start[3], local_no[3], local_o_start[3];
  alloc_local = pfft_local_size_dft_3d(n, comm, PFFT_TRANSPOSED_NONE,
      local_ni, local_i_start, local_no, local_o_start);
  pfft_complex *in = pfft_alloc_complex(alloc_local);
  pfft_complex *out = pfft_alloc_complex(alloc_local);
  if (!in || !out) return 1;
  pfft_plan plan_forw = pfft_plan_dft_3d(n, in, out, comm, PFFT_FORWARD,
      PFFT_TRANSPOSED_NONE | PFFT_MEASURE | PFFT_DESTROY_INPUT);
  pfft_plan plan_back = pfft_plan_dft_3d(n, out, in, comm, PFFT_BACKWARD,
      PFFT_TRANSPOSED_NONE | PFFT_MEASURE | PFFT_DESTROY_INPUT);
  if (!plan_forw || !plan_back) return 1;
  pfft_init_input_complex_3d(n, local_ni, local_i_start, in);
  pfft_execute(plan_forw);
  pfft_clear_input_complex_3d(n, local_ni, local_i_start, in);
  pfft_execute(plan_back);
  for (ptrdiff_t l = 0; l < local_ni[0] * local_ni[1] * local_ni[2]; l++)
    in[l] /= (n[0] * n[1] * n[2]);
  *err = pfft_check_output_complex_3d(n, local_ni, local_i_start, in, comm);
  pfft_printf(comm, "c2c n=(%td, %td, %td): maxerror = %6.2e;\n", n[0], n[1], n[2], *err);
  pfft_destroy_plan(plan_forw);
  pfft_destroy_plan(plan_back);
  pfft_free(in);
  pfft_free(out);
  return 0;
}

static int r2c(const ptrdiff_t *n, MPI_Comm comm, double *err) {
  ptrdiff_t alloc_local, local_ni[3], local_i_start[3], local_no[3], local_o_start[3];
  alloc_local = pfft_local_size_dft_r2c_3d(n, comm, PFFT_TRANSPOSED_NONE,
      local_ni, local_i_start, local_no, local_o_start);
  double *in = pfft_alloc_real(2 * alloc_local);
  pfft_complex *out = pfft_alloc_complex(alloc_local);
  if (!in || !out) return 1;
  pfft_plan plan_forw = pfft_plan_dft_r2c_3d(n, in, out, comm, PFFT_FORWARD,
      PFFT_TRANSPOSED_NONE | PFFT_MEASURE | PFFT_DESTROY_INPUT);
  pfft_plan plan_back = pfft_plan_dft_c2r_3d(n, out, in, comm, PFFT_BACKWARD,
      PFFT_TRANSPOSED_NONE | PFFT_MEASURE | PFFT_DESTROY_INPUT);
  if (!plan_forw || !plan_back) return 1;
  pfft_init_input_real(3, n, local_ni, local_i_start, in);
  pfft_execute(plan_forw);
  pfft_clear_input_real(3, n, local_ni, local_i_start, in);
  pfft_execute(plan_back);
  for (ptrdiff_t l = 0; l < local_ni[0] * local_ni[1] * local_ni[2]; l++)
    in[l] /= (n[0] * n[1] * n[2]);
  *err = pfft_check_output_real(3, n, local_ni, local_i_start, in, comm);
  pfft_printf(comm, "r2c n=(%td, %td, %td): maxerror = %6.2e;\n", n[0], n[1], n[2], *err);
  pfft_destroy_plan(plan_forw);
  pfft_destroy_plan(plan_back);
  pfft_free(in);
  pfft_free(out);
  return 0;
}

int main(void) {
  ptrdiff_t n[3] = {29, 27, 31};
  MPI_Comm comm_cart_2d;
  double e1 = 1, e2 = 1, e3 = 1;
  pfft_init();
  if (pfft_create_procmesh_2d(MPI_COMM_WORLD, 1, 1, &comm_cart_2d)) return 1;
  if (c2c(n, comm_cart_2d, &e1)) return 3;
  if (r2c(n, comm_cart_2d, &e2)) return 4;
  ptrdiff_t m[3] = {64, 48, 128};
  if (r2c(m, comm_cart_2d, &e3)) return 5;
  pfft_cleanup();
  return (e1 < 1e-12 && e2 < 1e-12 && e3 < 1e-12) ? 0 : 2;
}
'''


def test_reference_style_c_program_runs_on_gpu(tmp_path):
    c = str(tmp_path / "simple_check.c")
    exe = str(tmp_path / "simple_check")
    with open(c, "w") as f:
        f.write(SRC)
    libdir = os.path.dirname(P.LIB_PATH)
    r = subprocess.run(["gcc", "-std=gnu99", "-I" + os.path.join(ROOT, "include"), c, "-o", exe,
                        "-L" + libdir, "-lpfft", "-lm", "-Wl,-rpath," + libdir],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    env = dict(os.environ, PFFT_AMD_HOST_ALLOC="1")
    r = subprocess.run([exe], capture_output=True, text=True, env=env, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, (r.returncode, r.stdout, r.stderr)
    errs = [float(ln.split("maxerror =")[1].strip(" ;")) for ln in r.stdout.splitlines()
            if "maxerror" in ln]
    assert len(errs) == 3 and max(errs) < 1e-12, r.stdout
