/*
 * api.c -- exported pfft_ (double) and pfftf_ (float) entry points,
 * memory helpers and the reference's test-data tools.
 */
#include <complex.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "host.h"

static int g_nthreads = 1;

/* pfft_malloc: device memory, or pinned host memory when
 * PFFT_AMD_HOST_ALLOC=1 (for programs that fill arrays from host loops;
 * plans then stage through device buffers). */
static void *pa_alloc(size_t bytes) {
  void *p = NULL;
  const char *e = getenv("PFFT_AMD_HOST_ALLOC");
  if (e && e[0] == '1') {
    if (pfk_malloc_host(&p, bytes)) { pa_error("host allocation failed: %s", pfk_last_error()); return NULL; }
  } else if (pfk_malloc(&p, bytes)) {
    pa_error("device allocation of %zu bytes failed: %s", bytes, pfk_last_error());
    return NULL;
  }
  return p;
}

static void pa_dealloc(void *p) {
  if (!p) return;
  if (pfk_is_device_ptr(p)) pfk_free(p);
  else pfk_free_host(p);
}

static unsigned r2c_flag(unsigned pfft_flags) {
  return (pfft_flags & PFFT_PADDED_R2C) ? (TF_R2C | TF_PADDED) : TF_R2C;
}
static unsigned c2r_flag(unsigned pfft_flags) {
  return (pfft_flags & PFFT_PADDED_C2R) ? (TF_C2R | TF_PADDED) : TF_C2R;
}

/* ---- test-data tools: api/api-basic.c:60-189 ---- */
enum { PA_TOOL_INIT, PA_TOOL_CLEAR, PA_TOOL_CHECK };
enum { AT_REAL = 1, AT_COMPLEX = 2, AT_HERMITIAN = 3 };   /* PFFTI_ARRAYTYPE_* */

/* api-basic.c:254-264: note the index is k = k*(n[t]+1) + kvec[t] */
static INT plain_index(int rnk, const INT *n, const INT *kvec) {
  INT k = 0;
  for (int t = 0; t < rnk; t++) k += k * n[t] + kvec[t];
  return k;
}

/* api-basic.c:60-87, evaluated in the plan's real precision */
static double complex init_scalar(int prec, int rnk, const INT *n, const INT *ind) {
  INT per[PA_MAXRNK];
  for (int t = 0; t < rnk; t++) per[t] = ind[t] % n[t];
  const INT mi = plain_index(rnk, n, per);
  if (prec == PFK_F64) {
    const double m = (double)mi;
    if (m == 0) return 1500.0 + 1250.0 * I;
    return 1000.0 / (2 * m) + 1000.0 / (2 * m + 1) * I;
  } else {
    const float m = (float)mi;
    if (m == 0) return 1500.0 + 1250.0 * I;
    /* float complex result of a double expression with float m */
    const float complex v = (float complex)(1000.0 / (2 * m) + 1000.0 / (2 * m + 1) * I);
    return (double complex)v;
  }
}

static double pa_tool_array(int prec, int tool, int at, int rnk, const INT *n, const INT *ln,
                            const INT *ls, void *data) {
  if (rnk < 1 || rnk > PA_MAXRNK || !data) return 0;
  INT tot = 1;
  for (int t = 0; t < rnk; t++) tot *= ln[t];
  const size_t rsz = (prec == PFK_F64) ? sizeof(double) : sizeof(float);
  const size_t esz = (at == AT_REAL) ? rsz : 2 * rsz;
  const size_t bytes = (size_t)tot * esz;
  const int on_dev = pfk_is_device_ptr(data);
  void *h = data;
  if (on_dev) {
    h = malloc(bytes ? bytes : 1);
    if (!h) return 0;
    if (tool == PA_TOOL_CHECK && pfk_memcpy_d2h(h, data, bytes)) {
      pa_error("check_output: copy failed: %s", pfk_last_error());
      free(h);
      return 0;
    }
  }
  double maxerr = 0;
  INT kl[PA_MAXRNK], kg[PA_MAXRNK], km[PA_MAXRNK];
  for (INT k = 0; k < tot; k++) {
    INT l = k;
    for (int t = rnk - 1; t >= 0; t--) { kl[t] = l % ln[t]; l = (l - kl[t]) / ln[t]; }
    for (int t = 0; t < rnk; t++) { kg[t] = kl[t] + ls[t]; km[t] = n[t] - kg[t]; }
    double complex want = 0;
    int ignore = 0;
    if (tool != PA_TOOL_CLEAR) {
      const double complex d1 = init_scalar(prec, rnk, n, kg);
      if (at == AT_REAL) {
        ignore = !(kg[rnk - 1] < n[rnk - 1]);   /* padding element */
        want = ignore ? 0 : creal(d1);
      } else if (at == AT_COMPLEX) {
        want = d1;
      } else {
        const double complex d2 = init_scalar(prec, rnk, n, km);
        want = 0.5 * (d1 + conj(d2));
        if (prec != PFK_F64) want = (double complex)(float complex)want;
      }
    }
    if (tool == PA_TOOL_CHECK) {
      double err = 0;
      if (!ignore) {
        double complex got;
        if (prec == PFK_F64) got = (at == AT_REAL) ? ((double *)h)[k]
                                                   : ((double *)h)[2 * k] + I * ((double *)h)[2 * k + 1];
        else got = (at == AT_REAL) ? ((float *)h)[k]
                                   : ((float *)h)[2 * k] + I * ((float *)h)[2 * k + 1];
        err = cabs(got - want);
      }
      if (err > maxerr) maxerr = err;
    } else if (prec == PFK_F64) {
      if (at == AT_REAL) ((double *)h)[k] = creal(want);
      else { ((double *)h)[2 * k] = creal(want); ((double *)h)[2 * k + 1] = cimag(want); }
    } else {
      if (at == AT_REAL) ((float *)h)[k] = (float)creal(want);
      else { ((float *)h)[2 * k] = (float)creal(want); ((float *)h)[2 * k + 1] = (float)cimag(want); }
    }
  }
  if (on_dev) {
    if (tool != PA_TOOL_CHECK && pfk_memcpy_h2d(data, h, bytes))
      pa_error("init/clear: copy failed: %s", pfk_last_error());
    free(h);
  }
  return maxerr;
}

/* ---- double precision ---- */
#define PA(name) pfft_##name
#define R double
#define CPLX pfft_complex
#define PA_PREC PFK_F64
#include "api_impl.h"
#undef PA
#undef R
#undef CPLX
#undef PA_PREC

/* ---- single precision ---- */
#define PA(name) pfftf_##name
#define R float
#define CPLX pfftf_complex
#define PA_PREC PFK_F32
#include "api_impl.h"
#undef PA
#undef R
#undef CPLX
#undef PA_PREC

/* ---- extensions ---- */
int pfft_amd_memcpy_to_device(void *dst, const void *src, size_t bytes) {
  return pfk_memcpy_h2d(dst, src, bytes);
}
int pfft_amd_memcpy_to_host(void *dst, const void *src, size_t bytes) {
  return pfk_memcpy_d2h(dst, src, bytes);
}
const char *pfft_amd_version(void) { return "pfft-amd 0.1 (gfx950)"; }
const char *pfft_amd_last_error(void) { return pfk_last_error(); }

int pfft_amd_plan_npasses(const void *plan) { return pa_plan_npasses((const pa_plan *)plan); }
int pfft_amd_time_passes(const void *plan, void *in, void *out, int iters, double *ms,
                         double *bytes) {
  return pa_plan_time_passes((const pa_plan *)plan, in, out, iters, ms, bytes);
}
int pfft_amd_describe_plan(const void *plan, char *buf, size_t len) {
  if (!plan || !buf || !len) return 1;
  return pa_plan_describe((const pa_plan *)plan, buf, len);
}
int pfft_amd_execute_batch_noalloc(const void *const *plans, int count) {
  return pa_plans_execute_batch_noalloc((const pa_plan *const *)plans, count);
}

double pfft_amd_bw_probe(int mode, int width, int nt, long long grid, size_t bytes, int iters) {
  return pfk_bw_probe(mode, width, nt, grid, bytes, iters);
}

int pfft_amd_export_wisdom(const char *filename) { return pa_wisdom_export(filename); }
int pfft_amd_import_wisdom(const char *filename) { return pa_wisdom_import(filename); }
void pfft_amd_forget_wisdom(void) { pa_wisdom_forget(); }
int pfft_amd_wisdom_count(void) { return pa_wisdom_count(); }
int pfft_amd_tune_timings(void) { return pa_tune_timings(); }
