/*
 * plan.c -- planning and execution of the transform pipeline on one GPU.
 *
 * What a plan computes is the reference's execute_full
 * (api/api-basic.c:1044-1107) on a one-process mesh:
 *   conjugate (c2r forward) -> twiddle_input (SHIFTED_OUT) -> per step:
 *   embed -> serial FFT -> trunc, with local/global transposes ->
 *   twiddle_output (SHIFTED_IN) -> conjugate (r2c backward).
 * On one process every global transpose is an identity and every local
 * transpose is a permutation, so the whole chain is, per dimension d,
 *   y = M_out,d . Trunc_d . DFT_d . Embed_d . M_in,d  x
 * applied along d, plus a change of memory order between the input and
 * output layouts.  The plan therefore builds one device pass per
 * dimension (pfk.h), folding the +-1 shift modulations, the zero-pad and
 * the truncation into that pass's loads and stores, and the layout change
 * into the pass whose transform axis is contiguous on the changed side.
 * The two conjugations cancel the sign swaps the reference makes for
 * c2r forward / r2c backward (partrafo.c:426-443), so each pass simply
 * uses the user's sign.
 *
 * Pass order: r2c transforms the last dimension first (its input is
 * real), c2r last; a TRANSPOSED_IN plan transforms dimension 0 first and
 * reads the transposed input in that pass; otherwise dimension 0 is done
 * last and, for TRANSPOSED_OUT, writes the transposed output.
 */
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include "host.h"

enum { ROLE_IN = 0, ROLE_OUT = 1, ROLE_S0 = 2, ROLE_S1 = 3 };

typedef struct arrgeo {
  int rnk;
  int order[PA_MAXRNK];   /* memory order of the logical dims, outermost first */
  INT ext[PA_MAXRNK];     /* logical extents (element counts) */
  INT phys_last;          /* physical extent of dim rnk-1 (padding), 0 if = ext */
  INT h;                  /* interleaved tuple (howmany), innermost */
  INT opad;               /* elements added to every stride beyond the innermost dim */
  int opad_outer;         /* 1: opad on the strides beyond the row stride only */
  int is_real;
} arrgeo;

struct pa_plan {
  uint64_t id;              /* unique per created plan (batch cache key) */
  int prec;
  int rnk_n;
  INT n[PA_MAXRNK], ni[PA_MAXRNK], no[PA_MAXRNK];
  INT howmany;
  unsigned trafo_flag, pfft_flags, transp;
  int sign;
  int skip[PA_MAXRNK];
  INT local_ni[PA_MAXRNK], local_i_start[PA_MAXRNK];
  INT local_no[PA_MAXRNK], local_o_start[PA_MAXRNK];
  void *in, *out;
  arrgeo gin, gout;
  int64_t in_elems, out_elems;   /* in scalar elements of the array's type */
  size_t in_esz, out_esz;        /* bytes per element */
  int npass;
  pfk_pass pass[PA_MAXRNK];
  int src_role[PA_MAXRNK], dst_role[PA_MAXRNK];
  void *scratch[2];
  size_t scratch_bytes[2];
  const void *tw_adj[PA_MAXRNK], *rtw_adj[PA_MAXRNK];   /* opposite-sign tables */
  int is_r2r;
  void *tw_bufs[4 * PA_MAXRNK];
  int ntw;
  void *work[PA_MAXRNK];         /* long-line passes: lines * nf complex elements */
  void *work2[PA_MAXRNK];        /* long real-line passes: the same again */
  void *r2rw[2];                 /* r2r passes on the inner real FFT: its input / output */
  /* execution timer (kernel/timer.c fields): pass k adds its device time
   * to trafo[tslot[k]]; events tev[0..npass] bracket the passes */
  pa_timer *timer;
  int tslot[PA_MAXRNK];
  void *tev[PA_MAXRNK + 1];
};

void pa_error(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  fprintf(stderr, "PFFT ERROR: ");
  vfprintf(stderr, fmt, ap);
  fprintf(stderr, "\n");
  va_end(ap);
}

static int64_t geo_strides(const arrgeo *a, INT *stride) {
  INT run = a->h;
  for (int k = a->rnk - 1; k >= 0; --k) {
    const int d = a->order[k];
    stride[d] = run;
    run *= (d == a->rnk - 1 && a->phys_last) ? a->phys_last : a->ext[d];
    if (k > 0 && !(a->opad_outer && k == a->rnk - 1)) run += a->opad;
  }
  return (int64_t)run;
}

static void order_identity(int rnk, int *order) {
  for (int t = 0; t < rnk; t++) order[t] = t;
}

/* transposed decomposition n1 x .. x n_r x n0 x n_{r+1} x .. (doc/tutorial.tex:319-321) */
static void order_transposed(int rnk, int r, int *order) {
  int k = 0;
  for (int t = 1; t <= r; t++) order[k++] = t;
  order[k++] = 0;
  for (int t = r + 1; t < rnk; t++) order[k++] = t;
}

static int same_geo(const arrgeo *a, const arrgeo *b) {
  if (a->rnk != b->rnk || a->h != b->h || a->is_real != b->is_real) return 0;
  INT sa[PA_MAXRNK], sb[PA_MAXRNK];
  geo_strides(a, sa);
  geo_strides(b, sb);
  for (int t = 0; t < a->rnk; t++)
    if (a->ext[t] != b->ext[t] || sa[t] != sb[t]) return 0;
  return 1;
}


static const void *make_twiddles(pa_plan *p, INT N, INT count, int sign) {
  double *h = (double *)malloc(sizeof(double) * 2 * (size_t)(count > 0 ? count : 1));
  if (!h) return NULL;
  pa_twiddles((long)N, (long)count, sign, h);
  void *d = NULL;
  size_t bytes;
  if (p->prec == PFK_F64) {
    bytes = sizeof(double) * 2 * (size_t)count;
    if (pfk_malloc(&d, bytes) || pfk_memcpy_h2d(d, h, bytes)) { free(h); return NULL; }
  } else {
    float *f = (float *)malloc(sizeof(float) * 2 * (size_t)count);
    for (INT i = 0; i < 2 * count; i++) f[i] = (float)h[i];
    bytes = sizeof(float) * 2 * (size_t)count;
    if (pfk_malloc(&d, bytes) || pfk_memcpy_h2d(d, f, bytes)) { free(f); free(h); return NULL; }
    free(f);
  }
  free(h);
  p->tw_bufs[p->ntw++] = d;
  return d;
}

/* Table of a long line nf = n1 * n2 (pfk.h): W_n1^m (m < n1), W_n2^m
 * (m < n2), W_nf^l (l < 4096), W_nf^(4096 h) (h <= nf / 4096). */
static const void *make_long_twiddles(pa_plan *p, INT nf, INT n1, int sign) {
  const INT n2 = nf / n1, nh = nf / 4096 + 1;
  const INT count = n1 + n2 + 4096 + nh;
  double *h = (double *)malloc(sizeof(double) * 2 * (size_t)count);
  if (!h) return NULL;
  pa_twiddles((long)n1, (long)n1, sign, h);
  pa_twiddles((long)n2, (long)n2, sign, h + 2 * n1);
  pa_twiddles((long)nf, 4096, sign, h + 2 * (n1 + n2));
  pa_twiddles_step((long)nf, (long)nh, 4096, sign, h + 2 * (n1 + n2 + 4096));
  void *d = NULL;
  int err;
  if (p->prec == PFK_F64) {
    const size_t bytes = sizeof(double) * 2 * (size_t)count;
    err = pfk_malloc(&d, bytes) || pfk_memcpy_h2d(d, h, bytes);
  } else {
    float *f = (float *)malloc(sizeof(float) * 2 * (size_t)count);
    err = !f;
    if (f) {
      for (INT i = 0; i < 2 * count; i++) f[i] = (float)h[i];
      const size_t bytes = sizeof(float) * 2 * (size_t)count;
      err = pfk_malloc(&d, bytes) || pfk_memcpy_h2d(d, f, bytes);
      free(f);
    }
  }
  free(h);
  if (err) { pfk_free(d); return NULL; }
  p->tw_bufs[p->ntw++] = d;
  return d;
}

typedef struct bdim { INT ext, s_src, s_dst; } bdim;

static int cmp_bdim(const void *a, const void *b) {
  const bdim *x = (const bdim *)a, *y = (const bdim *)b;
  if (x->s_src != y->s_src) return (x->s_src > y->s_src) ? -1 : 1;
  return (x->s_dst > y->s_dst) ? -1 : (x->s_dst < y->s_dst);
}

/* Fill the geometry part of pass `ps` transforming dim d from S to D. */
static int fill_geometry(pfk_pass *ps, const arrgeo *S, const arrgeo *D, int d) {
  INT ss[PA_MAXRNK], sd[PA_MAXRNK];
  ps->src_elems = geo_strides(S, ss);
  ps->dst_elems = geo_strides(D, sd);
  ps->sin = ss[d];
  ps->sout = sd[d];
  bdim b[PA_MAXRNK + 1];
  int nb = 0;
  for (int e = 0; e < S->rnk; e++) {
    if (e == d || S->ext[e] == 1) continue;
    b[nb].ext = S->ext[e]; b[nb].s_src = ss[e]; b[nb].s_dst = sd[e]; nb++;
  }
  if (S->h > 1) { b[nb].ext = S->h; b[nb].s_src = 1; b[nb].s_dst = 1; nb++; }
  qsort(b, (size_t)nb, sizeof(bdim), cmp_bdim);
  /* merge dims that are contiguous with their inner neighbour on both sides */
  int m = 0;
  for (int i = 0; i < nb; i++) {
    if (m > 0) {
      bdim *o = &b[m - 1];
      if (o->s_src == b[i].ext * b[i].s_src && o->s_dst == b[i].ext * b[i].s_dst) {
        o->ext *= b[i].ext; o->s_src = b[i].s_src; o->s_dst = b[i].s_dst;
        continue;
      }
    }
    b[m++] = b[i];
  }
  nb = m;
  int tile = -1;
  if (ps->sin != 1)
    for (int i = 0; i < nb; i++) if (b[i].s_src == 1) { tile = i; break; }
  if (tile < 0 && ps->sout != 1)
    for (int i = 0; i < nb; i++) if (b[i].s_dst == 1) { tile = i; break; }
  if (tile < 0 && nb > 0) tile = nb - 1;   /* innermost remaining dim */
  if (tile >= 0) {
    ps->text = b[tile].ext; ps->tsin = b[tile].s_src; ps->tsout = b[tile].s_dst;
    for (int i = tile; i < nb - 1; i++) b[i] = b[i + 1];
    nb--;
  } else {
    ps->text = 1; ps->tsin = 0; ps->tsout = 0;
  }
  if (nb > PFK_MAXB) { pa_error("too many batch dimensions (%d)", nb); return 1; }
  ps->nb = nb;
  /* innermost batch dim first, for the kernel's div/mod decode */
  for (int i = 0; i < nb; i++) {
    ps->bext[i] = b[nb - 1 - i].ext;
    ps->bsin[i] = b[nb - 1 - i].s_src;
    ps->bsout[i] = b[nb - 1 - i].s_dst;
  }
  return 0;
}

/* highest element offset a pass touches on each side must fit the buffer */
static int check_bounds(const pfk_pass *ps) {
  int64_t ms = (ps->nin - 1) * ps->sin + (ps->text - 1) * ps->tsin;
  int64_t md = (ps->nout - 1) * ps->sout + (ps->text - 1) * ps->tsout;
  for (int i = 0; i < ps->nb; i++) {
    ms += (ps->bext[i] - 1) * ps->bsin[i];
    md += (ps->bext[i] - 1) * ps->bsout[i];
  }
  if (ps->sin < 0 || ps->sout < 0 || ms >= ps->src_elems || md >= ps->dst_elems) {
    pa_error("internal: pass geometry exceeds its buffers (%lld/%lld, %lld/%lld)",
             (long long)ms, (long long)ps->src_elems, (long long)md, (long long)ps->dst_elems);
    return 1;
  }
  return 0;
}

static int64_t geo_elems(const arrgeo *g) {
  INT s[PA_MAXRNK];
  return geo_strides(g, s);
}

static double wall_seconds(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

static void free_plan_memory(pa_plan *p) {
  for (int k = 0; k <= PA_MAXRNK; k++)
    if (p->tev[k]) { pfk_event_destroy(p->tev[k]); p->tev[k] = NULL; }
  free(p->timer);
  p->timer = NULL;
  for (int i = 0; i < p->ntw; i++) pfk_free(p->tw_bufs[i]);
  p->ntw = 0;
  for (int i = 0; i < 2; i++) { pfk_free(p->scratch[i]); p->scratch[i] = NULL; }
  for (int k = 0; k < PA_MAXRNK; k++) {
    pfk_free(p->work[k]);
    pfk_free(p->work2[k]);
    p->work[k] = p->work2[k] = NULL;
  }
  for (int i = 0; i < 2; i++) { pfk_free(p->r2rw[i]); p->r2rw[i] = NULL; }
}

/* Chooses every pass's launch variant (tune.c) on temporary device
 * buffers shaped like the plan's arrays, so planning leaves the user's
 * arrays untouched.  Without memory for the buffers the defaults stay. */
static void tune_plan(pa_plan *p) {
  const unsigned effort = p->pfft_flags & (PFFT_ESTIMATE | PFFT_PATIENT | PFFT_EXHAUSTIVE);
  int open_[PA_MAXRNK] = {0}, any = 0;
  for (int k = 0; k < p->npass; k++) any |= open_[k] = pa_tune_lookup(&p->pass[k], effort);
  if (!any) return;
  void *tin = NULL, *tout = NULL;
  const size_t ib = (size_t)(p->in_elems * (int64_t)p->in_esz);
  const size_t ob = (size_t)(p->out_elems * (int64_t)p->out_esz);
  if (pfk_malloc(&tin, ib) || pfk_malloc(&tout, ob) || pfk_memset(tin, 0, ib) ||
      pfk_memset(tout, 0, ob)) {
    pfk_free(tin);
    pfk_free(tout);
    return;
  }
  void *bufs[4] = {tin, tout, p->scratch[0], p->scratch[1]};
  for (int k = 0; k < p->npass; k++)
    if (open_[k] &&
        pa_tune_measure(&p->pass[k], bufs[p->src_role[k]], bufs[p->dst_role[k]], effort))
      break;
  (void)pfk_sync();
  pfk_free(tin);
  pfk_free(tout);
}

/* Extended complex FFT length of an r2r line of length n (pfk.h), 0 for
 * an unsupported kind. */
static INT r2r_length(int kind, INT n) {
  switch (kind) {
    case PFK_DHT: return n;
    case PFK_REDFT00: return 2 * (n - 1);
    case PFK_RODFT00: return 2 * (n + 1);
    case PFK_REDFT10: case PFK_REDFT01: case PFK_RODFT10: case PFK_RODFT01: return 4 * n;
    case PFK_REDFT11: case PFK_RODFT11: return 8 * n;
    default: return 0;
  }
}

pa_plan *pa_plan_create(int prec, int rnk_n, const INT *n, const INT *ni, const INT *no,
                        INT howmany, const INT *iblock, const INT *oblock, void *in, void *out,
                        MPI_Comm comm, int sign, const int *skip_trafos, unsigned tf,
                        unsigned pfft_flags, const int *kinds) {
  const unsigned transp = pfft_flags & (PFFT_TRANSPOSED_IN | PFFT_TRANSPOSED_OUT);
  pa_mesh mesh;

  /* validation: kernel/partrafo.c:337-373 */
  if (pfft_flags & (PFFT_SHIFTED_IN | PFFT_SHIFTED_OUT))
    for (int t = 0; t < rnk_n; t++)
      if (n[t] % 2) return NULL;
  if ((transp & PFFT_TRANSPOSED_OUT) && (transp & PFFT_TRANSPOSED_IN)) return NULL;
  if ((tf & TF_R2C) && (transp & PFFT_TRANSPOSED_IN)) return NULL;
  if ((tf & TF_C2R) && (transp & PFFT_TRANSPOSED_OUT)) return NULL;
  if (pa_mesh_of(comm, &mesh)) return NULL;
  const int r = mesh.rnk_pm;
  if (rnk_n < r) return NULL;
  if (rnk_n == r && rnk_n != 3) return NULL;

  /* limits of this implementation */
  if (rnk_n > PA_MAXRNK || rnk_n < 1) { pa_error("rank %d not supported", rnk_n); return NULL; }
  if (rnk_n == 3 && r == 3) {
    pa_error("3-d process meshes (3d-to-2d remap) are not supported on one GPU");
    return NULL;
  }
  if (tf & TF_R2R) {
    if (!kinds) { pa_error("r2r plans need one kind per dimension"); return NULL; }
    for (int t = 0; t < rnk_n; t++) {
      const INT m = r2r_length(kinds[t], n[t]);
      if (m <= 0) {
        pa_error("r2r kind %d of dimension %d is not supported (DHT, REDFT, RODFT)", kinds[t], t);
        return NULL;
      }
    }
  }
  if ((pfft_flags & PFFT_AMD_MFMA_F32) && (pfft_flags & PFFT_AMD_MFMA_BF16X2)) {
    pa_error("choose one of PFFT_AMD_MFMA_F32 and PFFT_AMD_MFMA_BF16X2");
    return NULL;
  }
  if ((pfft_flags & (PFFT_AMD_MFMA_F32 | PFFT_AMD_MFMA_BF16X2)) && prec != PFK_F32) {
    pa_error("the MFMA arithmetic forms are for complex64 (pfftf_) plans");
    return NULL;
  }
  if (mesh.size != 1) { pa_error("process mesh must have one process"); return NULL; }
  if (howmany < 1) return NULL;
  for (int t = 0; t < rnk_n; t++) {
    if (n[t] < 1 || ni[t] < 1 || no[t] < 1) { pa_error("sizes must be positive"); return NULL; }
    if (ni[t] > n[t] || no[t] > n[t]) {
      pa_error("ni and no must not exceed n (dim %d)", t);
      return NULL;
    }
  }
  if (skip_trafos && (tf & TF_RDFT) && skip_trafos[r]) {
    pa_error("skipping the r2c/c2r dimension is not supported");
    return NULL;
  }

  pa_plan *p = (pa_plan *)calloc(1, sizeof(pa_plan));
  if (!p) return NULL;
  static uint64_t next_id = 1;
  p->id = next_id++;
  p->prec = prec;
  p->rnk_n = rnk_n;
  p->howmany = howmany;
  p->trafo_flag = tf;
  p->pfft_flags = pfft_flags;
  p->transp = transp;
  p->sign = sign;
  p->in = in;
  p->out = out;
  for (int t = 0; t < rnk_n; t++) {
    p->n[t] = n[t]; p->ni[t] = ni[t]; p->no[t] = no[t];
    p->skip[t] = skip_trafos ? skip_trafos[t] : 0;
  }
  pa_local_size(rnk_n, n, ni, no, howmany, iblock, oblock, &mesh, tf, pfft_flags, p->local_ni,
                p->local_i_start, p->local_no, p->local_o_start);
  {
    /* a user block smaller than the data would leave part of the array on
     * processes that do not exist */
    INT lni[PA_MAXRNK], lno[PA_MAXRNK];
    for (int t = 0; t < rnk_n; t++) { lni[t] = p->local_ni[t]; lno[t] = p->local_no[t]; }
    const int last = rnk_n - 1;
    for (int t = 0; t < rnk_n; t++) {
      INT full_i = (t == last && (tf & TF_C2R)) ? ni[t] / 2 + 1 : ni[t];
      INT full_o = (t == last && (tf & TF_R2C)) ? no[t] / 2 + 1 : no[t];
      if (t == last && (tf & TF_R2C)) full_i = (tf & TF_PADDED) ? 2 * (ni[t] / 2 + 1) : ni[t];
      if (t == last && (tf & TF_C2R)) full_o = (tf & TF_PADDED) ? 2 * (no[t] / 2 + 1) : no[t];
      if (lni[t] != full_i || lno[t] != full_o) {
        pa_error("user block sizes must cover the whole array on a one-process mesh");
        free(p);
        return NULL;
      }
    }
  }

  const int R = rnk_n, last = R - 1;
  const int is_r2c = (tf & TF_R2C) != 0, is_c2r = (tf & TF_C2R) != 0;
  const int is_r2r = (tf & TF_R2R) != 0;
  const int padded = (tf & TF_PADDED) != 0;
  p->is_r2r = is_r2r;
  p->in_esz = (prec == PFK_F64 ? 8 : 4) * ((is_r2c || is_r2r) ? 1 : 2);
  p->out_esz = (prec == PFK_F64 ? 8 : 4) * ((is_c2r || is_r2r) ? 1 : 2);

  /* user arrays */
  p->gin.rnk = p->gout.rnk = R;
  p->gin.h = p->gout.h = howmany;
  if (transp & PFFT_TRANSPOSED_IN) order_transposed(R, r, p->gin.order);
  else order_identity(R, p->gin.order);
  if (transp & PFFT_TRANSPOSED_OUT) order_transposed(R, r, p->gout.order);
  else order_identity(R, p->gout.order);
  INT ext_in_c[PA_MAXRNK], ext_out_c[PA_MAXRNK];   /* extents of complex intermediates */
  for (int t = 0; t < R; t++) {
    p->gin.ext[t] = ni[t];
    p->gout.ext[t] = no[t];
    ext_in_c[t] = ni[t];
    ext_out_c[t] = no[t];
  }
  if (is_r2c) {
    p->gin.is_real = 1;
    if (padded) p->gin.phys_last = 2 * (ni[last] / 2 + 1);
    p->gout.ext[last] = no[last] / 2 + 1;
    ext_out_c[last] = no[last] / 2 + 1;
  }
  if (is_c2r) {
    p->gout.is_real = 1;
    if (padded) p->gout.phys_last = 2 * (no[last] / 2 + 1);
    p->gin.ext[last] = ni[last] / 2 + 1;
    ext_in_c[last] = ni[last] / 2 + 1;
  }
  if (is_r2r) p->gin.is_real = p->gout.is_real = 1;
  p->in_elems = geo_elems(&p->gin);
  p->out_elems = geo_elems(&p->gout);

  /* pass order */
  int ord[PA_MAXRNK];
  const int first0 = is_c2r || (!is_r2c && (transp & PFFT_TRANSPOSED_IN));
  for (int k = 0; k < R; k++) ord[k] = first0 ? k : R - 1 - k;
  const char *ev = getenv("PFFT_AMD_C2C_ORDER");
  const int c2c_order = ev ? (ev[0] == '1') : 1;
  if (c2c_order && !is_r2c && !is_c2r && !transp && R >= 3) {
    /* Untransposed complex (and r2r) plans: dimension R-2 first, then
     * 0 .. R-3, and the contiguous dimension R-1 last.  The first pass
     * then reads the user input with the short stride of one row, the
     * outer dimensions are transformed on the padded intermediates, and
     * the pass that writes the user output writes whole contiguous lines.
     * (Dimension 0 last would write the output with a stride of
     * n1*...*n_{R-1} elements, a large power of two for power-of-two
     * sizes, which the memory system serves ~20% slower.) */
    ord[0] = R - 2;
    for (int k = 1; k < R - 1; k++) ord[k] = k - 1;
    ord[R - 1] = R - 1;
  }

  ev = getenv("PFFT_AMD_PAD_ROWS");
  const int pad_rows = ev ? (ev[0] == '1') : 1;
  ev = getenv("PFFT_AMD_PAD_PLANES");
  const int pad_planes = ev ? (ev[0] == '1') : 1;
  ev = getenv("PFFT_AMD_ROW_OPAD");
  const int row_opad = ev ? (ev[0] == '1') : -1;   /* -1: by the row's segment count */
  ev = getenv("PFFT_AMD_XCD_REMAP");
  const int xcd_remap = ev ? (ev[0] == '1') : 0;
  ev = getenv("PFFT_AMD_INTER_ORDER");
  const int inter_order = ev ? (ev[0] == '1') : 0;
  int done[PA_MAXRNK] = {0};
  arrgeo cur = p->gin, nxt;
  int64_t scratch_elems[2] = {0, 0};
  int64_t r2r_elems[2] = {0, 0};   /* reals of the inner-FFT work arrays */
  p->npass = R;
  int src_role = ROLE_IN;
  arrgeo prev_inter;
  int have_prev = 0;
  for (int k = 0; k < R; k++) {
    const int d = ord[k];
    pfk_pass *ps = &p->pass[k];
    memset(ps, 0, sizeof(*ps));
    done[d] = 1;
    if (k == R - 1) {
      nxt = p->gout;
    } else {
      memset(&nxt, 0, sizeof(nxt));
      nxt.rnk = R;
      nxt.h = howmany;
      order_identity(R, nxt.order);
      /* Optional (PFFT_AMD_INTER_ORDER=1): place the next pass's line
       * dimension ord[k+1] just outside the innermost one, so that pass
       * reads its lines with a short stride (one padded row) and the long
       * stride moves to this pass's stores.  Measured on MI355X: 5% faster
       * for 1024x1024x512 complex128, slower for 512^3 r2c/c2r and for
       * complex64, so it is off by default. */
      const int dn = ord[k + 1];
      if (inter_order && dn != last) {
        int o = 0;
        for (int t = 0; t < R; t++)
          if (t != dn && t != last) nxt.order[o++] = t;
        nxt.order[o++] = dn;
        nxt.order[o++] = last;
      }
      for (int t = 0; t < R; t++) nxt.ext[t] = done[t] ? ext_out_c[t] : ext_in_c[t];
      nxt.is_real = is_r2r;   /* r2r intermediates are real */
      /* rows of the intermediate arrays start on 128-byte boundaries so the
       * 128-byte segments the pass kernels move are whole cache lines */
      const INT align = (prec == PFK_F64 ? 8 : 16) * (is_r2r ? 2 : 1);
      /* rows shorter than 16 segments (2 KiB) stay unpadded (this and the
       * pitches below): rounding a short row adds a partly filled segment
       * to a large share of its bytes (384 uneven complex64 c2c transforms
       * of sides 24..128: 1.45 ms per batch unpadded, 1.49 padded) */
      const INT pad_min = 16 * align;
      if (pad_rows && howmany == 1 && nxt.ext[last] >= pad_min && nxt.ext[last] % align)
        nxt.phys_last = (nxt.ext[last] + align - 1) / align * align;
      /* every plane pitch one 128-byte segment longer than its contents,
       * and every row pitch an odd number of segments (one more where the
       * rounded row has an even count), so that no stride of a strided
       * pass is a multiple of a large power of two: lines of a tile then
       * spread over the memory channels (a 4 MiB plane stride moved ~20%
       * slower than a padded one).  A row already odd keeps its length:
       * 512^3 r2c rows of 264 complex128 (33 segments) padded to 272 made
       * the column passes ~4% slower than at 264. */
      if (pad_rows && pad_planes && howmany == 1 && nxt.ext[last] >= pad_min) {
        const INT row = nxt.phys_last ? nxt.phys_last : nxt.ext[last];
        nxt.opad = align;
        nxt.opad_outer = row_opad >= 0 ? !row_opad : (row % align == 0 && (row / align) % 2 == 1);
      }
    }
    const int skipped = p->skip[d < r ? d : r] != 0;
    int kind = PFK_C2C;
    if (d == last && is_r2c) kind = PFK_R2C;
    else if (d == last && is_c2r) kind = PFK_C2R;
    else if (skipped) kind = PFK_COPY;
    else if (is_r2r) kind = PFK_R2R;
    ps->kind = kind;
    ps->xcd_remap = xcd_remap;
    ps->mfma = (pfft_flags & PFFT_AMD_MFMA_F32) ? 1 : (pfft_flags & PFFT_AMD_MFMA_BF16X2) ? 2 : 0;
    ps->prec = prec;
    ps->sign = (kind == PFK_R2R) ? PFFT_FORWARD : sign;
    ps->n = n[d];
    ps->nf = ((kind == PFK_R2C || kind == PFK_C2R) && n[d] % 2 == 0) ? n[d] / 2 : n[d];
    if (kind == PFK_R2R) {
      ps->r2r = kinds[d];
      ps->nf = r2r_length(kinds[d], n[d]);
    }
    /* embed / trunc offsets: kernel/ousample.c:241-322 */
    const int sh_in = (pfft_flags & PFFT_SHIFTED_IN) != 0;
    const int sh_out = (pfft_flags & PFFT_SHIFTED_OUT) != 0;
    ps->nin = ni[d];
    ps->zin = sh_in ? (n[d] - ni[d]) / 2 : 0;
    ps->nout = no[d];
    ps->zout = sh_out ? (n[d] - no[d]) / 2 : 0;
    if (kind == PFK_R2C) {
      ps->nout = no[d] / 2 + 1;
      ps->zout = (n[d] / 2 + 1) - (no[d] / 2 + 1);   /* keeps the upper end */
    }
    if (kind == PFK_C2R) {
      ps->nin = ni[d] / 2 + 1;
      ps->zin = (n[d] / 2 + 1) - (ni[d] / 2 + 1);    /* data at the upper end */
    }
    /* +-1 modulations: api/api-basic.c:1186-1285 */
    ps->imod.on = (pfft_flags & PFFT_SHIFTED_OUT) && !p->skip[d];
    ps->imod.start = p->local_i_start[d];
    ps->imod.half = ni[d] / 2;
    ps->imod.extra = (sh_in && ((n[d] / 2) % 2)) ? -1 : 1;
    ps->omod.on = (pfft_flags & PFFT_SHIFTED_IN) && !p->skip[d];
    ps->omod.start = p->local_o_start[d];
    ps->omod.half = no[d] / 2;
    ps->omod.extra = 1;
    if (kind == PFK_R2R && n[d] >= 8) {
      /* the inner real FFT of about the line's own length, where one
       * workgroup transforms it (else the extension below) */
      pfk_pass t = *ps, in;
      int64_t l1, l2;
      int path;
      t.n = n[d];
      t.text = 1;
      t.nb = 0;
      t.sin = t.sout = 1;
      if (!pfk_r2r_inner(&t, &in, &l1, &l2, &path) && in.nf <= PFK_MAXLINE &&
          !pfk_pass_check(&in))
        ps->r2r_path = path;
    }
    if (kind != PFK_COPY && !ps->r2r_path && ps->nf > PFK_MAXLINE) {
      /* long lines: two steps through a work buffer (c2c; even-length
       * r2c / c2r as their half-length complex line) */
      if ((kind == PFK_R2C || kind == PFK_C2R) && n[d] % 2) {
        pa_error("dimension %d: odd real lines of computed length %lld exceed %d", d,
                 (long long)ps->nf, PFK_MAXLINE);
        goto fail;
      }
      ps->n1 = pfk_long_split(ps->nf);
      if (!ps->n1) {
        pa_error("dimension %d: length %lld has no split into two factors <= %d", d,
                 (long long)ps->nf, PFK_MAXLINE);
        goto fail;
      }
    }
    if (kind != PFK_COPY) {
      ps->nstages = pfk_factorize(ps->nf, ps->radix);
      if (ps->nstages < 0) { pa_error("cannot factor length %lld", (long long)ps->nf); goto fail; }
      if (ps->r2r_path) {
        /* tables of the inner transform and the pre/post twiddles (pfk.h) */
        pfk_pass in;
        int64_t l1, l2;
        int path;
        (void)pfk_r2r_inner(ps, &in, &l1, &l2, &path);
        ps->tw = make_twiddles(p, in.nf, in.nf, in.sign);
        p->tw_adj[k] = ps->tw;   /* r2r plans have no adjoint */
        const int mode = pfk_r2r_mode(kinds[d]);
        if (path == 1 && mode != 2) {
          /* the half-spectrum combine: W_L^m (mode 0) or W_n^m (mode 1), m <= M */
          const INT L = mode == 0 ? pfk_r2r_inner_len(kinds[d], n[d]) : n[d];
          ps->rtw = make_twiddles(p, L, L / 2 + 1, in.sign);
          if (!ps->rtw) goto fail;
        }
        const int k10 = kinds[d] == PFK_REDFT10 || kinds[d] == PFK_RODFT10 ||
                        kinds[d] == PFK_REDFT01 || kinds[d] == PFK_RODFT01;
        const int k11 = kinds[d] == PFK_REDFT11 || kinds[d] == PFK_RODFT11;
        if (k10) ps->tw2 = make_twiddles(p, 4 * n[d], n[d] + 1, -1);
        if (k11) ps->tw2 = make_twiddles(p, 8 * n[d], 2 * n[d], -1);
        if ((k10 || k11) && !ps->tw2) goto fail;
      } else if (ps->n1) {
        ps->tw = make_long_twiddles(p, ps->nf, ps->n1, ps->sign);
        p->tw_adj[k] = make_long_twiddles(p, ps->nf, ps->n1, -ps->sign);
      } else {
        ps->tw = make_twiddles(p, ps->nf, ps->nf, ps->sign);
        p->tw_adj[k] = make_twiddles(p, ps->nf, ps->nf, -ps->sign);
      }
      if (!ps->tw || !p->tw_adj[k]) goto fail;
      if ((kind == PFK_R2C || kind == PFK_C2R) && n[d] % 2 == 0) {
        ps->rtw = make_twiddles(p, n[d], n[d] / 2 + 1, sign);   /* never r2r */
        p->rtw_adj[k] = make_twiddles(p, n[d], n[d] / 2 + 1, -sign);
        if (!ps->rtw || !p->rtw_adj[k]) goto fail;
      }
    }
    if (fill_geometry(ps, &cur, &nxt, d)) goto fail;
    if (check_bounds(ps)) goto fail;
    if (ps->n1) {
      int64_t lines = ps->text;
      for (int i = 0; i < ps->nb; i++) lines *= ps->bext[i];
      const size_t wb = (size_t)(lines * ps->nf) * (prec == PFK_F64 ? 16 : 8);
      if (kind == PFK_R2R) {
        /* the extended line lives in work2 */
        if (pfk_malloc(&p->work2[k], wb)) {
          pa_error("device allocation of %zu bytes failed: %s", wb, pfk_last_error());
          goto fail;
        }
        ps->work2 = p->work2[k];
      } else if (kind == PFK_R2C || kind == PFK_C2R) {
        /* the real side is read / written as complex pairs */
        const int r2c = kind == PFK_R2C;
        int even = (r2c ? ps->tsin : ps->tsout) % 2 == 0;
        for (int i = 0; i < ps->nb; i++) even &= (r2c ? ps->bsin[i] : ps->bsout[i]) % 2 == 0;
        const int plain = r2c ? (ps->sin == 1 && ps->nin == n[d] && ps->zin == 0 && !ps->imod.on)
                              : (ps->sout == 1 && ps->nout == n[d] && ps->zout == 0 &&
                                 !ps->omod.on);
        if (!even || !plain) {
          pa_error("dimension %d: real lines longer than %d need contiguous, unembedded, "
                   "unshifted real data (howmany 1)", d, 2 * PFK_MAXLINE);
          goto fail;
        }
        if (pfk_malloc(&p->work2[k], wb)) {
          pa_error("device allocation of %zu bytes failed: %s", wb, pfk_last_error());
          goto fail;
        }
        ps->work2 = p->work2[k];
      }
      if (pfk_malloc(&p->work[k], wb)) {
        pa_error("device allocation of %zu bytes failed: %s", wb, pfk_last_error());
        goto fail;
      }
      ps->work = p->work[k];
    } else if (ps->r2r_path) {
      pfk_pass in;
      int64_t l1, l2;
      int path;
      (void)pfk_r2r_inner(ps, &in, &l1, &l2, &path);
      int64_t lines = ps->text;
      for (int i = 0; i < ps->nb; i++) lines *= ps->bext[i];
      if (lines * l1 > r2r_elems[0]) r2r_elems[0] = lines * l1;
      if (lines * l2 > r2r_elems[1]) r2r_elems[1] = lines * l2;
    } else if (pfk_pass_check(ps)) {
      pa_error("dimension %d: %s", d, pfk_last_error());
      goto fail;
    }
    /* buffers */
    p->src_role[k] = src_role;
    int dst_role;
    if (k == R - 1) {
      dst_role = ROLE_OUT;
    } else {
      if (have_prev && (src_role == ROLE_S0 || src_role == ROLE_S1) && same_geo(&prev_inter, &nxt))
        dst_role = src_role;
      else
        dst_role = (src_role == ROLE_S0) ? ROLE_S1 : ROLE_S0;
      const int64_t e = geo_elems(&nxt);
      if (e > scratch_elems[dst_role - ROLE_S0]) scratch_elems[dst_role - ROLE_S0] = e;
      prev_inter = nxt;
      have_prev = 1;
    }
    p->dst_role[k] = dst_role;
    src_role = dst_role;
    cur = nxt;
  }
  for (int i = 0; i < 2; i++) {
    const size_t b = (size_t)r2r_elems[i] * (prec == PFK_F64 ? 8 : 4);
    if (b && pfk_malloc(&p->r2rw[i], b)) {
      pa_error("device allocation of %zu bytes failed: %s", b, pfk_last_error());
      goto fail;
    }
  }
  for (int k = 0; k < R; k++)
    if (p->pass[k].r2r_path) { p->pass[k].work = p->r2rw[0]; p->pass[k].work2 = p->r2rw[1]; }
  {
    const size_t cesz = (prec == PFK_F64 ? 16 : 8) / (is_r2r ? 2 : 1);
    for (int i = 0; i < 2; i++) {
      p->scratch_bytes[i] = (size_t)scratch_elems[i] * cesz;
      if (p->scratch_bytes[i] && pfk_malloc(&p->scratch[i], p->scratch_bytes[i])) {
        pa_error("device allocation of %zu bytes failed: %s", p->scratch_bytes[i],
                 pfk_last_error());
        goto fail;
      }
    }
  }
  /* timer (kernel/partrafo.c:895, kernel/timer.c:358-373).  Each pass does
   * one dimension's FFT with the remaps fused into its strides; its time
   * goes to the trafo slot of the reference step that transforms that
   * dimension: the T_OUT pipeline does dims r..R-1 in trafo[0] and dim d < r
   * in trafo[r-d]; the T_IN pipeline (TRANSPOSED_IN plans and c2r) dim
   * d < r in trafo[r+1+d] and dims r..R-1 in trafo[2r+1]
   * (api/api-basic.c:1071-1106, 1288-1311). */
  p->timer = pa_timer_new(r);
  if (!p->timer) goto fail;
  for (int k = 0; k < R; k++) {
    const int d = ord[k];
    p->tslot[k] = first0 ? (d >= r ? 2 * r + 1 : r + 1 + d) : (d >= r ? 0 : r - d);
  }
  for (int k = 0; k <= R; k++)
    if (pfk_event_create(&p->tev[k])) {
      pa_error("event creation failed: %s", pfk_last_error());
      goto fail;
    }
  tune_plan(p);
  return p;
fail:
  free_plan_memory(p);
  free(p);
  return NULL;
}

static void batch_cache_forget(uint64_t id);

void pa_plan_destroy(pa_plan *p) {
  if (!p) return;
  batch_cache_forget(p->id);
  free_plan_memory(p);
  free(p);
}

void *pa_plan_in(const pa_plan *p) { return p->in; }
void *pa_plan_out(const pa_plan *p) { return p->out; }

/* Runs the passes; `adjoint` runs them in reverse with swapped roles. */
static void run(const pa_plan *p, void *in, void *out, int adjoint) {
  if (!p) return;
  const double t0 = wall_seconds();
  if (adjoint && p->is_r2r) {
    pa_error("the adjoint pipeline is not available for r2r plans");
    return;
  }
  const size_t in_bytes = (size_t)(adjoint ? p->out_elems * p->out_esz : p->in_elems * p->in_esz);
  const size_t out_bytes = (size_t)(adjoint ? p->in_elems * p->in_esz : p->out_elems * p->out_esz);
  void *din = in, *dout = out, *stage_in = NULL, *stage_out = NULL;
  const int host_in = !pfk_is_device_ptr(in);
  const int host_out = !pfk_is_device_ptr(out);
  if (host_in) {
    if (pfk_malloc(&stage_in, in_bytes) || pfk_memcpy_h2d(stage_in, in, in_bytes)) {
      pa_error("staging input failed: %s", pfk_last_error());
      pfk_free(stage_in);
      return;
    }
    din = stage_in;
  }
  if (host_out) {
    if (in == out && host_in) {
      dout = stage_in;
      if (out_bytes > in_bytes) {
        pa_error("in-place host arrays: output larger than input");
        pfk_free(stage_in);
        return;
      }
    } else {
      if (pfk_malloc(&stage_out, out_bytes)) {
        pa_error("staging output failed: %s", pfk_last_error());
        pfk_free(stage_in);
        return;
      }
      dout = stage_out;
    }
  }
  void *bufs[4] = {din, dout, p->scratch[0], p->scratch[1]};
  int err = pfk_event_record(p->tev[0]);
  for (int i = 0; i < p->npass && !err; i++) {
    const int k = adjoint ? p->npass - 1 - i : i;
    const pfk_pass *fw = &p->pass[k];
    if (!adjoint) {
      err = pfk_launch_pass(fw, bufs[p->src_role[k]], bufs[p->dst_role[k]]);
    } else {
      pfk_pass a = *fw;
      a.sign = -fw->sign;
      a.nin = fw->nout; a.zin = fw->zout;
      a.nout = fw->nin; a.zout = fw->zin;
      a.imod = fw->omod; a.omod = fw->imod;
      a.sin = fw->sout; a.sout = fw->sin;
      a.tsin = fw->tsout; a.tsout = fw->tsin;
      for (int b = 0; b < a.nb; b++) { a.bsin[b] = fw->bsout[b]; a.bsout[b] = fw->bsin[b]; }
      a.src_elems = fw->dst_elems; a.dst_elems = fw->src_elems;
      if (fw->kind == PFK_R2C) { a.kind = PFK_C2R; a.hw = 1; }
      else if (fw->kind == PFK_C2R) { a.kind = PFK_R2C; a.hw = 2; }
      a.tw = p->tw_adj[k]; a.rtw = p->rtw_adj[k];
      /* map roles: forward IN <-> adjoint OUT */
      static const int swap_io[4] = {ROLE_OUT, ROLE_IN, ROLE_S0, ROLE_S1};
      err = pfk_launch_pass(&a, bufs[swap_io[p->dst_role[k]]], bufs[swap_io[p->src_role[k]]]);
    }
    if (!err) err = pfk_event_record(p->tev[i + 1]);
    if (err) pa_error("pass %d launch failed: %s", k, pfk_last_error());
  }
  if (!err && pfk_sync()) { pa_error("execution failed: %s", pfk_last_error()); err = 1; }
  if (!err) {
    for (int i = 0; i < p->npass; i++) {
      float ms = 0;
      const int k = adjoint ? p->npass - 1 - i : i;
      if (!pfk_event_elapsed_ms(p->tev[i], p->tev[i + 1], &ms))
        p->timer->trafo[p->tslot[k]] += 1e-3 * ms;
    }
  }
  if (host_out && !err) {
    if (pfk_memcpy_d2h(out, dout, out_bytes)) pa_error("copy-out failed: %s", pfk_last_error());
  }
  pfk_free(stage_in);
  pfk_free(stage_out);
  p->timer->whole += wall_seconds() - t0;
  p->timer->iter++;
}

void pa_plan_execute(const pa_plan *p, void *in, void *out) { run(p, in, out, 0); }

pa_timer *pa_plan_timer(const pa_plan *p) { return p->timer; }

void pa_plan_info(const pa_plan *p, int *rnk_n, const INT **n, const INT **ni, const INT **no,
                  unsigned *pfft_flags, int *sign) {
  *rnk_n = p->rnk_n;
  *n = p->n;
  *ni = p->ni;
  *no = p->no;
  *pfft_flags = p->pfft_flags;
  *sign = p->sign;
}

pa_timer *pa_timer_new(int rnk_pm) {
  pa_timer *t = (pa_timer *)calloc(1, sizeof(pa_timer));
  if (!t) return NULL;
  t->rnk_pm = rnk_pm;
  t->rnk_trafo = 2 * rnk_pm + 2;
  t->rnk_remap = 2 * rnk_pm;
  return t;
}
void pa_plan_execute_adjoint(const pa_plan *p, void *in, void *out) { run(p, in, out, 1); }

/* ---- batched execution ----------------------------------------------
 * All passes of all plans become tasks of one persistent device launch
 * (kernels_batch.hip): tasks are ordered by pass level (every plan's pass
 * 0, then every plan's pass 1, ...), each depending on the same plan's
 * previous pass.  The launch is captured once into a hipGraph per set of
 * plans and replayed on later calls with the same plans. */
#define PA_BATCH_CACHE 4
typedef struct {
  int used;
  int count;
  uint64_t *ids;
  void **ins, **outs;
  pfk_batch *b;
  uint64_t stamp;
} batch_entry;
static batch_entry g_bcache[PA_BATCH_CACHE];
static uint64_t g_bstamp = 0;

static void batch_entry_clear(batch_entry *e) {
  pfk_batch_destroy(e->b);
  free(e->ids);
  free(e->ins);
  free(e->outs);
  memset(e, 0, sizeof(*e));
}

static void batch_cache_forget(uint64_t id) {
  for (int i = 0; i < PA_BATCH_CACHE; i++) {
    batch_entry *e = &g_bcache[i];
    if (!e->used) continue;
    for (int k = 0; k < e->count; k++)
      if (e->ids[k] == id) { batch_entry_clear(e); break; }
  }
}

static pfk_batch *batch_build(const pa_plan *const *plans, int count) {
  int maxp = 0, ntasks = 0;
  for (int i = 0; i < count; i++) {
    if (plans[i]->npass > maxp) maxp = plans[i]->npass;
    ntasks += plans[i]->npass;
  }
  pfk_task *tasks = (pfk_task *)calloc((size_t)ntasks, sizeof(pfk_task));
  int *last = (int *)malloc(sizeof(int) * (size_t)count);
  for (int i = 0; i < count; i++) last[i] = -1;
  int k = 0;
  for (int lev = 0; lev < maxp; lev++)
    for (int i = 0; i < count; i++) {
      const pa_plan *p = plans[i];
      if (lev >= p->npass) continue;
      void *bufs[4] = {p->in, p->out, p->scratch[0], p->scratch[1]};
      tasks[k].pass = p->pass[lev];
      tasks[k].src = bufs[p->src_role[lev]];
      tasks[k].dst = bufs[p->dst_role[lev]];
      tasks[k].dep = last[i];
      last[i] = k;
      k++;
    }
  const char *ev = getenv("PFFT_AMD_BATCH_GRAPH");
  const int use_graph = ev ? (ev[0] == '1') : 1;
  pfk_batch *b = pfk_batch_create(tasks, ntasks, use_graph);
  free(tasks);
  free(last);
  return b;
}

void pa_plans_execute_batch(const pa_plan *const *plans, int count) {
  if (count <= 0 || !plans) return;
  int uniform = 1;
  for (int i = 0; i < count; i++) {
    if (!plans[i]) { fprintf(stderr, "!!! Error: Can not execute PFFT Plan == NULL !!!\n"); return; }
    if (plans[i]->prec != plans[0]->prec || !pfk_is_device_ptr(plans[i]->in) ||
        !pfk_is_device_ptr(plans[i]->out))
      uniform = 0;
    for (int k = 0; k < plans[i]->npass; k++)
      if (plans[i]->pass[k].n1 || plans[i]->pass[k].r2r_path || plans[i]->pass[k].mfma)
        uniform = 0;   /* passes that run as several launches, or in an MFMA
                          arithmetic form the batch kernels do not have */
  }
  if (!uniform) {
    /* host arrays, mixed precisions, long lines or MFMA forms: one plan
     * after another (each in its own arithmetic form) */
    for (int i = 0; i < count; i++) run(plans[i], plans[i]->in, plans[i]->out, 0);
    return;
  }
  batch_entry *hit = NULL;
  for (int i = 0; i < PA_BATCH_CACHE && !hit; i++) {
    batch_entry *e = &g_bcache[i];
    if (!e->used || e->count != count) continue;
    int same = 1;
    for (int k = 0; k < count && same; k++)
      same = e->ids[k] == plans[k]->id && e->ins[k] == plans[k]->in && e->outs[k] == plans[k]->out;
    if (same) hit = e;
  }
  if (!hit) {
    batch_entry *slot = &g_bcache[0];
    for (int i = 0; i < PA_BATCH_CACHE; i++) {
      if (!g_bcache[i].used) { slot = &g_bcache[i]; break; }
      if (g_bcache[i].stamp < slot->stamp) slot = &g_bcache[i];
    }
    if (slot->used) batch_entry_clear(slot);
    pfk_batch *b = batch_build(plans, count);
    if (!b) {
      pa_error("batch setup failed (%s); executing plans one by one", pfk_last_error());
      for (int i = 0; i < count; i++) run(plans[i], plans[i]->in, plans[i]->out, 0);
      return;
    }
    slot->used = 1;
    slot->count = count;
    slot->b = b;
    slot->ids = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)count);
    slot->ins = (void **)malloc(sizeof(void *) * (size_t)count);
    slot->outs = (void **)malloc(sizeof(void *) * (size_t)count);
    for (int k = 0; k < count; k++) {
      slot->ids[k] = plans[k]->id;
      slot->ins[k] = plans[k]->in;
      slot->outs[k] = plans[k]->out;
    }
    hit = slot;
  }
  hit->stamp = ++g_bstamp;
  if (pfk_batch_run(hit->b)) pa_error("batched execution failed: %s", pfk_last_error());
}

/* The batch for callers that cannot allocate: the same kernels and
 * work-list as pa_plans_execute_batch, launched directly (no graph), with
 * the task list built in a fixed host array and the device work-list in the
 * library's fixed device array (pfk_batch_run_noalloc).  Returns 0 on success; 1 (with a
 * message) when the plans need more than PA_NOALLOC_TASKS passes, hold host
 * arrays, differ in precision or have passes the batch kernels do not run
 * (long lines, r2r inner transforms, MFMA forms). */
#define PA_NOALLOC_TASKS 2048
static pfk_task g_noalloc_tasks[PA_NOALLOC_TASKS];

int pa_plans_execute_batch_noalloc(const pa_plan *const *plans, int count) {
  if (count <= 0) return 0;
  if (!plans) { pa_error("execute_batch_noalloc: no plans"); return 1; }
  int ntasks = 0, maxp = 0;
  for (int i = 0; i < count; i++) {
    const pa_plan *p = plans[i];
    if (!p) { pa_error("execute_batch_noalloc: plan %d is NULL", i); return 1; }
    if (p->prec != plans[0]->prec || !pfk_is_device_ptr(p->in) || !pfk_is_device_ptr(p->out)) {
      pa_error("execute_batch_noalloc: plans must share one precision and use device arrays");
      return 1;
    }
    for (int k = 0; k < p->npass; k++)
      if (p->pass[k].n1 || p->pass[k].r2r_path || p->pass[k].mfma) {
        pa_error("execute_batch_noalloc: plan %d has passes the batch kernels do not run", i);
        return 1;
      }
    ntasks += p->npass;
    if (p->npass > maxp) maxp = p->npass;
  }
  if (ntasks > PA_NOALLOC_TASKS) {
    pa_error("execute_batch_noalloc: %d passes exceed the fixed work-list of %d", ntasks,
             PA_NOALLOC_TASKS);
    return 1;
  }
  /* level order (every plan's pass 0, then every pass 1, ...), each task
   * depending on the previous pass of its plan: within the block of level
   * lev-1 the plans with at least lev passes appear in plan order */
  int k = 0, prev_start = 0;
  for (int lev = 0; lev < maxp; lev++) {
    const int cur_start = k;
    int a = prev_start;
    for (int i = 0; i < count; i++) {
      const pa_plan *p = plans[i];
      const int in_prev = lev > 0 && p->npass > lev - 1;
      if (lev < p->npass) {
        void *bufs[4] = {p->in, p->out, p->scratch[0], p->scratch[1]};
        pfk_task *t = &g_noalloc_tasks[k++];
        memset(t, 0, sizeof(*t));
        t->pass = p->pass[lev];
        t->src = bufs[p->src_role[lev]];
        t->dst = bufs[p->dst_role[lev]];
        t->dep = lev > 0 ? a : -1;
      }
      if (in_prev) a++;
    }
    prev_start = cur_start;
  }
  if (pfk_batch_run_noalloc(g_noalloc_tasks, ntasks)) {
    pa_error("execute_batch_noalloc failed: %s", pfk_last_error());
    return 1;
  }
  return 0;
}

/* Times every pass of the plan over `iters` back-to-back executions on
 * device arrays, after one untimed execution: one chain of HIP events on
 * the library stream, one between every two passes, read after a single
 * synchronisation.  (Reading the events after each execution would let the
 * stream run dry, and the next execution's first event would then carry
 * the host's launch latency into pass 0.)  ms[k] receives the mean
 * duration of pass k; bytes[k] the bytes pass k reads plus writes (one
 * sweep of its source and destination). */
int pa_plan_time_passes(const pa_plan *p, void *in, void *out, int iters, double *ms,
                        double *bytes) {
  if (!p || iters < 1 || !pfk_is_device_ptr(in) || !pfk_is_device_ptr(out)) return 1;
  void *bufs[4] = {in, out, p->scratch[0], p->scratch[1]};
  const int nev = iters * p->npass + 1;
  void **ev = (void **)calloc((size_t)nev, sizeof(void *));
  if (!ev) return 1;
  int err = 0, made = 0;
  for (; made < nev && !err; made++)
    if (pfk_event_create(&ev[made])) err = 1;
  for (int k = 0; k < p->npass && !err; k++)
    err = pfk_launch_pass(&p->pass[k], bufs[p->src_role[k]], bufs[p->dst_role[k]]);
  if (!err) err = pfk_event_record(ev[0]);
  for (int it = 0; it < iters && !err; it++)
    for (int k = 0; k < p->npass && !err; k++) {
      err = pfk_launch_pass(&p->pass[k], bufs[p->src_role[k]], bufs[p->dst_role[k]]);
      if (!err) err = pfk_event_record(ev[it * p->npass + k + 1]);
    }
  for (int k = 0; k < p->npass; k++) ms[k] = 0;
  for (int it = 0; it < iters && !err; it++)
    for (int k = 0; k < p->npass && !err; k++) {
      float t = 0;
      const int e0 = it * p->npass + k;
      if (pfk_event_elapsed_ms(ev[e0], ev[e0 + 1], &t)) err = 1;
      ms[k] += t;
    }
  for (int k = 0; k < p->npass; k++) {
    ms[k] /= iters;
    bytes[k] = pa_pass_bytes(&p->pass[k]);
  }
  for (int i = 0; i < made; i++)
    if (ev[i]) pfk_event_destroy(ev[i]);
  free(ev);
  return err;
}

int pa_plan_npasses(const pa_plan *p) { return p ? p->npass : 0; }

static const char *kernel_name(const pfk_pass *ps) {
  if (ps->kind == PFK_COPY) return "copy";
  if (ps->n1) return "long-two-step";
  if (ps->r2r_path) return "r2r-inner-fft";
  const int fam = pfk_pass_kernels(ps);
  const int lm = (ps->variant & 8) && ps->kind != PFK_R2R &&
                 (ps->prec == PFK_F32 || pfk_pass_plain(ps));
  if (fam & PFK_HAS_MFMA) {
    static const char *nm[2][2] = {{"mfma-f32", "mfma-f32-lm"}, {"mfma-bf16x2", "mfma-bf16x2-lm"}};
    return nm[ps->mfma == 2][lm && ps->nf <= 1024];
  }
  if (ps->variant == 1) return "general";
  if (ps->variant == 3 && (fam & PFK_HAS_SMALL)) return "small";
  if (fam & PFK_HAS_POW2) {
    static const char *nm[2][4] = {
        {"fast", "fast-narrow", "fast-grid", "fast-narrow-grid"},
        {"fast-lm", "fast-narrow-lm", "fast-grid-lm", "fast-narrow-grid-lm"}};
    return nm[lm][((ps->variant & 3) == 2 ? 1 : 0) + ((ps->variant & 4) ? 2 : 0)];
  }
  if (fam & PFK_HAS_SMALL) return "small";
  return "general";
}

/* One line per pass: kind, length, strides, batch and tile geometry,
 * the kernel that runs it. */
int pa_plan_describe(const pa_plan *p, char *buf, size_t len) {
  static const char *kn[] = {"c2c", "r2c", "c2r", "copy", "r2r"};   /* enum pfk_kind */
  size_t o = 0;
  buf[0] = 0;
  for (int k = 0; k < p->npass && o < len; k++) {
    const pfk_pass *ps = &p->pass[k];
    o += (size_t)snprintf(buf + o, len - o,
                          "pass %d: %s n=%lld nf=%lld nin=%lld@%lld nout=%lld@%lld sin=%lld "
                          "sout=%lld tile=%lld(%lld,%lld) nb=%d roles=%d->%d kernel=%s%s\n",
                          k, (ps->kind >= 0 && ps->kind <= PFK_R2R) ? kn[ps->kind] : "?",
                          (long long)ps->n, (long long)ps->nf,
                          (long long)ps->nin, (long long)ps->zin, (long long)ps->nout,
                          (long long)ps->zout, (long long)ps->sin, (long long)ps->sout,
                          (long long)ps->text, (long long)ps->tsin, (long long)ps->tsout, ps->nb,
                          p->src_role[k], p->dst_role[k], kernel_name(ps),
                          (ps->xcd_remap && !strncmp(kernel_name(ps), "fast", 4)) ? "+xcd" : "");
  }
  return 0;
}
