/*
 * api_impl.h -- the pfft.h entry points for one precision.  Included by
 * api.c twice, with
 *   PA(name)  the exported name (pfft_name / pfftf_name)
 *   R, CPLX   real and complex element types (CPLX is R[2])
 *   PA_PREC   PFK_F64 or PFK_F32
 * Each function cites the reference function whose contract it keeps.
 */

/* ---------------- setup: api/api-basic.c:198-228 ---------------- */
void PA(init)(void) {
  if (pfk_init()) pa_error("device initialisation failed: %s", pfk_last_error());
}
void PA(cleanup)(void) { (void)pfk_sync(); }
void PA(plan_with_nthreads)(int nthreads) { g_nthreads = nthreads > 0 ? nthreads : 1; }
int PA(get_nthreads)(void) { return g_nthreads; }

/* ---------------- memory: kernel/malloc.c:28-43 ---------------- */
void *PA(malloc)(size_t n) { return pa_alloc(n); }
R *PA(alloc_real)(size_t n) { return (R *)pa_alloc(sizeof(R) * n); }
CPLX *PA(alloc_complex)(size_t n) { return (CPLX *)pa_alloc(sizeof(CPLX) * n); }
void PA(free)(void *p) { pa_dealloc(p); }

/* ---------------- plans: api/api-basic.c:945-991, api/api-adv.c:91-217 ---------------- */
static PA(plan) PA(mk)(int rnk_n, const INT *n, const INT *ni, const INT *no, INT howmany,
                       const INT *iblock, const INT *oblock, void *in, void *out,
                       MPI_Comm comm, int sign, const int *skip, unsigned tf,
                       unsigned pfft_flags) {
  return (PA(plan))pa_plan_create(PA_PREC, rnk_n, n, ni, no, howmany, iblock, oblock, in, out,
                                  comm, sign, skip, tf, pfft_flags, NULL);
}

/* r2r plans: api/api-adv.c:141-155, 204-217 (sign 0, one kind per dim) */
PA(plan) PA(plan_many_r2r_skipped)(int rnk_n, const INT *n, const INT *ni, const INT *no,
                                   INT howmany, const INT *iblock, const INT *oblock,
                                   const int *skip_trafos, R *in, R *out, MPI_Comm comm_cart,
                                   const PA(r2r_kind) *kinds, unsigned pfft_flags) {
  int k[PA_MAXRNK];
  if (rnk_n < 1 || rnk_n > PA_MAXRNK || !kinds) {
    pa_error("r2r plans need 1..%d dimensions and one kind per dimension", PA_MAXRNK);
    return NULL;
  }
  for (int t = 0; t < rnk_n; t++) k[t] = (int)kinds[t];
  return (PA(plan))pa_plan_create(PA_PREC, rnk_n, n, ni, no, howmany, iblock, oblock, in, out,
                                  comm_cart, 0, skip_trafos, TF_R2R, pfft_flags, k);
}
PA(plan) PA(plan_many_r2r)(int rnk_n, const INT *n, const INT *ni, const INT *no, INT howmany,
                           const INT *iblock, const INT *oblock, R *in, R *out,
                           MPI_Comm comm_cart, const PA(r2r_kind) *kinds, unsigned pfft_flags) {
  return PA(plan_many_r2r_skipped)(rnk_n, n, ni, no, howmany, iblock, oblock, NULL, in, out,
                                   comm_cart, kinds, pfft_flags);
}
PA(plan) PA(plan_r2r)(int rnk_n, const INT *n, R *in, R *out, MPI_Comm comm_cart,
                      const PA(r2r_kind) *kinds, unsigned pfft_flags) {
  return PA(plan_many_r2r)(rnk_n, n, n, n, 1, PFFT_DEFAULT_BLOCKS, PFFT_DEFAULT_BLOCKS, in, out,
                           comm_cart, kinds, pfft_flags);
}
PA(plan) PA(plan_r2r_3d)(const INT *n, R *in, R *out, MPI_Comm comm_cart,
                         const PA(r2r_kind) *kinds, unsigned pfft_flags) {
  return PA(plan_r2r)(3, n, in, out, comm_cart, kinds, pfft_flags);
}

PA(plan) PA(plan_many_dft)(int rnk_n, const INT *n, const INT *ni, const INT *no, INT howmany,
                           const INT *iblock, const INT *oblock, CPLX *in, CPLX *out,
                           MPI_Comm comm_cart, int sign, unsigned pfft_flags) {
  return PA(mk)(rnk_n, n, ni, no, howmany, iblock, oblock, in, out, comm_cart, sign, NULL, TF_C2C,
                pfft_flags);
}
PA(plan) PA(plan_many_dft_r2c)(int rnk_n, const INT *n, const INT *ni, const INT *no,
                               INT howmany, const INT *iblock, const INT *oblock, R *in,
                               CPLX *out, MPI_Comm comm_cart, int sign, unsigned pfft_flags) {
  return PA(mk)(rnk_n, n, ni, no, howmany, iblock, oblock, in, out, comm_cart, sign, NULL,
                r2c_flag(pfft_flags), pfft_flags);
}
PA(plan) PA(plan_many_dft_c2r)(int rnk_n, const INT *n, const INT *ni, const INT *no,
                               INT howmany, const INT *iblock, const INT *oblock, CPLX *in,
                               R *out, MPI_Comm comm_cart, int sign, unsigned pfft_flags) {
  return PA(mk)(rnk_n, n, ni, no, howmany, iblock, oblock, in, out, comm_cart, sign, NULL,
                c2r_flag(pfft_flags), pfft_flags);
}
PA(plan) PA(plan_many_dft_skipped)(int rnk_n, const INT *n, const INT *ni, const INT *no,
                                   INT howmany, const INT *iblock, const INT *oblock,
                                   const int *skip_trafos, CPLX *in, CPLX *out,
                                   MPI_Comm comm_cart, int sign, unsigned pfft_flags) {
  return PA(mk)(rnk_n, n, ni, no, howmany, iblock, oblock, in, out, comm_cart, sign, skip_trafos,
                TF_C2C, pfft_flags);
}
PA(plan) PA(plan_many_dft_r2c_skipped)(int rnk_n, const INT *n, const INT *ni, const INT *no,
                                       INT howmany, const INT *iblock, const INT *oblock,
                                       const int *skip_trafos, R *in, CPLX *out,
                                       MPI_Comm comm_cart, int sign, unsigned pfft_flags) {
  return PA(mk)(rnk_n, n, ni, no, howmany, iblock, oblock, in, out, comm_cart, sign, skip_trafos,
                r2c_flag(pfft_flags), pfft_flags);
}
PA(plan) PA(plan_many_dft_c2r_skipped)(int rnk_n, const INT *n, const INT *ni, const INT *no,
                                       INT howmany, const INT *iblock, const INT *oblock,
                                       const int *skip_trafos, CPLX *in, R *out,
                                       MPI_Comm comm_cart, int sign, unsigned pfft_flags) {
  return PA(mk)(rnk_n, n, ni, no, howmany, iblock, oblock, in, out, comm_cart, sign, skip_trafos,
                c2r_flag(pfft_flags), pfft_flags);
}
PA(plan) PA(plan_dft)(int rnk_n, const INT *n, CPLX *in, CPLX *out, MPI_Comm comm_cart,
                      int sign, unsigned pfft_flags) {
  return PA(plan_many_dft)(rnk_n, n, n, n, 1, PFFT_DEFAULT_BLOCKS, PFFT_DEFAULT_BLOCKS, in, out,
                           comm_cart, sign, pfft_flags);
}
PA(plan) PA(plan_dft_r2c)(int rnk_n, const INT *n, R *in, CPLX *out, MPI_Comm comm_cart,
                          int sign, unsigned pfft_flags) {
  return PA(plan_many_dft_r2c)(rnk_n, n, n, n, 1, PFFT_DEFAULT_BLOCKS, PFFT_DEFAULT_BLOCKS, in,
                               out, comm_cart, sign, pfft_flags);
}
PA(plan) PA(plan_dft_c2r)(int rnk_n, const INT *n, CPLX *in, R *out, MPI_Comm comm_cart,
                          int sign, unsigned pfft_flags) {
  return PA(plan_many_dft_c2r)(rnk_n, n, n, n, 1, PFFT_DEFAULT_BLOCKS, PFFT_DEFAULT_BLOCKS, in,
                               out, comm_cart, sign, pfft_flags);
}
PA(plan) PA(plan_dft_3d)(const INT *n, CPLX *in, CPLX *out, MPI_Comm comm_cart, int sign,
                         unsigned pfft_flags) {
  return PA(plan_dft)(3, n, in, out, comm_cart, sign, pfft_flags);
}
PA(plan) PA(plan_dft_r2c_3d)(const INT *n, R *in, CPLX *out, MPI_Comm comm_cart, int sign,
                             unsigned pfft_flags) {
  return PA(plan_dft_r2c)(3, n, in, out, comm_cart, sign, pfft_flags);
}
PA(plan) PA(plan_dft_c2r_3d)(const INT *n, CPLX *in, R *out, MPI_Comm comm_cart, int sign,
                             unsigned pfft_flags) {
  return PA(plan_dft_c2r)(3, n, in, out, comm_cart, sign, pfft_flags);
}

/* ---------------- execution: api/api-basic.c:1109-1147, 239-249 ---------------- */
void PA(execute)(const PA(plan) plan) {
  if (!plan) { fprintf(stderr, "!!! Error: Can not execute PFFT Plan == NULL !!!\n"); return; }
  pa_plan_execute((const pa_plan *)plan, pa_plan_in((const pa_plan *)plan),
                  pa_plan_out((const pa_plan *)plan));
}
void PA(execute_dft)(const PA(plan) plan, CPLX *in, CPLX *out) {
  if (!plan) { fprintf(stderr, "!!! Error: Can not execute PFFT Plan == NULL !!!\n"); return; }
  pa_plan_execute((const pa_plan *)plan, in, out);
}
void PA(execute_dft_r2c)(const PA(plan) plan, R *in, CPLX *out) {
  if (!plan) { fprintf(stderr, "!!! Error: Can not execute PFFT Plan == NULL !!!\n"); return; }
  pa_plan_execute((const pa_plan *)plan, in, out);
}
void PA(execute_dft_c2r)(const PA(plan) plan, CPLX *in, R *out) {
  if (!plan) { fprintf(stderr, "!!! Error: Can not execute PFFT Plan == NULL !!!\n"); return; }
  pa_plan_execute((const pa_plan *)plan, in, out);
}
void PA(execute_r2r)(const PA(plan) plan, R *in, R *out) {
  if (!plan) { fprintf(stderr, "!!! Error: Can not execute PFFT Plan == NULL !!!\n"); return; }
  pa_plan_execute((const pa_plan *)plan, in, out);
}
void PA(execute_adjoint)(const PA(plan) plan, void *in, void *out) {
  if (!plan) { fprintf(stderr, "!!! Error: Can not execute PFFT Plan == NULL !!!\n"); return; }
  pa_plan_execute_adjoint((const pa_plan *)plan, in, out);
}
void PA(execute_batch)(const PA(plan) *plans, int count) {
  pa_plans_execute_batch((const pa_plan *const *)plans, count);
}
void PA(destroy_plan)(PA(plan) plan) {
  if (!plan) {
    fprintf(stderr, "!!! Error: Can not finalize PFFT Plan == NULL !!!\n");
    return;
  }
  pa_plan_destroy((pa_plan *)plan);
}

/* ---------------- layout: api/api-basic.c:736-773,887-941,993-1039; api/api-adv.c:30-88,327-385 ---------------- */
static INT PA(ls)(int rnk_n, const INT *n, const INT *ni, const INT *no, INT howmany,
                  const INT *iblock, const INT *oblock, MPI_Comm comm, unsigned tf,
                  unsigned pfft_flags, INT *lni, INT *lis, INT *lno, INT *los) {
  pa_mesh m;
  if (pa_mesh_of(comm, &m)) return 0;
  return pa_local_size(rnk_n, n, ni, no, howmany, iblock, oblock, &m, tf, pfft_flags, lni, lis,
                       lno, los);
}
INT PA(local_size_many_dft)(int rnk_n, const INT *n, const INT *ni, const INT *no, INT howmany,
                            const INT *iblock, const INT *oblock, MPI_Comm comm_cart,
                            unsigned pfft_flags, INT *lni, INT *lis, INT *lno, INT *los) {
  return PA(ls)(rnk_n, n, ni, no, howmany, iblock, oblock, comm_cart, TF_C2C, pfft_flags, lni,
                lis, lno, los);
}
INT PA(local_size_many_dft_r2c)(int rnk_n, const INT *n, const INT *ni, const INT *no,
                                INT howmany, const INT *iblock, const INT *oblock,
                                MPI_Comm comm_cart, unsigned pfft_flags, INT *lni, INT *lis,
                                INT *lno, INT *los) {
  return PA(ls)(rnk_n, n, ni, no, howmany, iblock, oblock, comm_cart, r2c_flag(pfft_flags),
                pfft_flags, lni, lis, lno, los);
}
INT PA(local_size_many_dft_c2r)(int rnk_n, const INT *n, const INT *ni, const INT *no,
                                INT howmany, const INT *iblock, const INT *oblock,
                                MPI_Comm comm_cart, unsigned pfft_flags, INT *lni, INT *lis,
                                INT *lno, INT *los) {
  return PA(ls)(rnk_n, n, ni, no, howmany, iblock, oblock, comm_cart, c2r_flag(pfft_flags),
                pfft_flags, lni, lis, lno, los);
}
INT PA(local_size_dft)(int rnk_n, const INT *n, MPI_Comm comm_cart, unsigned pfft_flags, INT *lni,
                       INT *lis, INT *lno, INT *los) {
  return PA(local_size_many_dft)(rnk_n, n, n, n, 1, PFFT_DEFAULT_BLOCKS, PFFT_DEFAULT_BLOCKS,
                                 comm_cart, pfft_flags, lni, lis, lno, los);
}
INT PA(local_size_dft_r2c)(int rnk_n, const INT *n, MPI_Comm comm_cart, unsigned pfft_flags,
                           INT *lni, INT *lis, INT *lno, INT *los) {
  return PA(local_size_many_dft_r2c)(rnk_n, n, n, n, 1, PFFT_DEFAULT_BLOCKS,
                                     PFFT_DEFAULT_BLOCKS, comm_cart, pfft_flags, lni, lis, lno,
                                     los);
}
INT PA(local_size_dft_c2r)(int rnk_n, const INT *n, MPI_Comm comm_cart, unsigned pfft_flags,
                           INT *lni, INT *lis, INT *lno, INT *los) {
  return PA(local_size_many_dft_c2r)(rnk_n, n, n, n, 1, PFFT_DEFAULT_BLOCKS,
                                     PFFT_DEFAULT_BLOCKS, comm_cart, pfft_flags, lni, lis, lno,
                                     los);
}
INT PA(local_size_dft_3d)(const INT *n, MPI_Comm comm_cart, unsigned pfft_flags, INT *lni,
                          INT *lis, INT *lno, INT *los) {
  return PA(local_size_dft)(3, n, comm_cart, pfft_flags, lni, lis, lno, los);
}
INT PA(local_size_dft_r2c_3d)(const INT *n, MPI_Comm comm_cart, unsigned pfft_flags, INT *lni,
                              INT *lis, INT *lno, INT *los) {
  return PA(local_size_dft_r2c)(3, n, comm_cart, pfft_flags, lni, lis, lno, los);
}
INT PA(local_size_dft_c2r_3d)(const INT *n, MPI_Comm comm_cart, unsigned pfft_flags, INT *lni,
                              INT *lis, INT *lno, INT *los) {
  return PA(local_size_dft_c2r)(3, n, comm_cart, pfft_flags, lni, lis, lno, los);
}

/* r2r layout: api/api-adv.c:76-88, api/api-basic.c:775-785, 929-940 */
INT PA(local_size_many_r2r)(int rnk_n, const INT *n, const INT *ni, const INT *no, INT howmany,
                            const INT *iblock, const INT *oblock, MPI_Comm comm_cart,
                            unsigned pfft_flags, INT *lni, INT *lis, INT *lno, INT *los) {
  return PA(ls)(rnk_n, n, ni, no, howmany, iblock, oblock, comm_cart, TF_R2R, pfft_flags, lni,
                lis, lno, los);
}
INT PA(local_size_r2r)(int rnk_n, const INT *n, MPI_Comm comm_cart, unsigned pfft_flags, INT *lni,
                       INT *lis, INT *lno, INT *los) {
  return PA(local_size_many_r2r)(rnk_n, n, n, n, 1, PFFT_DEFAULT_BLOCKS, PFFT_DEFAULT_BLOCKS,
                                 comm_cart, pfft_flags, lni, lis, lno, los);
}
INT PA(local_size_r2r_3d)(const INT *n, MPI_Comm comm_cart, unsigned pfft_flags, INT *lni,
                          INT *lis, INT *lno, INT *los) {
  return PA(local_size_r2r)(3, n, comm_cart, pfft_flags, lni, lis, lno, los);
}

static void PA(lb)(int rnk_n, const INT *ni, const INT *no, const INT *iblock, const INT *oblock,
                   MPI_Comm comm, int pid, unsigned tf, unsigned pfft_flags, INT *lni, INT *lis,
                   INT *lno, INT *los) {
  pa_mesh m;
  if (pa_mesh_of(comm, &m)) return;
  pa_local_block(rnk_n, ni, no, iblock, oblock, &m, pid, tf, pfft_flags, lni, lis, lno, los);
}
void PA(local_block_many_dft)(int rnk_n, const INT *ni, const INT *no, const INT *iblock,
                              const INT *oblock, MPI_Comm comm_cart, int pid,
                              unsigned pfft_flags, INT *lni, INT *lis, INT *lno, INT *los) {
  PA(lb)(rnk_n, ni, no, iblock, oblock, comm_cart, pid, TF_C2C, pfft_flags, lni, lis, lno, los);
}
void PA(local_block_many_dft_r2c)(int rnk_n, const INT *ni, const INT *no, const INT *iblock,
                                  const INT *oblock, MPI_Comm comm_cart, int pid,
                                  unsigned pfft_flags, INT *lni, INT *lis, INT *lno, INT *los) {
  PA(lb)(rnk_n, ni, no, iblock, oblock, comm_cart, pid, r2c_flag(pfft_flags), pfft_flags, lni,
         lis, lno, los);
}
void PA(local_block_many_dft_c2r)(int rnk_n, const INT *ni, const INT *no, const INT *iblock,
                                  const INT *oblock, MPI_Comm comm_cart, int pid,
                                  unsigned pfft_flags, INT *lni, INT *lis, INT *lno, INT *los) {
  PA(lb)(rnk_n, ni, no, iblock, oblock, comm_cart, pid, c2r_flag(pfft_flags), pfft_flags, lni,
         lis, lno, los);
}
/* r2r blocks: api/api-adv.c:373-385, api/api-basic.c:872-882, 1029-1037 */
void PA(local_block_many_r2r)(int rnk_n, const INT *ni, const INT *no, const INT *iblock,
                              const INT *oblock, MPI_Comm comm_cart, int pid,
                              unsigned pfft_flags, INT *lni, INT *lis, INT *lno, INT *los) {
  PA(lb)(rnk_n, ni, no, iblock, oblock, comm_cart, pid, TF_R2R, pfft_flags, lni, lis, lno, los);
}
void PA(local_block_r2r)(int rnk_n, const INT *n, MPI_Comm comm_cart, int pid,
                         unsigned pfft_flags, INT *lni, INT *lis, INT *lno, INT *los) {
  PA(local_block_many_r2r)(rnk_n, n, n, PFFT_DEFAULT_BLOCKS, PFFT_DEFAULT_BLOCKS, comm_cart, pid,
                           pfft_flags, lni, lis, lno, los);
}
void PA(local_block_r2r_3d)(const INT *n, MPI_Comm comm_cart, int pid, unsigned pfft_flags,
                            INT *lni, INT *lis, INT *lno, INT *los) {
  PA(local_block_r2r)(3, n, comm_cart, pid, pfft_flags, lni, lis, lno, los);
}
void PA(local_block_dft)(int rnk_n, const INT *n, MPI_Comm comm_cart, int pid,
                         unsigned pfft_flags, INT *lni, INT *lis, INT *lno, INT *los) {
  PA(local_block_many_dft)(rnk_n, n, n, PFFT_DEFAULT_BLOCKS, PFFT_DEFAULT_BLOCKS, comm_cart, pid,
                           pfft_flags, lni, lis, lno, los);
}
void PA(local_block_dft_r2c)(int rnk_n, const INT *n, MPI_Comm comm_cart, int pid,
                             unsigned pfft_flags, INT *lni, INT *lis, INT *lno, INT *los) {
  PA(local_block_many_dft_r2c)(rnk_n, n, n, PFFT_DEFAULT_BLOCKS, PFFT_DEFAULT_BLOCKS, comm_cart,
                               pid, pfft_flags, lni, lis, lno, los);
}
void PA(local_block_dft_c2r)(int rnk_n, const INT *n, MPI_Comm comm_cart, int pid,
                             unsigned pfft_flags, INT *lni, INT *lis, INT *lno, INT *los) {
  PA(local_block_many_dft_c2r)(rnk_n, n, n, PFFT_DEFAULT_BLOCKS, PFFT_DEFAULT_BLOCKS, comm_cart,
                               pid, pfft_flags, lni, lis, lno, los);
}
void PA(local_block_dft_3d)(const INT *n, MPI_Comm comm_cart, int pid, unsigned pfft_flags,
                            INT *lni, INT *lis, INT *lno, INT *los) {
  PA(local_block_dft)(3, n, comm_cart, pid, pfft_flags, lni, lis, lno, los);
}
void PA(local_block_dft_r2c_3d)(const INT *n, MPI_Comm comm_cart, int pid, unsigned pfft_flags,
                                INT *lni, INT *lis, INT *lno, INT *los) {
  PA(local_block_dft_r2c)(3, n, comm_cart, pid, pfft_flags, lni, lis, lno, los);
}
void PA(local_block_dft_c2r_3d)(const INT *n, MPI_Comm comm_cart, int pid, unsigned pfft_flags,
                                INT *lni, INT *lis, INT *lno, INT *los) {
  PA(local_block_dft_c2r)(3, n, comm_cart, pid, pfft_flags, lni, lis, lno, los);
}

/* ---------------- process mesh: kernel/procmesh.c:36-89 ---------------- */
int PA(create_procmesh)(int rnk_n, MPI_Comm comm, const int *np, MPI_Comm *comm_cart) {
  return pa_create_procmesh(rnk_n, comm, np, comm_cart);
}
int PA(create_procmesh_1d)(MPI_Comm comm, int np0, MPI_Comm *comm_cart_1d) {
  int np[1] = {np0};
  return pa_create_procmesh(1, comm, np, comm_cart_1d);
}
int PA(create_procmesh_2d)(MPI_Comm comm, int np0, int np1, MPI_Comm *comm_cart_2d) {
  int np[2] = {np0, np1};
  return pa_create_procmesh(2, comm, np, comm_cart_2d);
}

/* ---------------- integer vectors: util/util.c:103-173 ---------------- */
INT PA(prod_INT)(int d, const INT *vec) {
  INT p = 1;
  for (int t = 0; t < d; t++) p *= vec[t];
  return p;
}
INT PA(sum_INT)(int d, const INT *vec) {
  INT s = 0;
  for (int t = 0; t < d; t++) s += vec[t];
  return s;
}
int PA(equal_INT)(int d, const INT *v1, const INT *v2) {
  int eq = 1;
  for (int t = 0; t < d; t++) eq *= (v1[t] == v2[t]);
  return eq;
}
void PA(vcopy_INT)(int d, const INT *v1, INT *v2) {
  for (int t = 0; t < d; t++) v2[t] = v1[t];
}
void PA(vadd_INT)(int d, const INT *v1, const INT *v2, INT *sum) {
  for (int t = 0; t < d; t++) sum[t] = v1[t] + v2[t];
}
void PA(vsub_INT)(int d, const INT *v1, const INT *v2, INT *sum) {
  for (int t = 0; t < d; t++) sum[t] = v1[t] - v2[t];
}

/* ---------------- rank-0 printing: util/util.c:175-201 ---------------- */
void PA(vfprintf)(MPI_Comm comm, FILE *stream, const char *format, va_list ap) {
  (void)comm;
  vfprintf(stream, format, ap);
}
void PA(fprintf)(MPI_Comm comm, FILE *stream, const char *format, ...) {
  va_list ap;
  va_start(ap, format);
  PA(vfprintf)(comm, stream, format, ap);
  va_end(ap);
}
void PA(printf)(MPI_Comm comm, const char *format, ...) {
  va_list ap;
  va_start(ap, format);
  PA(vfprintf)(comm, stdout, format, ap);
  va_end(ap);
}

/* ---------------- test-data tools: api/api-basic.c:60-189, 295-479 ---------------- */
void PA(init_input_complex)(int rnk_n, const INT *n, const INT *ln, const INT *ls, CPLX *data) {
  pa_tool_array(PA_PREC, PA_TOOL_INIT, AT_COMPLEX, rnk_n, n, ln, ls, data);
}
void PA(init_input_complex_hermitian)(int rnk_n, const INT *n, const INT *ln, const INT *ls,
                                      CPLX *data) {
  pa_tool_array(PA_PREC, PA_TOOL_INIT, AT_HERMITIAN, rnk_n, n, ln, ls, data);
}
void PA(init_input_real)(int rnk_n, const INT *n, const INT *ln, const INT *ls, R *data) {
  pa_tool_array(PA_PREC, PA_TOOL_INIT, AT_REAL, rnk_n, n, ln, ls, data);
}
void PA(clear_input_complex)(int rnk_n, const INT *n, const INT *ln, const INT *ls, CPLX *data) {
  pa_tool_array(PA_PREC, PA_TOOL_CLEAR, AT_COMPLEX, rnk_n, n, ln, ls, data);
}
void PA(clear_input_complex_hermitian)(int rnk_n, const INT *n, const INT *ln, const INT *ls,
                                       CPLX *data) {
  pa_tool_array(PA_PREC, PA_TOOL_CLEAR, AT_HERMITIAN, rnk_n, n, ln, ls, data);
}
void PA(clear_input_real)(int rnk_n, const INT *n, const INT *ln, const INT *ls, R *data) {
  pa_tool_array(PA_PREC, PA_TOOL_CLEAR, AT_REAL, rnk_n, n, ln, ls, data);
}
R PA(check_output_complex)(int rnk_n, const INT *n, const INT *ln, const INT *ls,
                           const CPLX *data, MPI_Comm comm) {
  (void)comm;
  return (R)pa_tool_array(PA_PREC, PA_TOOL_CHECK, AT_COMPLEX, rnk_n, n, ln, ls, (void *)data);
}
R PA(check_output_complex_hermitian)(int rnk_n, const INT *n, const INT *ln, const INT *ls,
                                     const CPLX *data, MPI_Comm comm) {
  (void)comm;
  return (R)pa_tool_array(PA_PREC, PA_TOOL_CHECK, AT_HERMITIAN, rnk_n, n, ln, ls, (void *)data);
}
R PA(check_output_real)(int rnk_n, const INT *n, const INT *ln, const INT *ls, const R *data,
                        MPI_Comm comm) {
  (void)comm;
  return (R)pa_tool_array(PA_PREC, PA_TOOL_CHECK, AT_REAL, rnk_n, n, ln, ls, (void *)data);
}
void PA(init_input_complex_3d)(const INT *n, const INT *ln, const INT *ls, CPLX *data) {
  PA(init_input_complex)(3, n, ln, ls, data);
}
void PA(init_input_complex_hermitian_3d)(const INT *n, const INT *ln, const INT *ls, CPLX *data) {
  PA(init_input_complex_hermitian)(3, n, ln, ls, data);
}
void PA(init_input_real_3d)(const INT *n, const INT *ln, const INT *ls, R *data) {
  PA(init_input_real)(3, n, ln, ls, data);
}
void PA(clear_input_complex_3d)(const INT *n, const INT *ln, const INT *ls, CPLX *data) {
  PA(clear_input_complex)(3, n, ln, ls, data);
}
void PA(clear_input_complex_hermitian_3d)(const INT *n, const INT *ln, const INT *ls,
                                         CPLX *data) {
  PA(clear_input_complex_hermitian)(3, n, ln, ls, data);
}
void PA(clear_input_real_3d)(const INT *n, const INT *ln, const INT *ls, R *data) {
  PA(clear_input_real)(3, n, ln, ls, data);
}
R PA(check_output_complex_3d)(const INT *n, const INT *ln, const INT *ls, const CPLX *data,
                              MPI_Comm comm) {
  return PA(check_output_complex)(3, n, ln, ls, data, comm);
}
R PA(check_output_complex_hermitian_3d)(const INT *n, const INT *ln, const INT *ls,
                                        const CPLX *data, MPI_Comm comm) {
  return PA(check_output_complex_hermitian)(3, n, ln, ls, data, comm);
}
R PA(check_output_real_3d)(const INT *n, const INT *ln, const INT *ls, const R *data,
                           MPI_Comm comm) {
  return PA(check_output_real)(3, n, ln, ls, data, comm);
}

/* ---------------- ghost cells: api/api-basic.c:482-590, api/api-adv.c:223-326 ---------------- */
INT PA(local_size_many_gc)(int rnk_n, const INT *local_n, const INT *local_start, INT howmany,
                           const INT *gc_below, const INT *gc_above, INT *local_ngc,
                           INT *local_gc_start) {
  return pa_local_size_gc(rnk_n, local_n, local_start, howmany, gc_below, gc_above, local_ngc,
                          local_gc_start);
}
INT PA(local_size_gc)(int rnk_n, const INT *local_n, const INT *local_start,
                      const INT *gc_below, const INT *gc_above, INT *local_ngc,
                      INT *local_gc_start) {
  return PA(local_size_many_gc)(rnk_n, local_n, local_start, 1, gc_below, gc_above, local_ngc,
                                local_gc_start);
}
INT PA(local_size_gc_3d)(const INT *local_n, const INT *local_start, const INT *gc_below,
                         const INT *gc_above, INT *local_ngc, INT *local_gc_start) {
  return PA(local_size_gc)(3, local_n, local_start, gc_below, gc_above, local_ngc,
                           local_gc_start);
}
PA(gcplan) PA(plan_many_rgc)(int rnk_n, const INT *n, INT howmany, const INT *block,
                             const INT *gc_below, const INT *gc_above, R *data,
                             MPI_Comm comm_cart, unsigned gc_flags) {
  return (PA(gcplan))pa_gcplan_create(PA_PREC, rnk_n, n, howmany, block, gc_below, gc_above,
                                      data, comm_cart, gc_flags, 0);
}
PA(gcplan) PA(plan_many_cgc)(int rnk_n, const INT *n, INT howmany, const INT *block,
                             const INT *gc_below, const INT *gc_above, CPLX *data,
                             MPI_Comm comm_cart, unsigned gc_flags) {
  return (PA(gcplan))pa_gcplan_create(PA_PREC, rnk_n, n, howmany, block, gc_below, gc_above,
                                      data, comm_cart, gc_flags, 1);
}
PA(gcplan) PA(plan_rgc)(int rnk_n, const INT *n, const INT *gc_below, const INT *gc_above,
                        R *data, MPI_Comm comm_cart, unsigned gc_flags) {
  return PA(plan_many_rgc)(rnk_n, n, 1, PFFT_DEFAULT_BLOCKS, gc_below, gc_above, data,
                           comm_cart, gc_flags);
}
PA(gcplan) PA(plan_cgc)(int rnk_n, const INT *n, const INT *gc_below, const INT *gc_above,
                        CPLX *data, MPI_Comm comm_cart, unsigned gc_flags) {
  return PA(plan_many_cgc)(rnk_n, n, 1, PFFT_DEFAULT_BLOCKS, gc_below, gc_above, data,
                           comm_cart, gc_flags);
}
PA(gcplan) PA(plan_rgc_3d)(const INT *n, const INT *gc_below, const INT *gc_above, R *data,
                           MPI_Comm comm_cart, unsigned gc_flags) {
  return PA(plan_rgc)(3, n, gc_below, gc_above, data, comm_cart, gc_flags);
}
PA(gcplan) PA(plan_cgc_3d)(const INT *n, const INT *gc_below, const INT *gc_above, CPLX *data,
                           MPI_Comm comm_cart, unsigned gc_flags) {
  return PA(plan_cgc)(3, n, gc_below, gc_above, data, comm_cart, gc_flags);
}
/* a NULL plan (zero ghost cells) does nothing: api-basic.c:556-578 */
void PA(exchange)(PA(gcplan) ths) {
  if (ths) pa_gc_exchange((pa_gcplan *)ths);
}
void PA(reduce)(PA(gcplan) ths) {
  if (ths) pa_gc_reduce((pa_gcplan *)ths);
}
void PA(destroy_gcplan)(PA(gcplan) ths) { pa_gcplan_destroy((pa_gcplan *)ths); }

/* ---------------- ghost-cell timers: gcell/gctimer.c ---------------- */
static PA(gctimer) PA(gctimer_new)(const pa_gctimer *src) {
  pa_gctimer *t = (pa_gctimer *)calloc(1, sizeof(pa_gctimer));
  if (t && src) *t = *src;
  return (PA(gctimer))t;
}
void PA(reset_gctimers)(PA(gcplan) ths) {
  if (!ths) return;
  memset(pa_gcplan_timer((pa_gcplan *)ths, 0), 0, sizeof(pa_gctimer));
  memset(pa_gcplan_timer((pa_gcplan *)ths, 1), 0, sizeof(pa_gctimer));
}
PA(gctimer) PA(get_gctimer_exg)(const PA(gcplan) ths) {
  return ths ? PA(gctimer_new)(pa_gcplan_timer((pa_gcplan *)ths, 0)) : NULL;
}
PA(gctimer) PA(get_gctimer_red)(const PA(gcplan) ths) {
  return ths ? PA(gctimer_new)(pa_gcplan_timer((pa_gcplan *)ths, 1)) : NULL;
}
PA(gctimer) PA(copy_gctimer)(const PA(gctimer) orig) {
  return PA(gctimer_new)((const pa_gctimer *)orig);
}
void PA(average_gctimer)(PA(gctimer) ths) {
  pa_gctimer *t = (pa_gctimer *)ths;
  if (!t || t->iter <= 0) return;
  t->whole /= t->iter;
  t->pad_zeros /= t->iter;
  t->exchange /= t->iter;
  t->iter = 1;
}
PA(gctimer) PA(add_gctimers)(const PA(gctimer) sum1, const PA(gctimer) sum2) {
  pa_gctimer *t = (pa_gctimer *)PA(copy_gctimer)(sum1);
  const pa_gctimer *b = (const pa_gctimer *)sum2;
  if (t && b) {
    t->iter += b->iter;
    t->whole += b->whole;
    t->pad_zeros += b->pad_zeros;
    t->exchange += b->exchange;
  }
  return (PA(gctimer))t;
}
/* one process: the maximum over ranks is the timer itself */
PA(gctimer) PA(reduce_max_gctimer)(const PA(gctimer) ths, MPI_Comm comm) {
  (void)comm;
  return PA(copy_gctimer)(ths);
}
void PA(convert_gctimer2vec)(const PA(gctimer) ths, double *times) {
  const pa_gctimer *t = (const pa_gctimer *)ths;
  times[0] = (double)t->iter;
  times[1] = t->whole;
  times[2] = t->pad_zeros;
  times[3] = t->exchange;
}
PA(gctimer) PA(convert_vec2gctimer)(const double *times) {
  pa_gctimer t = {(int)times[0], times[1], times[2], times[3]};
  return PA(gctimer_new)(&t);
}
void PA(destroy_gctimer)(PA(gctimer) ths) { free(ths); }

/* output format of gcell/gctimer.c:92-145, 274-294 (Matlab assignments;
 * index 1 = log2(1 process) + 1) */
static void PA(gc_print)(const PA(gcplan) ths, MPI_Comm comm, FILE *f, int header, int adv) {
  INT n[3], gb[3], ga[3];
  pa_gcplan_describe((const pa_gcplan *)ths, n, gb, ga);
  if (header) {
    PA(fprintf)(comm, f, "\n%% n - FFT size\n");
    PA(fprintf)(comm, f, "%% np - process grid\n");
    PA(fprintf)(comm, f, "%% procs - number of processes\n");
    PA(fprintf)(comm, f, "%% index(i) = log(procs(i)) + 1\n");
  }
  PA(fprintf)(comm, f, "\nindex(%d) = %d;  procs(%d) = %d;  ", 1, 1, 1, 1);
  PA(fprintf)(comm, f, "np_gc(%d, 1:3) = [%d %d %d];\n", 1, 1, 1, 1);
  PA(fprintf)(comm, f, "n_gc(%d, 1:3) = [%td %td %td];  ", 1, n[0], n[1], n[2]);
  PA(fprintf)(comm, f, "gc_below(%d, 1:3) = [%td %td %td];  ", 1, gb[0], gb[1], gb[2]);
  PA(fprintf)(comm, f, "gc_above(%d, 1:3) = [%td %td %td];\n", 1, ga[0], ga[1], ga[2]);
  for (int pass = 0; pass < (adv ? 2 : 1); pass++) {
    for (int w = 0; w < 2; w++) {
      const char *prefix = w ? "gcells_red" : "gcells_exg";
      const pa_gctimer *t = pa_gcplan_timer((pa_gcplan *)ths, w);
      pa_gctimer m = *t;
      PA(average_gctimer)((PA(gctimer))&m);
      if (pass == 0) {
        PA(fprintf)(comm, f, "%s_gc_iter(%d)      = %d;  ", prefix, 1, t->iter);
        PA(fprintf)(comm, f, "%s_gcells(%d)       = %.3e;\n", prefix, 1, m.whole);
      } else {
        PA(fprintf)(comm, f, "%s_gc_pad_zeros(%d) = %.3e;  ", prefix, 1, m.pad_zeros);
        PA(fprintf)(comm, f, "%s_gc_exchange(%d)  = %.3e;\n", prefix, 1, m.exchange);
      }
    }
  }
}
void PA(print_average_gctimer)(const PA(gcplan) ths, MPI_Comm comm) {
  if (ths) PA(gc_print)(ths, comm, stdout, 1, 0);
}
void PA(print_average_gctimer_adv)(const PA(gcplan) ths, MPI_Comm comm) {
  if (ths) PA(gc_print)(ths, comm, stdout, 1, 1);
}
static void PA(gc_write)(const PA(gcplan) ths, const char *name, MPI_Comm comm, int adv) {
  if (!ths || !name) return;
  FILE *probe = fopen(name, "r");
  const int newfile = (probe == NULL);
  if (probe) fclose(probe);
  FILE *f = fopen(name, "a+");
  if (!f) {
    PA(fprintf)(comm, stderr, "Error: Cannot open file %s.\n", name);
    return;
  }
  PA(gc_print)(ths, comm, f, newfile, adv);
  fclose(f);
}
void PA(write_average_gctimer)(const PA(gcplan) ths, const char *name, MPI_Comm comm) {
  PA(gc_write)(ths, name, comm, 0);
}
void PA(write_average_gctimer_adv)(const PA(gcplan) ths, const char *name, MPI_Comm comm) {
  PA(gc_write)(ths, name, comm, 1);
}

/* ---------------- plan timers: kernel/timer.c, api/pfft.h:386-412 ---------------- */
static PA(timer) PA(timer_dup)(const pa_timer *src) {
  pa_timer *t = (pa_timer *)malloc(sizeof(pa_timer));
  if (t && src) *t = *src;
  return (PA(timer))t;
}
void PA(reset_timer)(PA(plan) ths) {
  if (!ths) return;
  pa_timer *t = pa_plan_timer((const pa_plan *)ths);
  pa_timer *z = pa_timer_new(t->rnk_pm);
  if (z) { *t = *z; free(z); }
}
PA(timer) PA(get_timer)(const PA(plan) ths) {
  return ths ? PA(timer_dup)(pa_plan_timer((const pa_plan *)ths)) : NULL;
}
PA(timer) PA(copy_timer)(const PA(timer) orig) {
  return orig ? PA(timer_dup)((const pa_timer *)orig) : NULL;
}
void PA(average_timer)(PA(timer) ths) {
  pa_timer *t = (pa_timer *)ths;
  if (!t || t->iter <= 0) return;
  t->whole /= t->iter;
  for (int k = 0; k < t->rnk_trafo; k++) t->trafo[k] /= t->iter;
  for (int k = 0; k < t->rnk_remap; k++) t->remap[k] /= t->iter;
  for (int k = 0; k < 2; k++) t->remap_3dto2d[k] /= t->iter;
  t->itwiddle /= t->iter;
  t->otwiddle /= t->iter;
  t->iter = 1;
}
PA(timer) PA(add_timers)(const PA(timer) sum1, const PA(timer) sum2) {
  pa_timer *r = (pa_timer *)PA(copy_timer)(sum1);
  const pa_timer *b = (const pa_timer *)sum2;
  if (!r || !b) return (PA(timer))r;
  r->iter += b->iter;
  r->whole += b->whole;
  for (int k = 0; k < r->rnk_trafo; k++) r->trafo[k] += b->trafo[k];
  for (int k = 0; k < r->rnk_remap; k++) r->remap[k] += b->remap[k];
  for (int k = 0; k < 2; k++) r->remap_3dto2d[k] += b->remap_3dto2d[k];
  r->itwiddle += b->itwiddle;
  r->otwiddle += b->otwiddle;
  return (PA(timer))r;
}
/* one process: the maximum over ranks is the timer itself */
PA(timer) PA(reduce_max_timer)(const PA(timer) ths, MPI_Comm comm) {
  (void)comm;
  return PA(copy_timer)(ths);
}
/* [rnk_pm, rnk_trafo, rnk_remap, iter, whole, trafo..., remap...,
 *  remap_3dto2d[2], itwiddle, otwiddle] (kernel/timer.c:297-340); the
 * caller frees the vector */
double *PA(convert_timer2vec)(const PA(timer) ths) {
  const pa_timer *t = (const pa_timer *)ths;
  double *v = (double *)malloc(sizeof(double) * (size_t)(t->rnk_trafo + t->rnk_remap + 9));
  if (!v) return NULL;
  int m = 0;
  v[m++] = t->rnk_pm;
  v[m++] = t->rnk_trafo;
  v[m++] = t->rnk_remap;
  v[m++] = t->iter;
  v[m++] = t->whole;
  for (int k = 0; k < t->rnk_trafo; k++) v[m++] = t->trafo[k];
  for (int k = 0; k < t->rnk_remap; k++) v[m++] = t->remap[k];
  v[m++] = t->remap_3dto2d[0];
  v[m++] = t->remap_3dto2d[1];
  v[m++] = t->itwiddle;
  v[m++] = t->otwiddle;
  return v;
}
PA(timer) PA(convert_vec2timer)(const double *times) {
  pa_timer *t = pa_timer_new((int)times[0]);
  if (!t || t->rnk_pm < 0 || t->rnk_pm > PA_MAXRNK) { free(t); return NULL; }
  int m = 3;
  t->iter = (int)times[m++];
  t->whole = times[m++];
  for (int k = 0; k < t->rnk_trafo; k++) t->trafo[k] = times[m++];
  for (int k = 0; k < t->rnk_remap; k++) t->remap[k] = times[m++];
  t->remap_3dto2d[0] = times[m++];
  t->remap_3dto2d[1] = times[m++];
  t->itwiddle = times[m++];
  t->otwiddle = times[m++];
  return (PA(timer))t;
}
void PA(destroy_timer)(PA(timer) ths) { free(ths); }

/* output format of kernel/timer.c:114-220, 402-455 (Matlab assignments;
 * index 1 = log2(1 process) + 1) */
static void PA(timer_print)(const PA(plan) ths, MPI_Comm comm, FILE *f, int header, int adv) {
  int rnk, sign;
  const INT *n, *ni, *no;
  unsigned fl;
  pa_plan_info((const pa_plan *)ths, &rnk, &n, &ni, &no, &fl, &sign);
  const pa_timer *t = pa_plan_timer((const pa_plan *)ths);
  if (header) {
    PA(fprintf)(comm, f, "\n%% n - FFT size\n");
    PA(fprintf)(comm, f, "%% np - process grid\n");
    PA(fprintf)(comm, f, "%% procs - number of processes\n");
    PA(fprintf)(comm, f, "%% pfft - PFFT runtime\n");
    PA(fprintf)(comm, f, "%% index(i) = log(procs(i)) + 1\n");
  }
  const char *effort = (fl & PFFT_ESTIMATE) ? "ESTIMATE" : (fl & PFFT_PATIENT) ? "PATIENT"
                       : (fl & PFFT_EXHAUSTIVE) ? "EXHAUSTIVE" : "MEASURE";
  PA(fprintf)(comm, f, "%% pfft_flags == PFFT_%s", effort);
  if (fl & PFFT_TRANSPOSED_IN) PA(fprintf)(comm, f, " | PFFT_TRANSPOSED_IN");
  if (fl & PFFT_TRANSPOSED_OUT) PA(fprintf)(comm, f, " | PFFT_TRANSPOSED_OUT");
  if (fl & PFFT_SHIFTED_IN) PA(fprintf)(comm, f, " | PFFT_SHIFTED_IN");
  if (fl & PFFT_SHIFTED_OUT) PA(fprintf)(comm, f, " | PFFT_SHIFTED_OUT");
  PA(fprintf)(comm, f, (fl & PFFT_TUNE) ? " | PFFT_TUNE" : " | PFFT_NO_TUNE");
  if (fl & PFFT_PRESERVE_INPUT) PA(fprintf)(comm, f, " | PFFT_PRESERVE_INPUT");
  if (fl & PFFT_DESTROY_INPUT) PA(fprintf)(comm, f, " | PFFT_DESTROY_INPUT");
  if (fl & PFFT_BUFFERED_INPLACE) PA(fprintf)(comm, f, " | PFFT_BUFFERED_INPLACE");
  PA(fprintf)(comm, f, "\n%% fftw_flags == FFTW_%s\n", effort);
  PA(fprintf)(comm, f, "index(%d) = %d;  procs(%d) = %d;  ", 1, 1, 1, 1);
  PA(fprintf)(comm, f, "np_pfft(%d, 1:%d) = [", 1, t->rnk_pm);
  for (int k = 0; k < t->rnk_pm; k++) PA(fprintf)(comm, f, "%d ", 1);
  PA(fprintf)(comm, f, "];\n");
  const INT *vs[3] = {n, ni, no};
  const char *nm[3] = {"n_pfft", "ni_pfft", "no_pfft"};
  for (int v = 0; v < 3; v++) {
    PA(fprintf)(comm, f, "%s(%d, 1:%d) = [", nm[v], 1, rnk);
    for (int k = 0; k < rnk; k++) PA(fprintf)(comm, f, "%td ", vs[v][k]);
    PA(fprintf)(comm, f, v < 2 ? "];  " : "];\n");
  }
  const char *pre = (sign == PFFT_FORWARD) ? "pfft_forw" : "pfft_back";
  pa_timer m = *t;
  PA(average_timer)((PA(timer))&m);
  PA(fprintf)(comm, f, "%s_iter(%d)    = %d;  ", pre, 1, t->iter);
  PA(fprintf)(comm, f, "%s(%d)   = %.3e;\n", pre, 1, m.whole);
  if (!adv) return;
  int k = 0, l = 0;
  PA(fprintf)(comm, f, "%s_itwiddle(%d)   = %.3e;\n", pre, 1, m.itwiddle);
  PA(fprintf)(comm, f, "%s_remap_3dto2d(%d, 2)   = %.3e;\n", pre, 1, m.remap_3dto2d[0]);
  for (int h = 0; h < 2; h++) {   /* the T_OUT steps, then the T_IN steps */
    for (int q = 0; q < m.rnk_pm; q++, k++, l++) {
      PA(fprintf)(comm, f, "%s_trafo%d(%d, 2)   = %.3e;  ", pre, k + 1, 1, m.trafo[k]);
      PA(fprintf)(comm, f, "%s_remap%d(%d, 1) = %.3e;\n", pre, l + 1, 1, m.remap[l]);
    }
    PA(fprintf)(comm, f, "%s_trafo%d(%d, 2)   = %.3e;\n", pre, k + 1, 1, m.trafo[k]);
    k++;
  }
  PA(fprintf)(comm, f, "%s_remap_2dto3d(%d, 2)   = %.3e;\n", pre, 1, m.remap_3dto2d[1]);
  PA(fprintf)(comm, f, "%s_otwiddle(%d)   = %.3e;\n", pre, 1, m.otwiddle);
}
void PA(print_average_timer)(const PA(plan) ths, MPI_Comm comm) {
  if (ths) PA(timer_print)(ths, comm, stdout, 1, 0);
}
void PA(print_average_timer_adv)(const PA(plan) ths, MPI_Comm comm) {
  if (ths) PA(timer_print)(ths, comm, stdout, 1, 1);
}
static void PA(timer_write)(const PA(plan) ths, const char *name, MPI_Comm comm, int adv) {
  if (!ths || !name) return;
  FILE *probe = fopen(name, "r");
  const int newfile = (probe == NULL);
  if (probe) fclose(probe);
  FILE *f = fopen(name, "a+");
  if (!f) {
    PA(fprintf)(comm, stderr, "Error: Cannot open file %s.\n", name);
    return;
  }
  PA(timer_print)(ths, comm, f, newfile, adv);
  fclose(f);
}
void PA(write_average_timer)(const PA(plan) ths, const char *name, MPI_Comm comm) {
  PA(timer_write)(ths, name, comm, 0);
}
void PA(write_average_timer_adv)(const PA(plan) ths, const char *name, MPI_Comm comm) {
  PA(timer_write)(ths, name, comm, 1);
}

/* ---------------- array printing: api/api-basic.c:584-710 ---------------- */
/* One rank; device arrays are copied to the host first.  Format of the
 * reference: a header line per k0 slice, then rows of "%.2e" values. */
static void PA(apr)(const R *data, const INT *ln, const INT *ls, const char *name, int cplx) {
  const INT cnt = ln[0] * ln[1] * ln[2];
  if (cnt < 1 || !data) return;
  const size_t bytes = sizeof(R) * (size_t)cnt * (cplx ? 2 : 1);
  const R *h = data;
  R *copy = NULL;
  if (pfk_is_device_ptr(data)) {
    copy = (R *)malloc(bytes);
    if (!copy || pfk_memcpy_d2h(copy, data, bytes)) { free(copy); return; }
    h = copy;
  }
  fflush(stdout);
  printf("Rank %d:\n", 0);
  INT l = 0;
  for (INT k0 = 0; k0 < ln[0]; k0++) {
    printf("%s(%td,%td:%td,%td:%td):\n", name, ls[0] + k0, ls[1], ls[1] + ln[1] - 1, ls[2],
           ls[2] + ln[2] - 1);
    for (INT k1 = 0; k1 < ln[1]; k1++) {
      printf("  ");
      for (INT k2 = 0; k2 < ln[2]; k2++, l++) {
        if (cplx) printf("  %.2e + %.2ei,", (double)h[2 * l], (double)h[2 * l + 1]);
        else printf("  %.2e,", (double)h[l]);
      }
      printf("\n");
    }
  }
  printf("\n");
  fflush(stdout);
  free(copy);
}
void PA(apr_complex_3d)(const CPLX *data, const INT *local_n, const INT *local_start,
                        const char *name, MPI_Comm comm) {
  (void)comm;
  PA(apr)((const R *)data, local_n, local_start, name, 1);
}
void PA(apr_real_3d)(const R *data, const INT *local_n, const INT *local_start,
                     const char *name, MPI_Comm comm) {
  (void)comm;
  PA(apr)(data, local_n, local_start, name, 0);
}
void PA(apr_complex_permuted_3d)(const CPLX *data, const INT *local_n, const INT *local_start,
                                 int perm0, int perm1, int perm2, const char *name,
                                 MPI_Comm comm) {
  const int pm[3] = {perm0, perm1, perm2};
  INT n[3], s[3];
  for (int t = 0; t < 3; t++) { n[t] = local_n[pm[t]]; s[t] = local_start[pm[t]]; }
  PA(apr_complex_3d)(data, n, s, name, comm);
}
void PA(apr_real_permuted_3d)(const R *data, const INT *local_n, const INT *local_start,
                              int perm0, int perm1, int perm2, const char *name,
                              MPI_Comm comm) {
  const int pm[3] = {perm0, perm1, perm2};
  INT n[3], s[3];
  for (int t = 0; t < 3; t++) { n[t] = local_n[pm[t]]; s[t] = local_start[pm[t]]; }
  PA(apr_real_3d)(data, n, s, name, comm);
}

/* ---------------- command-line options: util/getargs.c:41-130 ---------------- */
/* `name` followed by neededArgs values of `type` (PFFT_INT ... PFFT_UNSIGNED)
 * is read into parameter[0..neededArgs); PFFT_SWITCH sets *(int*)parameter
 * to 1 when `name` is present.  Values are counted up to the next argument
 * that starts with '-'.  An unsupported type reports an error and reads
 * nothing. */
void PA(get_args)(int argc, char **argv, const char *name, int neededArgs, unsigned type,
                  void *parameter) {
  for (int it = 0; it < argc; it++) {
    if (strcmp(argv[it], name)) continue;
    if (type == PFFT_SWITCH) { *(int *)parameter = 1; return; }
    int given = 0;
    while (it + given + 1 < argc && argv[it + given + 1] && argv[it + given + 1][0] != '-')
      given++;
    if (given < neededArgs) {
      PA(printf)(MPI_COMM_WORLD, "!!! Warning: Not enough command line arguments for %s !!!\n",
                 name);
      return;
    }
    for (int t = 0; t < neededArgs; t++) {
      const char *a = argv[it + 1 + t];
      switch (type) {
        case PFFT_INT: ((int *)parameter)[t] = (int)strtol(a, NULL, 0); break;
        case PFFT_PTRDIFF_T: ((ptrdiff_t *)parameter)[t] = (ptrdiff_t)strtol(a, NULL, 0); break;
        case PFFT_FLOAT: ((float *)parameter)[t] = strtof(a, NULL); break;
        case PFFT_DOUBLE: ((double *)parameter)[t] = strtod(a, NULL); break;
        case PFFT_LDOUBLE: ((long double *)parameter)[t] = strtold(a, NULL); break;
        case PFFT_UNSIGNED: ((unsigned *)parameter)[t] = (unsigned)strtoul(a, NULL, 0); break;
        default:
          PA(fprintf)(MPI_COMM_WORLD, stderr,
                      "!!! Error: PFFT_DATATYPE of %s not supported. !!!\n", name);
          return;
      }
    }
    return;
  }
}
