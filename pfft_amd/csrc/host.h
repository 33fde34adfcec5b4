/*
 * host.h -- internal declarations of the C host library (planner,
 * executor, API tools).  Not installed; the public surface is
 * include/pfft.h.
 */
#ifndef PFFT_AMD_HOST_H
#define PFFT_AMD_HOST_H 1

#include <stddef.h>
#include <stdint.h>
#include "pfft.h"
#include "pfk.h"

typedef ptrdiff_t INT;

#define PA_MAXRNK 8

/* internal transform flags: same meaning as the reference's PFFTI_TRAFO_*
 * (kernel/ipfft.h:137-163) */
#define TF_C2C      (1U << 0)
#define TF_R2C      (1U << 1)
#define TF_C2R      (1U << 2)
#define TF_R2R      (0x7F8U)     /* RE00..RO11, bits 3..10 */
#define TF_SKIP     (1U << 11)
#define TF_PHANTOM  (1U << 12)
#define TF_PADDED   (1U << 13)
#define TF_RDFT     (TF_R2C | TF_C2R)

#define PA_MAX(a, b) ((a) > (b) ? (a) : (b))

/* ---- process mesh (comm.c) ---- */
typedef struct pa_mesh {
  int rnk_pm;          /* Cartesian dimension (a non-Cartesian comm is a 1-d mesh) */
  int np[PA_MAXRNK];   /* processes per mesh dimension (all 1 on one GPU) */
  int size;            /* total number of processes */
} pa_mesh;

/* Fills *m for `comm`; returns 0 if the communicator is usable by this
 * one-process library. */
int pa_mesh_of(MPI_Comm comm, pa_mesh *m);
int pa_create_procmesh(int rnk, MPI_Comm comm, const int *np, MPI_Comm *comm_cart);

/* ---- layout machinery (layout.c) ---- */
INT pa_local_size(int rnk_n, const INT *n, const INT *ni, const INT *no, INT howmany,
                  const INT *iblock, const INT *oblock, const pa_mesh *m,
                  unsigned trafo_flag, unsigned pfft_flags,
                  INT *lni, INT *lis, INT *lno, INT *los);
void pa_local_block(int rnk_n, const INT *ni, const INT *no, const INT *iblock,
                    const INT *oblock, const pa_mesh *m, int pid, unsigned trafo_flag,
                    unsigned pfft_flags, INT *lni, INT *lis, INT *lno, INT *los);
INT pa_physical_size_1d(INT n, unsigned trafo_flag);

/* ---- plans (plan.c) ---- */
typedef struct pa_plan pa_plan;
pa_plan *pa_plan_create(int prec, int rnk_n, const INT *n, const INT *ni, const INT *no,
                        INT howmany, const INT *iblock, const INT *oblock, void *in,
                        void *out, MPI_Comm comm, int sign, const int *skip_trafos,
                        unsigned trafo_flag, unsigned pfft_flags, const int *kinds);
void pa_plan_execute(const pa_plan *p, void *in, void *out);
void pa_plan_execute_adjoint(const pa_plan *p, void *in, void *out);
void pa_plan_destroy(pa_plan *p);
void *pa_plan_in(const pa_plan *p);
void *pa_plan_out(const pa_plan *p);
void pa_plans_execute_batch(const pa_plan *const *plans, int count);
int pa_plans_execute_batch_noalloc(const pa_plan *const *plans, int count);
int pa_plan_time_passes(const pa_plan *p, void *in, void *out, int iters, double *ms,
                        double *bytes);
int pa_plan_npasses(const pa_plan *p);
int pa_plan_describe(const pa_plan *p, char *buf, size_t len);

/* execution timer: the fields of the reference's timer (kernel/ipfft.h:
 * 208-219, kernel/timer.c:358-373), in seconds */
typedef struct pa_timer {
  int rnk_pm, rnk_trafo, rnk_remap, iter;
  double whole;
  double trafo[2 * PA_MAXRNK + 2];
  double remap[2 * PA_MAXRNK];
  double remap_3dto2d[2];
  double itwiddle, otwiddle;
} pa_timer;
pa_timer *pa_timer_new(int rnk_pm);
pa_timer *pa_plan_timer(const pa_plan *p);
void pa_plan_info(const pa_plan *p, int *rnk_n, const INT **n, const INT **ni, const INT **no,
                  unsigned *pfft_flags, int *sign);

/* ---- autotuner and wisdom (tune.c) ---- */
int pa_tune_lookup(pfk_pass *ps, unsigned effort);
int pa_tune_measure(pfk_pass *ps, const void *src, void *dst, unsigned effort);
double pa_pass_bytes(const pfk_pass *ps);
int pa_wisdom_export(const char *filename);
int pa_wisdom_import(const char *filename);
void pa_wisdom_forget(void);
int pa_wisdom_count(void);
int pa_tune_timings(void);

/* ---- ghost cells (gcell.c) ---- */
/* exchange / reduce timer: gcell/gctimer.c fields (seconds) */
typedef struct pa_gctimer {
  int iter;
  double whole, pad_zeros, exchange;
} pa_gctimer;
typedef struct pa_gcplan pa_gcplan;
INT pa_local_size_gc(int rnk_n, const INT *local_n, const INT *local_start, INT howmany,
                     const INT *gc_below, const INT *gc_above, INT *local_ngc,
                     INT *local_gc_start);
pa_gcplan *pa_gcplan_create(int prec, int rnk_n, const INT *n, INT howmany, const INT *block,
                            const INT *gc_below, const INT *gc_above, void *data,
                            MPI_Comm comm, unsigned gc_flags, int complex_data);
void pa_gc_exchange(pa_gcplan *p);
void pa_gc_reduce(pa_gcplan *p);
void pa_gcplan_destroy(pa_gcplan *p);
pa_gctimer *pa_gcplan_timer(pa_gcplan *p, int which);   /* 0 exchange, 1 reduce */
void pa_gcplan_describe(const pa_gcplan *p, INT *n, INT *gb, INT *ga);

/* ---- twiddles (twiddle.c) ---- */
/* exp(sign*2*pi*i*m/N) for m in [0, count), rounded to double, computed in
 * long double with exact integer quadrant reduction. */
void pa_twiddles(long N, long count, int sign, double *out_interleaved);
/* the same for m = step*j, j in [0, count) */
void pa_twiddles_step(long N, long count, long step, int sign, double *out_interleaved);

/* ---- errors ---- */
void pa_error(const char *fmt, ...);

#endif
