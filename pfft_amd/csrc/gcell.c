/*
 * gcell.c -- ghost-cell plans on a one-process mesh: data distribution,
 * plan geometry, execution (device kernels in kernels_gcell.hip) and the
 * exchange/reduce timers.  The reference's module is gcell/gcells_plan.c,
 * gcell/gcells_sendrecv.c, gcell/gctimer.c with the wrappers in
 * api/api-basic.c:482-590 and api/api-adv.c:223-326.
 *
 * One process owns the whole array, so the neighbours of every process
 * are itself and the ghost cells are periodic images of the local block
 * (doc/reference.tex:558-586: "the data is wrapped periodically").
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include "host.h"

struct pa_gcplan {
  int prec;
  int rnk_pm;
  INT n[3], gb[3], ga[3], ngc[3];
  INT tuple;
  void *data;
  void *scratch;    /* device: the local block, tuple*n0*n1*n2 reals */
  void *stage;      /* device copy of a host array (box size), made on first use */
  void *ev[3];
  pfk_gc geo;
  pa_gctimer exg, red;
};

static double wall_seconds(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

/* gcell/gcells_plan.c:48-74 with util/util.c:85-94 (NULL: no ghost cells) */
INT pa_local_size_gc(int rnk_n, const INT *local_n, const INT *local_start, INT howmany,
                     const INT *gc_below, const INT *gc_above, INT *local_ngc,
                     INT *local_gc_start) {
  INT mem = howmany;
  for (int t = 0; t < rnk_n; t++) {
    const INT b = gc_below ? gc_below[t] : 0, a = gc_above ? gc_above[t] : 0;
    local_ngc[t] = b + local_n[t] + a;
    local_gc_start[t] = local_start[t] - b;
    mem *= local_ngc[t];
  }
  return mem;
}

/* gcells_plan.c:222-235: the first rnk_pm sizes in transposed order */
static void transposed_copy(int rnk_pm, unsigned gc_flags, const INT *v, INT *cp) {
  for (int t = 0; t < 3; t++) cp[t] = v[t];
  if (gc_flags & PFFT_GC_TRANSPOSED)
    for (int t = 0; t < rnk_pm; t++) cp[t] = v[(t + 1) % rnk_pm];
}

/* api/api-adv.c:232-326 and gcells_plan.c:77-200 */
pa_gcplan *pa_gcplan_create(int prec, int rnk_n, const INT *n, INT howmany, const INT *block,
                            const INT *gc_below, const INT *gc_above, void *data,
                            MPI_Comm comm, unsigned gc_flags, int complex_data) {
  pa_mesh mesh;
  INT pn[3], gb[3], ga[3];
  INT tuple = howmany;
  if (pa_mesh_of(comm, &mesh)) return NULL;
  /* NULL for zero ghost cells (gcells_plan.c:97-107), before the rank check */
  int nothing = 1;
  for (int t = 0; t < rnk_n; t++)
    if ((gc_below && gc_below[t]) || (gc_above && gc_above[t])) nothing = 0;
  if (nothing) return NULL;
  if (rnk_n != 3) {
    pa_error("ghost-cell plans are available for three dimensions only");
    return NULL;
  }
  if (mesh.size != 1) { pa_error("process mesh must have one process"); return NULL; }
  if (mesh.rnk_pm >= 3) {
    pa_error("3-d process meshes (3d-to-2d remap) are not supported on one GPU");
    return NULL;
  }
  if (howmany < 1) { pa_error("howmany must be positive"); return NULL; }
  for (int t = 0; t < 3; t++) {
    pn[t] = n[t];
    gb[t] = gc_below ? gc_below[t] : 0;
    ga[t] = gc_above ? gc_above[t] : 0;
    if (n[t] < 1 || gb[t] < 0 || ga[t] < 0) {
      pa_error("ghost-cell plan: sizes must be positive and ghost cells non-negative");
      return NULL;
    }
  }
  if (complex_data) {
    /* api-adv.c:304-316: complex data are 2*howmany reals; r2c arrays hold
     * n/2+1 complex values along the last dimension */
    gc_flags &= ~PFFT_GC_PADDED;
    if (gc_flags & PFFT_GC_R2C) {
      pn[2] = n[2] / 2 + 1;
      gc_flags &= ~PFFT_GC_R2C;
    }
    tuple = 2 * howmany;
  } else if (gc_flags & PFFT_GC_PADDED) {
    pn[2] = 2 * (n[2] / 2 + 1);   /* api-adv.c:249-250 */
  }
  pa_gcplan *p = (pa_gcplan *)calloc(1, sizeof(pa_gcplan));
  if (!p) return NULL;
  p->prec = prec;
  p->rnk_pm = mesh.rnk_pm;
  p->tuple = tuple;
  p->data = data;
  transposed_copy(mesh.rnk_pm, gc_flags, pn, p->n);
  transposed_copy(mesh.rnk_pm, gc_flags, gb, p->gb);
  transposed_copy(mesh.rnk_pm, gc_flags, ga, p->ga);
  /* blocks (given in the transposed order, gcells_plan.c:124-131): one
   * process must own every distributed dimension whole */
  for (int t = 0; t < mesh.rnk_pm; t++) {
    if (block && block[t] != PFFT_DEFAULT_BLOCK && block[t] < p->n[t]) {
      pa_error("block size %lld of dimension %d does not cover the array on one process",
               (long long)block[t], t);
      free(p);
      return NULL;
    }
  }
  INT L = tuple;
  for (int t = 0; t < 3; t++) {
    p->ngc[t] = p->gb[t] + p->n[t] + p->ga[t];
    p->geo.n[t] = p->n[t];
    p->geo.gb[t] = p->gb[t];
    p->geo.ngc[t] = p->ngc[t];
    L *= p->n[t];
  }
  p->geo.prec = prec;
  p->geo.tuple = tuple;
  const size_t rsz = (prec == PFK_F64) ? sizeof(double) : sizeof(float);
  if (pfk_malloc(&p->scratch, rsz * (size_t)L)) {
    pa_error("ghost-cell plan: device allocation failed: %s", pfk_last_error());
    free(p);
    return NULL;
  }
  for (int k = 0; k < 3; k++) {
    if (pfk_event_create(&p->ev[k])) {
      pa_error("ghost-cell plan: event creation failed: %s", pfk_last_error());
      pa_gcplan_destroy(p);
      return NULL;
    }
  }
  return p;
}

static INT box_reals(const pa_gcplan *p) {
  return p->tuple * p->ngc[0] * p->ngc[1] * p->ngc[2];
}
static INT block_reals(const pa_gcplan *p) {
  return p->tuple * p->n[0] * p->n[1] * p->n[2];
}

/* A host array is staged through a device copy of the box. */
static void *device_view(pa_gcplan *p, INT upload_reals) {
  if (pfk_is_device_ptr(p->data)) return p->data;
  const size_t rsz = (p->prec == PFK_F64) ? sizeof(double) : sizeof(float);
  if (!p->stage && pfk_malloc(&p->stage, rsz * (size_t)box_reals(p))) {
    pa_error("ghost-cell staging allocation failed: %s", pfk_last_error());
    return NULL;
  }
  if (pfk_memcpy_h2d(p->stage, p->data, rsz * (size_t)upload_reals)) {
    pa_error("ghost-cell staging copy failed: %s", pfk_last_error());
    return NULL;
  }
  return p->stage;
}

static int finish(pa_gcplan *p, void *dev, pa_gctimer *tm, double t0, int pad_first) {
  if (pfk_sync()) { pa_error("ghost-cell execution failed: %s", pfk_last_error()); return 1; }
  if (dev != p->data) {
    const size_t rsz = (p->prec == PFK_F64) ? sizeof(double) : sizeof(float);
    if (pfk_memcpy_d2h(p->data, dev, rsz * (size_t)box_reals(p))) {
      pa_error("ghost-cell staging copy failed: %s", pfk_last_error());
      return 1;
    }
  }
  float a = 0, b = 0;
  (void)pfk_event_elapsed_ms(p->ev[0], p->ev[1], &a);
  (void)pfk_event_elapsed_ms(p->ev[1], p->ev[2], &b);
  /* exchange: pad (block copy) then exchange; reduce: exchange then
   * truncate (gcells_plan.c:373-405) */
  tm->pad_zeros += 1e-3 * (pad_first ? a : b);
  tm->exchange += 1e-3 * (pad_first ? b : a);
  tm->whole += wall_seconds() - t0;
  tm->iter++;
  return 0;
}

void pa_gc_exchange(pa_gcplan *p) {
  const double t0 = wall_seconds();
  void *dev = device_view(p, block_reals(p));
  if (!dev) return;
  if (pfk_gc_exchange(&p->geo, dev, p->scratch, p->ev)) {
    pa_error("ghost-cell exchange failed: %s", pfk_last_error());
    return;
  }
  (void)finish(p, dev, &p->exg, t0, 1);
}

void pa_gc_reduce(pa_gcplan *p) {
  const double t0 = wall_seconds();
  void *dev = device_view(p, box_reals(p));
  if (!dev) return;
  if (pfk_gc_reduce(&p->geo, dev, p->scratch, p->ev)) {
    pa_error("ghost-cell reduce failed: %s", pfk_last_error());
    return;
  }
  (void)finish(p, dev, &p->red, t0, 0);
}

void pa_gcplan_destroy(pa_gcplan *p) {
  if (!p) return;
  for (int k = 0; k < 3; k++)
    if (p->ev[k]) pfk_event_destroy(p->ev[k]);
  if (p->scratch) pfk_free(p->scratch);
  if (p->stage) pfk_free(p->stage);
  free(p);
}

pa_gctimer *pa_gcplan_timer(pa_gcplan *p, int which) { return which ? &p->red : &p->exg; }

void pa_gcplan_describe(const pa_gcplan *p, INT *n, INT *gb, INT *ga) {
  for (int t = 0; t < 3; t++) {
    n[t] = p->n[t];
    gb[t] = p->gb[t];
    ga[t] = p->ga[t];
  }
}
