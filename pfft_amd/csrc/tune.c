/*
 * tune.c -- plan-time selection of launch variants ("autotuner") and the
 * wisdom cache that remembers the selections.
 *
 * The reference hands each serial step to FFTW's planner with the effort
 * flag the user chose (kernel/partrafo.c:991-1012: PFFT_ESTIMATE, the
 * default PFFT_MEASURE, PFFT_PATIENT, PFFT_EXHAUSTIVE).  Here the same
 * flags decide whether a pass's launch variant is chosen by timing:
 *   ESTIMATE   no timing; the built-in default of each pass;
 *   MEASURE    every candidate launched 1+3 times, fastest mean kept;
 *   PATIENT    1+6 launches per candidate;
 *   EXHAUSTIVE 1+10 launches per candidate.
 * Candidates of a pass (pfk_pass.variant, pfk_pass.xcd_remap), each where
 * a kernel exists for the pass shape:
 *   (0,0) power-of-two register kernel, default tile width;
 *   (0,1) the same with the XCD-aware workgroup order;
 *   (2,0) (2,1) power-of-two register kernel, half the tile width;
 *   (4,r) (6,r) as (0,r) (2,r) with one workgroup per tile instead of
 *         the persistent grid;
 *   (8,r) (12,r) as (0,r) (4,r) with line-major lane order (complex64
 *         passes whose lines are contiguous on at least one side);
 *   (3,0) short-length kernel, one line per lane;
 *   (1,0) the general kernel.
 * A pass with one candidate is not timed.  All candidates compute the
 * same transform (the power-of-two variants bit-identically, the others
 * with their own factorisations).  A pass run by an MFMA kernel keeps it
 * (its arithmetic form is the user's choice, not a tuning parameter); only
 * its workgroup order is timed.
 *
 * Timing never touches the user's arrays: the plan passes the tuner
 * temporary device buffers of the same sizes.  Passes that move fewer
 * than PA_TUNE_MIN_BYTES are not timed (launch overhead dominates and the
 * default is used).
 *
 * Wisdom: every timed selection is stored under a key made of the pass's
 * geometry; a later plan with a pass of the same geometry reuses it
 * without timing.  The cache is process-wide and can be written to and
 * read from a text file (pfft_amd_export_wisdom / pfft_amd_import_wisdom);
 * with PFFT_AMD_WISDOM=<file> it is loaded before the first tuned plan and
 * rewritten whenever a selection is added.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "host.h"

#define PA_TUNE_MIN_BYTES (16.0 * 1024 * 1024)
#define PA_WISDOM_MAGIC "pfft-amd-wisdom 1"
#define PA_KEY_LEN 320

typedef struct {
  char key[PA_KEY_LEN];
  int variant, remap;
  double ms;
} wisdom_entry;

static wisdom_entry *g_wis = NULL;
static int g_nwis = 0, g_capwis = 0;
static int g_env_loaded = 0;
static int g_timed = 0;   /* candidate timings performed (statistics) */

/* Geometry key of a pass: everything that decides which launch is
 * fastest, nothing that only changes the values (twiddle pointers,
 * modulation starts). */
static void pass_key(const pfk_pass *p, char *key) {
  int o = snprintf(key, PA_KEY_LEN,
                   "k%d.p%d.n%lld.f%lld.i%lld@%lld.o%lld@%lld.s%lld:%lld.t%lld:%lld:%lld.h%d.m%d.mo%d%d",
                   p->kind, p->prec, (long long)p->n, (long long)p->nf, (long long)p->nin,
                   (long long)p->zin, (long long)p->nout, (long long)p->zout, (long long)p->sin,
                   (long long)p->sout, (long long)p->text, (long long)p->tsin,
                   (long long)p->tsout, p->hw, p->mfma, p->imod.on, p->omod.on);
  for (int b = 0; b < p->nb && o < PA_KEY_LEN; b++)
    o += snprintf(key + o, (size_t)(PA_KEY_LEN - o), ".b%lld:%lld:%lld", (long long)p->bext[b],
                  (long long)p->bsin[b], (long long)p->bsout[b]);
}

static wisdom_entry *wisdom_find(const char *key) {
  for (int i = 0; i < g_nwis; i++)
    if (!strcmp(g_wis[i].key, key)) return &g_wis[i];
  return NULL;
}

static int wisdom_put(const char *key, int variant, int remap, double ms) {
  wisdom_entry *e = wisdom_find(key);
  if (!e) {
    if (g_nwis == g_capwis) {
      const int cap = g_capwis ? 2 * g_capwis : 64;
      wisdom_entry *n = (wisdom_entry *)realloc(g_wis, sizeof(wisdom_entry) * (size_t)cap);
      if (!n) return 1;
      g_wis = n;
      g_capwis = cap;
    }
    e = &g_wis[g_nwis++];
    strncpy(e->key, key, PA_KEY_LEN - 1);
    e->key[PA_KEY_LEN - 1] = 0;
  }
  e->variant = variant;
  e->remap = remap;
  e->ms = ms;
  return 0;
}

int pa_wisdom_export(const char *filename) {
  if (!filename) return 1;
  FILE *f = fopen(filename, "w");
  if (!f) { pa_error("cannot write wisdom file %s", filename); return 1; }
  fprintf(f, "%s %d\n", PA_WISDOM_MAGIC, g_nwis);
  for (int i = 0; i < g_nwis; i++)
    fprintf(f, "%s %d %d %.6f\n", g_wis[i].key, g_wis[i].variant, g_wis[i].remap, g_wis[i].ms);
  const int bad = ferror(f);
  return (fclose(f) != 0) || bad;
}

/* Reads a file written by pa_wisdom_export; entries replace cached ones
 * with the same key.  Returns 0 on success, 1 if the file is missing or
 * malformed (the cache then keeps what was read before the error). */
int pa_wisdom_import(const char *filename) {
  if (!filename) return 1;
  FILE *f = fopen(filename, "r");
  if (!f) return 1;
  char line[PA_KEY_LEN + 64];
  int count = -1, ok = 1;
  if (!fgets(line, sizeof line, f) || strncmp(line, PA_WISDOM_MAGIC, strlen(PA_WISDOM_MAGIC)) ||
      sscanf(line + strlen(PA_WISDOM_MAGIC), "%d", &count) != 1 || count < 0)
    ok = 0;
  for (int i = 0; ok && i < count; i++) {
    char key[PA_KEY_LEN];
    int v, r;
    double ms;
    if (!fgets(line, sizeof line, f) || sscanf(line, "%319s %d %d %lf", key, &v, &r, &ms) != 4 ||
        v < 0 || v > 14 || ((v & 1) && v != 1 && v != 3) || r < 0 || r > 1) {
      ok = 0;
      break;
    }
    if (wisdom_put(key, v, r, ms)) ok = 0;
  }
  fclose(f);
  if (!ok) pa_error("malformed wisdom file %s", filename);
  return !ok;
}

void pa_wisdom_forget(void) {
  free(g_wis);
  g_wis = NULL;
  g_nwis = g_capwis = 0;
}

int pa_wisdom_count(void) { return g_nwis; }
int pa_tune_timings(void) { return g_timed; }

static void wisdom_env_load(void) {
  if (g_env_loaded) return;
  g_env_loaded = 1;
  const char *w = getenv("PFFT_AMD_WISDOM");
  if (w && *w) (void)pa_wisdom_import(w);
}

static void wisdom_env_save(void) {
  const char *w = getenv("PFFT_AMD_WISDOM");
  if (w && *w) (void)pa_wisdom_export(w);
}

double pa_pass_bytes(const pfk_pass *ps) {
  const double rs = (ps->prec == PFK_F64) ? 8.0 : 4.0;
  const double src_sz = (ps->kind == PFK_R2C || ps->kind == PFK_R2R) ? rs : 2 * rs;
  const double dst_sz = (ps->kind == PFK_C2R || ps->kind == PFK_R2R) ? rs : 2 * rs;
  double nl = (double)ps->text;
  for (int b = 0; b < ps->nb; b++) nl *= (double)ps->bext[b];
  return nl * ((double)ps->nin * src_sz + (double)ps->nout * dst_sz);
}

/* Mean duration of `reps` launches after one warm-up launch; < 0 on a
 * launch error. */
static double time_launches(const pfk_pass *ps, const void *src, void *dst, int reps, void *e0,
                            void *e1) {
  if (pfk_launch_pass(ps, src, dst)) return -1;
  if (pfk_event_record(e0)) return -1;
  for (int i = 0; i < reps; i++)
    if (pfk_launch_pass(ps, src, dst)) return -1;
  if (pfk_event_record(e1)) return -1;
  float t = 0;
  if (pfk_event_elapsed_ms(e0, e1, &t)) return -1;
  g_timed++;
  return (double)t / reps;
}

/* Settles a pass without timing where possible: returns 0 if the pass
 * keeps its default (tuning off, PFFT_ESTIMATE, small pass, no variants)
 * or took its variant from the wisdom cache; 1 if it must be timed with
 * pa_tune_measure.  `effort` is the plan's PFFT_ESTIMATE / PFFT_PATIENT /
 * PFFT_EXHAUSTIVE bits (none = PFFT_MEASURE). */
int pa_tune_lookup(pfk_pass *ps, unsigned effort) {
  const char *ev = getenv("PFFT_AMD_TUNE");
  const int mode = ev ? atoi(ev) : 1;   /* 0 off, 1 on, 2 on for every pass size */
  if (mode == 0 || (effort & PFFT_ESTIMATE) || ps->kind == PFK_COPY) return 0;
  const int fam = pfk_pass_kernels(ps);
  if (fam == 0) return 0;   /* general kernel only */
  if (mode != 2 && pa_pass_bytes(ps) < PA_TUNE_MIN_BYTES) return 0;
  wisdom_env_load();
  char key[PA_KEY_LEN];
  pass_key(ps, key);
  const wisdom_entry *w = wisdom_find(key);
  if (!w) return 1;
  ps->variant = w->variant;
  ps->xcd_remap = w->remap;
  return 0;
}

/* Times the candidates of a pass that pa_tune_lookup left open and keeps
 * the fastest in the pass and in the wisdom cache.  src and dst are device
 * buffers the pass may read and overwrite.  On a device error the pass
 * keeps its default and 1 is returned. */
int pa_tune_measure(pfk_pass *ps, const void *src, void *dst, unsigned effort) {
  char key[PA_KEY_LEN];
  pass_key(ps, key);
  const char *rv = getenv("PFFT_AMD_XCD_REMAP");
  const int remap_fixed = (rv && *rv) ? (rv[0] == '1') : -1;
  const int reps = (effort & PFFT_EXHAUSTIVE) ? 10 : (effort & PFFT_PATIENT) ? 6 : 3;
  static const int cand[14][2] = {{0, 0}, {0, 1}, {2, 0},  {2, 1},  {4, 0},  {4, 1},  {6, 0},
                                  {6, 1}, {8, 0}, {8, 1}, {12, 0}, {12, 1}, {3, 0}, {1, 0}};
  /* line-major lane order: c2c passes whose lines are contiguous on one
   * side and contiguous or adjacent along the tile on the other (the
   * kernels switch lane order at an LDS exchange), and real passes whose
   * complex side is contiguous too; complex128 on plain passes only */
  const int lm_shape =
      (ps->kind == PFK_C2C && (ps->sin == 1 || ps->sout == 1) && (ps->sin == 1 || ps->tsin == 1) &&
       (ps->sout == 1 || ps->tsout == 1)) ||
      (ps->kind == PFK_R2C && ps->sout == 1) || (ps->kind == PFK_C2R && ps->sin == 1);
  const int lm_ok = lm_shape && (ps->prec == PFK_F32 || pfk_pass_plain(ps));
  const int fam = pfk_pass_kernels(ps);
  void *e0 = NULL, *e1 = NULL;
  if (pfk_event_create(&e0) || pfk_event_create(&e1)) {
    if (e0) pfk_event_destroy(e0);
    return 1;
  }
  const int v0 = ps->variant, r0 = ps->xcd_remap;
  int best_v = v0, best_r = r0, err = 0;
  double best = -1;
  for (int c = 0; c < 14 && !err; c++) {
    const int v = cand[c][0], r = cand[c][1];
    const int pow2 = v != 1 && v != 3;
    if ((v & 8) && !lm_ok) continue;
    if (fam & PFK_HAS_MFMA) {   /* the user's arithmetic form: grid, lane order, XCD order vary */
      if ((v & 3) || (remap_fixed >= 0 && r != remap_fixed)) continue;
    } else if (pow2 && !(fam & PFK_HAS_POW2)) continue;
    if (v == 3 && !(fam & PFK_HAS_SMALL)) continue;
    if (pow2 && remap_fixed >= 0 && r != remap_fixed) continue;
    ps->variant = v;
    ps->xcd_remap = r;
    const double t = time_launches(ps, src, dst, reps, e0, e1);
    if (t < 0) { err = 1; break; }
    if (best < 0 || t < best) { best = t; best_v = v; best_r = r; }
  }
  pfk_event_destroy(e0);
  pfk_event_destroy(e1);
  if (pfk_sync()) err = 1;
  if (err) {
    ps->variant = v0;
    ps->xcd_remap = r0;
    return 1;
  }
  ps->variant = best_v;
  ps->xcd_remap = best_r;
  if (!wisdom_put(key, best_v, best_r, best)) wisdom_env_save();
  return 0;
}
