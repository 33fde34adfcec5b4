/*
 * layout.c -- the integer layout contract of the PFFT API on one process:
 * block sizes, local array sizes/starts and the memory each plan needs.
 *
 * This restates the reference's planner arithmetic for a mesh whose
 * every dimension has one process (coordinate 0), so each value is
 * bit-exact with what the reference reports on a 1x1 (or 1) mesh:
 *   pa_local_size      <- PX(local_size_partrafo)  kernel/partrafo.c:202-314
 *   pa_local_block     <- PX(local_block_partrafo) kernel/partrafo.c:99-199
 *   step_local_size    <- PX(local_size_partrafo_transposed)
 *                                                   kernel/partrafo-transposed.c:88-182
 *   split_steps        <- init_param_size_and_trafo_flags kernel/partrafo.c:735-834
 *   eval_blocks        <- evaluate_blocks          kernel/partrafo.c:652-701
 *   block_of           <- PX(local_block_size/offset), default_block_size
 *                                                   kernel/block.c:31-95
 * The only third-party value, FFTW-MPI's local_size_many_transposed,
 * enters through ftw_local_size_transposed() below (one process).
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "host.h"

INT pa_physical_size_1d(INT n, unsigned tf) {
  /* util/util.c:224-227 */
  return (tf & TF_RDFT) ? n / 2 + 1 : n;
}

static void phys(int rnk, const INT *n, unsigned tf, INT *pn) {
  /* util/util.c:207-221: only the last dimension changes for r2c/c2r */
  for (int t = 0; t < rnk - 1; t++) pn[t] = n[t];
  pn[rnk - 1] = pa_physical_size_1d(n[rnk - 1], tf);
}

static INT prod(int d, const INT *v) {
  INT p = 1;
  for (int t = 0; t < d; t++) p *= v[t];
  return p;
}

/* kernel/block.c:74-95 -- global block size (user value or the default,
 * rounded up) */
static INT global_block(INT N, INT user_blk, int P) {
  return (user_blk == PFFT_DEFAULT_BLOCK) ? (N + P - 1) / P : user_blk;
}

/* kernel/block.c:31-72 -- size and offset of block `which` */
static void block_of(INT N, INT blk, int which, INT *ln, INT *ls) {
  INT blocks = (blk > 0) ? (N + blk - 1) / blk : 0;
  if (which >= blocks) { *ln = 0; *ls = 0; return; }
  *ls = (INT)which * blk;
  *ln = (which == blocks - 1) ? N - (INT)which * blk : blk;
}

/* util/util.c:72-83 */
static void eval_user_blocks(int r, const INT *pn, const INT *block, const int *np, INT *out) {
  for (int t = 0; t < r; t++)
    out[t] = global_block(pn[t], block ? block[t] : PFFT_DEFAULT_BLOCK, np[t]);
}

/* kernel/partrafo.c:652-701 */
static void eval_blocks(int rnk_n, const INT *ni, const INT *no, const INT *iblock_user,
                        const INT *oblock_user, int r, const int *np, unsigned tf,
                        unsigned transp, INT *iblk, INT *mblk, INT *oblk) {
  INT pni[PA_MAXRNK], pno[PA_MAXRNK];
  phys(rnk_n, ni, tf, pni);
  phys(rnk_n, no, tf, pno);
  const INT *pnm = ((tf & TF_C2R) || (transp & PFFT_TRANSPOSED_IN)) ? pni : pno;
  const INT *mblock_user = PFFT_DEFAULT_BLOCKS;
  if (transp & PFFT_TRANSPOSED_IN) mblock_user = iblock_user;
  if (transp & PFFT_TRANSPOSED_OUT) mblock_user = oblock_user;
  for (int t = 0; t < r; t++) iblk[t] = mblk[t] = oblk[t] = 0;
  if (!(transp & PFFT_TRANSPOSED_IN)) eval_user_blocks(r, pni, iblock_user, np, iblk);
  eval_user_blocks(r, pnm + 1, mblock_user, np, mblk);
  if (!(transp & PFFT_TRANSPOSED_OUT)) eval_user_blocks(r, pno, oblock_user, np, oblk);
}

/* partrafo-transposed.c:363-411 -- non-transposed: dims 0..r-1 are
 * distributed; transposed: dims 1..r are distributed */
static void decompose(int rnk_n, const INT *n, const INT *blk, int r, unsigned tf,
                      int transposed, INT *ln, INT *ls) {
  INT pn[PA_MAXRNK];
  phys(rnk_n, n, tf, pn);
  for (int t = 0; t < rnk_n; t++) { ln[t] = pn[t]; ls[t] = 0; }
  const int off = transposed ? 1 : 0;
  for (int t = 0; t < r; t++) block_of(pn[t + off], blk[t], 0, &ln[t + off], &ls[t + off]);
}

/* partrafo-transposed.c:341-361 */
static void local_size_transposed(int rnk_n, const INT *ni, const INT *no, const INT *iblock,
                                  const INT *oblock, int r, unsigned tf, unsigned transp,
                                  INT *lni, INT *lis, INT *lno, INT *los) {
  if (transp & PFFT_TRANSPOSED_OUT) {
    decompose(rnk_n, ni, iblock, r, tf, 0, lni, lis);
    decompose(rnk_n, no, oblock, r, tf, 1, lno, los);
  } else {
    decompose(rnk_n, ni, iblock, r, tf, 1, lni, lis);
    decompose(rnk_n, no, oblock, r, tf, 0, lno, los);
  }
}

/* kernel/ousample.c:55-75 */
static INT ousam_size(INT nb, int rnk, const INT *ni, const INT *no, INT hm, unsigned tf) {
  INT pni[PA_MAXRNK], pno[PA_MAXRNK];
  phys(rnk, ni, tf, pni);
  phys(rnk, no, tf, pno);
  return PA_MAX(nb * prod(rnk, pni) * hm, nb * prod(rnk, pno) * hm);
}

/* kernel/outrafo.c:93-115 (sertrafo part: kernel/sertrafo.c:152-169) */
static INT outrafo_size(INT nb, int rnk, const INT *n, const INT *ni, const INT *no, INT hm,
                        unsigned tf) {
  INT pn[PA_MAXRNK];
  INT mem = 1;
  mem = PA_MAX(mem, ousam_size(nb, rnk, ni, n, hm, tf));
  phys(rnk, n, tf, pn);
  mem = PA_MAX(mem, nb * prod(rnk, pn) * hm);
  mem = PA_MAX(mem, ousam_size(nb, rnk, n, no, hm, tf));
  return mem;
}

/* FFTW-MPI XM(local_size_many_transposed)(2, N, hm, b0, b1, comm) on one
 * process: the local input block is min(b0,N0) x N1 and the local
 * (transposed) output block is min(b1,N1) x N0; FFTW returns hm times the
 * larger of the two (FFTW 3.3.4 mpi/api.c, published semantics). */
static INT ftw_local_size_transposed(const INT *N, INT hm, INT b0, INT b1) {
  INT l0 = (b0 > 0 && b0 < N[0]) ? b0 : N[0];
  INT l1 = (b1 > 0 && b1 < N[1]) ? b1 : N[1];
  return hm * PA_MAX(l0 * N[1], l1 * N[0]);
}

/* kernel/transpose.c:115-137 */
static INT global_transp_size(INT N0, INT N1, INT h0, INT h1, INT hm, INT blk0, INT blk1) {
  INT N[2], blk[2];
  N[0] = N1 * h1; blk[0] = blk1 * h1;
  N[1] = N0 * h0; blk[1] = blk0 * h0;
  if (N[0] * N[1] * h0 * h1 * hm == 0) return 1;
  INT mem = ftw_local_size_transposed(N, hm, blk[0], blk[1]);
  return PA_MAX(mem, 1);
}

/* kernel/transpose.c:63-112 (the T_IN branch swaps the roles) */
static void gtransp_param_core(int step, int r, const INT *ni, const INT *no, const INT *lni,
                               const INT *lno, INT tuple, const INT *iblock, const INT *oblock,
                               unsigned tf, INT *N0, INT *N1, INT *h0, INT *h1, INT *hm,
                               INT *blk0, INT *blk1) {
  *N0 = no[r - step];
  *N1 = ni[r - 1 - step];
  *h0 = 1;
  for (int t = 0; t < step; t++) *h0 *= lno[r - t];
  for (int t = 0; t < r - 1 - step; t++) *h0 *= lni[t];
  *h1 = 1;
  *hm = tuple * ((tf & TF_R2R) ? 1 : 2);
  *blk0 = oblock[r - 1 - step];
  *blk1 = iblock[r - 1 - step];
}

static void gtransp_param(int step, int r, const INT *ni, const INT *no, const INT *lni,
                          const INT *lno, INT tuple, const INT *iblock, const INT *oblock,
                          unsigned tf, unsigned transp, INT *N0, INT *N1, INT *h0, INT *h1,
                          INT *hm, INT *blk0, INT *blk1) {
  if (transp & PFFT_TRANSPOSED_IN)
    gtransp_param_core(step, r, no, ni, lno, lni, tuple, oblock, iblock, tf, N1, N0, h1, h0, hm,
                       blk1, blk0);
  else
    gtransp_param_core(step, r, ni, no, lni, lno, tuple, iblock, oblock, tf, N0, N1, h0, h1, hm,
                       blk0, blk1);
}

/* kernel/outrafo.c:46-91 (the T_IN branch swaps the local sizes) */
static void outrafo_param(int step, int r, const INT *n, const INT *ni, const INT *no,
                          const INT *lni, const INT *lno, unsigned transp, const unsigned *tfs,
                          INT *Nb, INT *N, INT *Ni, INT *No, unsigned *tf) {
  if (transp & PFFT_TRANSPOSED_IN) {
    const INT *tmp = lni;
    lni = lno;
    lno = tmp;
  }
  *Nb = 1;
  for (int t = 0; t < step; t++) *Nb *= lno[r - t];
  for (int t = 0; t < r - step; t++) *Nb *= lni[t];
  *N = n[r - step];
  *Ni = ni[r - step];
  *No = no[r - step];
  *tf = tfs[r - step];
}

/* partrafo-transposed.c:294-337 */
static INT initial_nb(int r, const INT *lni, const INT *lno, unsigned transp) {
  const INT *ln = (transp & PFFT_TRANSPOSED_IN) ? lno : lni;
  INT Nb = ln[0];
  for (int t = 1; t < r; t++) Nb *= ln[t];
  return Nb;
}

static INT tuple_size(int rnk_n, const INT *ni, const INT *no, int r, INT howmany, unsigned tf,
                      unsigned transp) {
  const INT *n = (transp & PFFT_TRANSPOSED_OUT) ? no : ni;
  INT tpl = howmany;
  if (r + 1 < rnk_n) tpl *= pa_physical_size_1d(n[rnk_n - 1], tf);
  for (int t = r + 1; t < rnk_n - 1; t++) tpl *= n[t];
  return tpl;
}

/* Overwrite of the r2c input / c2r output last dimension with real counts
 * (partrafo-transposed.c:67-85 and 158-176). */
static void real_count_overwrite(int rnk_n, const INT *ni, const INT *no, unsigned tf, INT *lni,
                                 INT *lno) {
  if (tf & TF_R2C) {
    if (tf & TF_PADDED) lni[rnk_n - 1] *= 2;
    else lni[rnk_n - 1] = ni[rnk_n - 1];
  }
  if (tf & TF_C2R) {
    if (tf & TF_PADDED) lno[rnk_n - 1] *= 2;
    else lno[rnk_n - 1] = no[rnk_n - 1];
  }
}

/* kernel/partrafo-transposed.c:88-182 */
static INT step_local_size(int rnk_n, const INT *n, const INT *ni, const INT *no, INT howmany,
                           const INT *iblock, const INT *oblock, int r, unsigned transp,
                           const unsigned *tfs, INT *lni, INT *lis, INT *lno, INT *los) {
  unsigned tf = tfs[r];
  INT mem = 1;
  local_size_transposed(rnk_n, ni, no, iblock, oblock, r, tf, transp, lni, lis, lno, los);
  INT Nb = initial_nb(r, lni, lno, transp);
  INT tpl = tuple_size(rnk_n, ni, no, r, howmany, tf, transp);
  mem = PA_MAX(mem, outrafo_size(Nb, rnk_n - r, n + r, ni + r, no + r, howmany, tf));
  INT pni[PA_MAXRNK], pno[PA_MAXRNK];
  phys(rnk_n, ni, tf, pni);
  phys(rnk_n, no, tf, pno);
  if (tf & TF_RDFT) tf = TF_C2C;
  for (int t = 0; t < r; t++) {
    INT N0, N1, h0, h1, hm, b0, b1, N, Ni, No;
    gtransp_param(t, r, pni, pno, lni, lno, tpl, iblock, oblock, tf, transp, &N0, &N1, &h0, &h1,
                  &hm, &b0, &b1);
    hm = 1; /* partrafo-transposed.c:142: mem counts complex/real units */
    mem = PA_MAX(mem, global_transp_size(N0, N1, h0, h1, hm, b0, b1));
    outrafo_param(t + 1, r, n, ni, no, lni, lno, transp, tfs, &Nb, &N, &Ni, &No, &tf);
    mem = PA_MAX(mem, outrafo_size(Nb, 1, &N, &Ni, &No, tpl, tf));
  }
  real_count_overwrite(rnk_n, ni, no, tfs[r], lni, lno);
  return mem;
}

/* kernel/partrafo.c:836-844 */
static unsigned skip_flag(const int *skip, int t) {
  return (skip && skip[t]) ? TF_SKIP : 0;
}

/* kernel/partrafo.c:735-834 -- split into the transposed-out (to) and
 * transposed-in (ti) steps; decides which step computes the FFTs. */
void pa_split_steps(int rnk_n, const INT *n, const INT *ni, const INT *no, int r, unsigned tf,
                    unsigned transp, const int *skip, INT *n_to, INT *ni_to, INT *no_to,
                    unsigned *tfs_to, INT *n_ti, INT *ni_ti, INT *no_ti, unsigned *tfs_ti) {
  int t;
  for (t = 0; t < r + 1; t++) tfs_to[t] = tf | skip_flag(skip, t);
  memcpy(ni_to, ni, sizeof(INT) * rnk_n);
  memcpy(n_to, n, sizeof(INT) * rnk_n);
  memcpy(no_to, no, sizeof(INT) * rnk_n);
  for (t = 0; t < r + 1; t++) tfs_ti[t] = tf | TF_SKIP;
  memcpy(ni_ti, no, sizeof(INT) * rnk_n);
  memcpy(n_ti, no, sizeof(INT) * rnk_n);
  memcpy(no_ti, no, sizeof(INT) * rnk_n);

  if (transp & PFFT_TRANSPOSED_IN) {
    for (t = 0; t < r + 1; t++) tfs_ti[t] = tf | skip_flag(skip, t);
    memcpy(ni_ti, ni, sizeof(INT) * rnk_n);
    memcpy(n_ti, n, sizeof(INT) * rnk_n);
    memcpy(no_ti, no, sizeof(INT) * rnk_n);
  }
  if (tf & TF_R2C) {
    for (t = 0; t < r + 1; t++) tfs_to[t] = ((t == r) ? tf : TF_C2C) | skip_flag(skip, t);
    memcpy(ni_to, ni, sizeof(INT) * rnk_n);
    memcpy(n_to, n, sizeof(INT) * rnk_n);
    memcpy(no_to, no, sizeof(INT) * rnk_n);
    for (t = 0; t < r + 1; t++) tfs_ti[t] = TF_C2C | TF_SKIP;
    for (t = 0; t < rnk_n - 1; t++) ni_ti[t] = n_ti[t] = no_ti[t] = no[t];
    ni_ti[t] = n_ti[t] = no_ti[t] = no[t] / 2 + 1;
  }
  if (tf & TF_C2R) {
    for (t = 0; t < r + 1; t++) tfs_to[t] = TF_C2C | TF_SKIP;
    for (t = 0; t < rnk_n - 1; t++) ni_to[t] = n_to[t] = no_to[t] = ni[t];
    ni_to[t] = n_to[t] = no_to[t] = ni[t] / 2 + 1;
    for (t = 0; t < r + 1; t++) tfs_ti[t] = ((t == r) ? tf : TF_C2C) | skip_flag(skip, t);
    memcpy(ni_ti, ni, sizeof(INT) * rnk_n);
    memcpy(n_ti, n, sizeof(INT) * rnk_n);
    memcpy(no_ti, no, sizeof(INT) * rnk_n);
  }
  if (transp & PFFT_TRANSPOSED_OUT) {
    for (t = 0; t < r + 1; t++) tfs_ti[t] = 0;
    for (t = 0; t < rnk_n; t++) ni_ti[t] = n_ti[t] = no_ti[t] = 0;
  }
  if (transp & PFFT_TRANSPOSED_IN) {
    /* the reference assigns 0 to its local pointer here (partrafo.c:815),
     * leaving the caller's to-flags untouched */
    for (t = 0; t < rnk_n; t++) ni_to[t] = n_to[t] = no_to[t] = 0;
  }
  /* partrafo.c:828 tests `~transp_flag & (T_IN || T_OUT)`, i.e. bit 0:
   * true unless TRANSPOSED_IN is set. */
  if (!(transp & 1U)) {
    if (tf & TF_C2R) tfs_to[0] = TF_PHANTOM;
    else tfs_ti[0] = TF_PHANTOM;
  }
}

static unsigned transp_of(unsigned pfft_flags) {
  return pfft_flags & (PFFT_TRANSPOSED_IN | PFFT_TRANSPOSED_OUT);
}

/* kernel/partrafo.c:703-732: which step's arrays are the user's */
static void alias_local(INT *lni, INT *lis, INT *dln, INT *dls, INT *lno, INT *los,
                        unsigned transp, INT **lni_to, INT **lis_to, INT **lno_to,
                        INT **los_to, INT **lni_ti, INT **lis_ti, INT **lno_ti, INT **los_ti) {
  *lni_to = lni; *lis_to = lis; *lno_to = dln; *los_to = dls;
  *lni_ti = dln; *lis_ti = dls; *lno_ti = lno; *los_ti = los;
  if (transp & PFFT_TRANSPOSED_OUT) {
    *lno_to = lno; *los_to = los;
    *lni_ti = dln; *lis_ti = dls; *lno_ti = dln; *los_ti = dls;
  }
  if (transp & PFFT_TRANSPOSED_IN) {
    *lni_to = dln; *lis_to = dls; *lno_to = dln; *los_to = dls;
    *lni_ti = lni; *lis_ti = lis;
  }
}

static void shift_starts(int rnk_n, const INT *ni, const INT *no, unsigned pfft_flags, INT *lis,
                         INT *los) {
  /* kernel/partrafo.c:287-299 (the reference prints an error for odd
   * sizes and carries on) */
  if (pfft_flags & PFFT_SHIFTED_IN)
    for (int t = 0; t < rnk_n; t++) {
      if (ni[t] % 2)
        fprintf(stderr, "PFFT ERROR: flag PFFT_SHIFTED_IN requires ni to be an even number\n");
      lis[t] -= ni[t] / 2;
    }
  if (pfft_flags & PFFT_SHIFTED_OUT)
    for (int t = 0; t < rnk_n; t++) {
      if (no[t] % 2)
        fprintf(stderr, "PFFT ERROR: flag PFFT_SHIFTED_OUT requires no to be an even number\n");
      los[t] -= no[t] / 2;
    }
}

static int mesh_rank_for(int rnk_n, const pa_mesh *m) {
  /* util/util.c:253-260: a 3-d FFT on a 3-d mesh is remapped to 2-d */
  if (rnk_n == 3 && m->rnk_pm == 3) return 2;
  return m->rnk_pm;
}

INT pa_local_size(int rnk_n, const INT *n, const INT *ni, const INT *no, INT howmany,
                  const INT *iblock, const INT *oblock, const pa_mesh *m, unsigned tf,
                  unsigned pfft_flags, INT *lni, INT *lis, INT *lno, INT *los) {
  const unsigned transp = transp_of(pfft_flags);
  const int r = mesh_rank_for(rnk_n, m);
  INT iblk[PA_MAXRNK], mblk[PA_MAXRNK], oblk[PA_MAXRNK];
  INT n_to[PA_MAXRNK], ni_to[PA_MAXRNK], no_to[PA_MAXRNK];
  INT n_ti[PA_MAXRNK], ni_ti[PA_MAXRNK], no_ti[PA_MAXRNK];
  unsigned tfs_to[PA_MAXRNK + 1], tfs_ti[PA_MAXRNK + 1];
  INT dln[PA_MAXRNK], dls[PA_MAXRNK];
  INT *lni_to, *lis_to, *lno_to, *los_to, *lni_ti, *lis_ti, *lno_ti, *los_ti;
  INT mem = 1;

  if (rnk_n < 1 || rnk_n > PA_MAXRNK || rnk_n < r || (rnk_n == r && rnk_n != 3) ||
      transp == (PFFT_TRANSPOSED_IN | PFFT_TRANSPOSED_OUT)) {
    /* combinations the reference cannot plan (kernel/partrafo.c:348-373);
     * its local_size would read past the size arrays here */
    pa_error("local_size: unsupported rank/mesh/flag combination");
    for (int t = 0; t < rnk_n && t < PA_MAXRNK; t++) lni[t] = lis[t] = lno[t] = los[t] = 0;
    return 0;
  }
  eval_blocks(rnk_n, ni, no, iblock, oblock, r, m->np, tf, transp, iblk, mblk, oblk);
  pa_split_steps(rnk_n, n, ni, no, r, tf, transp, NULL, n_to, ni_to, no_to, tfs_to, n_ti, ni_ti,
                 no_ti, tfs_ti);
  alias_local(lni, lis, dln, dls, lno, los, transp, &lni_to, &lis_to, &lno_to, &los_to, &lni_ti,
              &lis_ti, &lno_ti, &los_ti);
  if (!(transp & PFFT_TRANSPOSED_IN))
    mem = PA_MAX(mem, step_local_size(rnk_n, n_to, ni_to, no_to, howmany, iblk, mblk, r,
                                      PFFT_TRANSPOSED_OUT, tfs_to, lni_to, lis_to, lno_to,
                                      los_to));
  if (!(transp & PFFT_TRANSPOSED_OUT))
    mem = PA_MAX(mem, step_local_size(rnk_n, n_ti, ni_ti, no_ti, howmany, mblk, oblk, r,
                                      PFFT_TRANSPOSED_IN, tfs_ti, lni_ti, lis_ti, lno_ti,
                                      los_ti));
  shift_starts(rnk_n, ni, no, pfft_flags, lis, los);
  return mem;
}

/* kernel/partrafo.c:99-199 */
void pa_local_block(int rnk_n, const INT *ni, const INT *no, const INT *iblock,
                    const INT *oblock, const pa_mesh *m, int pid, unsigned tf,
                    unsigned pfft_flags, INT *lni, INT *lis, INT *lno, INT *los) {
  const unsigned transp = transp_of(pfft_flags);
  const int r = mesh_rank_for(rnk_n, m);
  INT iblk[PA_MAXRNK], mblk[PA_MAXRNK], oblk[PA_MAXRNK];
  INT n_to[PA_MAXRNK], ni_to[PA_MAXRNK], no_to[PA_MAXRNK];
  INT n_ti[PA_MAXRNK], ni_ti[PA_MAXRNK], no_ti[PA_MAXRNK];
  unsigned tfs_to[PA_MAXRNK + 1], tfs_ti[PA_MAXRNK + 1];
  INT dln[PA_MAXRNK], dls[PA_MAXRNK];
  INT *lni_to, *lis_to, *lno_to, *los_to, *lni_ti, *lis_ti, *lno_ti, *los_ti;

  if (rnk_n < 1 || rnk_n > PA_MAXRNK || rnk_n < r || (rnk_n == r && rnk_n != 3) ||
      transp == (PFFT_TRANSPOSED_IN | PFFT_TRANSPOSED_OUT)) {
    pa_error("local_block: unsupported rank/mesh/flag combination");
    for (int t = 0; t < rnk_n && t < PA_MAXRNK; t++) lni[t] = lis[t] = lno[t] = los[t] = 0;
    return;
  }
  if (pid != 0) {
    /* a one-process mesh has only rank 0 */
    for (int t = 0; t < rnk_n; t++) lni[t] = lis[t] = lno[t] = los[t] = 0;
    fprintf(stderr, "PFFT ERROR: local_block: pid %d does not exist on a one-process mesh\n", pid);
    return;
  }
  eval_blocks(rnk_n, ni, no, iblock, oblock, r, m->np, tf, transp, iblk, mblk, oblk);
  /* local_block passes ni as both n and ni (partrafo.c:143-146) */
  pa_split_steps(rnk_n, ni, ni, no, r, tf, transp, NULL, n_to, ni_to, no_to, tfs_to, n_ti, ni_ti,
                 no_ti, tfs_ti);
  alias_local(lni, lis, dln, dls, lno, los, transp, &lni_to, &lis_to, &lno_to, &los_to, &lni_ti,
              &lis_ti, &lno_ti, &los_ti);
  if (!(transp & PFFT_TRANSPOSED_IN)) {
    local_size_transposed(rnk_n, ni_to, no_to, iblk, mblk, r, tfs_to[r], PFFT_TRANSPOSED_OUT,
                          lni_to, lis_to, lno_to, los_to);
    real_count_overwrite(rnk_n, ni_to, no_to, tfs_to[r], lni_to, lno_to);
  }
  if (!(transp & PFFT_TRANSPOSED_OUT)) {
    local_size_transposed(rnk_n, ni_ti, no_ti, mblk, oblk, r, tfs_ti[r], PFFT_TRANSPOSED_IN,
                          lni_ti, lis_ti, lno_ti, los_ti);
    real_count_overwrite(rnk_n, ni_ti, no_ti, tfs_ti[r], lni_ti, lno_ti);
  }
  shift_starts(rnk_n, ni, no, pfft_flags, lis, los);
}
