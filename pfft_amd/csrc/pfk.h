/*
 * pfk.h -- thin internal C ABI between the C host library (planner,
 * executor, API) and the hand-written HIP kernels for gfx950.
 *
 * Everything here is plain C: plain pointers, sizes and a POD pass
 * descriptor.  No HIP or torch types cross this boundary, so the host
 * side (the .c files of pfft_amd/csrc) is compiled by gcc and only the .hip files
 * by hipcc.
 *
 * A "pass" transforms one logical dimension of a strided multi-dimensional
 * array: for every line along that dimension it
 *   1. loads nin elements (stride sin), multiplies element j by the
 *      input modulation factor (+-1), and embeds them at FFT index zin+j
 *      of a zero line of length n  (the reference's ousam embed,
 *      kernel/ousample.c:391-456, and twiddle_input, api/api-basic.c:1186);
 *   2. computes the unnormalised length-n DFT with the given sign
 *      (FFTW convention: sign -1 forward; kernel/sertrafo.c:489-554);
 *      kind R2C / C2R are FFTW's half-spectrum real transforms;
 *   3. keeps nout outputs starting at FFT index zout (the reference's
 *      ousam trunc, kernel/ousample.c:461-506), multiplies output k' by the
 *      output modulation factor (api/api-basic.c:1242) and stores them with
 *      stride sout.
 * The lines are indexed by up to PFK_MAXB batch dimensions plus one "tile"
 * dimension whose consecutive values are processed together by one
 * workgroup so that global loads/stores coalesce along it.
 */
#ifndef PFK_H
#define PFK_H 1

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* helpers below are shared by the C host code and the HIP kernels */
#ifdef __HIPCC__
#define PFK_HD __host__ __device__
#else
#define PFK_HD
#endif

#define PFK_MAXB 6
#define PFK_MAXSTAGES 24

enum pfk_kind { PFK_C2C = 0, PFK_R2C = 1, PFK_C2R = 2, PFK_COPY = 3, PFK_R2R = 4 };

/* Real-to-real kinds of a PFK_R2R pass (the FFTW kind numbers of
 * include/pfft.h).  A length-n r2r line is computed as a forward complex
 * DFT of length nf of a real sequence u built from the line:
 *   DHT      nf = n      u_p = x_p                         Y_k = Re F_k - Im F_k
 *   REDFT00  nf = 2(n-1) u_p = u_{nf-p} = x_p              Y_k = Re F_k
 *   REDFT10  nf = 4n     u_{2p+1} = u_{nf-2p-1} = x_p      Y_k = Re F_k
 *   REDFT01  nf = 4n     u_p = u_{nf-p} = x_p              Y_k = Re F_{2k+1}
 *   REDFT11  nf = 8n     u_{2p+1} = u_{nf-2p-1} = x_p      Y_k = Re F_{2k+1}
 *   RODFT00  nf = 2(n+1) u_{p+1} = -u_{nf-p-1} = x_p       Y_k = -Im F_{k+1}
 *   RODFT10  nf = 4n     u_{2p+1} = -u_{nf-2p-1} = x_p     Y_k = -Im F_{k+1}
 *   RODFT01  nf = 4n     u_{p+1} = -u_{nf-p-1} = x_p (p < n-1),
 *                        u_n = x_{n-1} (unmirrored)        Y_k = -Im F_{2k+1}
 *   RODFT11  nf = 8n     u_{2p+1} = -u_{nf-2p-1} = x_p     Y_k = -Im F_{2k+1}
 * (all other u entries zero).  Each identity is the kind's defining sum
 * with the cosine/sine written as the sum of two exponentials. */
enum pfk_r2r { PFK_DHT = 2, PFK_REDFT00 = 3, PFK_REDFT01 = 4, PFK_REDFT10 = 5,
               PFK_REDFT11 = 6, PFK_RODFT00 = 7, PFK_RODFT01 = 8, PFK_RODFT10 = 9,
               PFK_RODFT11 = 10 };
enum pfk_prec { PFK_F64 = 0, PFK_F32 = 1 };

/* +-1 modulation of a line index:
 *   f(j) = (on && j+start < half) ? ((j+start) odd ? -1 : 1) * extra : 1  */
typedef struct pfk_mod {
  int32_t on;
  int32_t extra;      /* +1 or -1 */
  int64_t start;
  int64_t half;
} pfk_mod;

typedef struct pfk_pass {
  int32_t kind;       /* enum pfk_kind */
  int32_t prec;       /* enum pfk_prec */
  int32_t sign;       /* -1 or +1 */
  int32_t nstages;    /* radix factorisation of the computed FFT length */
  int32_t radix[PFK_MAXSTAGES];
  int64_t n;          /* logical transform length */
  int64_t nf;         /* computed complex FFT length (n, n/2 for even r2c/c2r, the
                         extended length for r2r) */
  int64_t nin, zin;   /* input count along the line and embed offset */
  int64_t nout, zout; /* kept output count and first kept index */
  int64_t sin, sout;  /* line strides, in elements of the src/dst type */
  pfk_mod imod, omod;
  int32_t nb;         /* number of non-tile batch dims */
  int32_t hw;         /* Hermitian weights: 0 none; 1 c2r input x1/2 off DC/Nyquist;
                         2 r2c output x2 off DC/Nyquist (adjoint passes) */
  int64_t bext[PFK_MAXB], bsin[PFK_MAXB], bsout[PFK_MAXB];
  int64_t text, tsin, tsout;   /* tile dimension (text==1: none) */
  int32_t xcd_remap;  /* map consecutive tiles to one XCD (see kernels_pow2.hip) */
  int32_t mfma;       /* complex64 c2c arithmetic form: 0 plain, 1 FP32 MFMA, 2 split-BF16 MFMA */
  int32_t variant;    /* launch variant: 0 default, 1 general kernel, 2 narrow power-of-two
                         tiles, 3 short-length register kernel; + 4 (with 0 or 2): one
                         workgroup per tile instead of the persistent grid; + 8: line-major
                         lane order (complex64 c2c) */
  int32_t r2r;        /* enum pfk_r2r of a PFK_R2R pass */
  /* Long lines (nf > PFK_MAXLINE): nf = n1 * n2 is computed in two
   * steps through `work` (lines * nf complex elements, owned by the plan):
   * length-n1 DFTs over the decimated lines j2, j2 + n2, ... times
   * W_nf^(j2*k1), then length-n2 DFTs (Cooley-Tukey four-step).  The twiddle
   * table holds W_n1 (n1), W_n2 (n2), W_nf^l (l < 4096) and W_nf^(4096 h).
   * Real lines (r2c / c2r, n even, nf = n/2; contiguous and unembedded on
   * the real side) pass through `work2` as the half-length complex line
   * z_j = x_2j + i x_2j+1 (r2c: transform then combine the half spectrum;
   * c2r: combine then transform). */
  int64_t n1;         /* 0: the line fits one workgroup */
  void *work;
  void *work2;        /* real lines (r2c / c2r, nf = n/2): the half-length complex line */
  /* Fields of the two steps (set by the long-line launcher; 0 otherwise):
   * dec_in:  element j of line c is index j*dec_in + c of the original line,
   *          whose embed (nin, zin) and input modulation apply;
   * dec_out: output k of line c is index k*dec_out + c of the original
   *          output line, whose trunc (nout, zout) and modulation apply;
   * twn:     output k of line c is multiplied by W_twn^(c*k) (table tw2). */
  int64_t dec_in, dec_out, twn;
  const void *tw2;
  /* r2r passes: 0 = transform of the extension (table above); 1, 2 =
   * inner FFT of about the line's own length between elementwise pre and
   * post steps (pfk_r2r_inner), through `work` and `work2` */
  int32_t r2r_path;
  int32_t pad3;
  /* device twiddle tables (owned by the host plan):
   *   tw  : W_nf^m = exp(sign*2*pi*i*m/nf), m in [0,nf)
   *   rtw : W_n^k  = exp(sign*2*pi*i*k/n),  k in [0,n/2] (even r2c/c2r only) */
  const void *tw;
  const void *rtw;
  /* element counts of the src/dst buffers, for host-side bounds checks */
  int64_t src_elems, dst_elems;
} pfk_pass;

/* Longest computed line one workgroup transforms. */
#define PFK_MAXLINE 4096

/* Radix factorisation of a computed length: 4s first, then 2, then odd
 * primes; the number of stages, or -1. */
static inline int pfk_factorize(int64_t nf, int32_t *radix) {
  int ns = 0;
  int64_t m = nf;
  while (m % 4 == 0 && ns < PFK_MAXSTAGES) { radix[ns++] = 4; m /= 4; }
  while (m % 2 == 0 && ns < PFK_MAXSTAGES) { radix[ns++] = 2; m /= 2; }
  for (int64_t q = 3; m > 1 && ns < PFK_MAXSTAGES; q += 2)
    while (m % q == 0 && ns < PFK_MAXSTAGES) { radix[ns++] = (int32_t)q; m /= q; }
  return m == 1 ? ns : -1;
}

/* Split of a long computed length nf = n1 * n2 with both factors at most
 * PFK_MAXLINE: the most balanced one (smallest max/min) among those with a
 * power-of-two n2 >= 64 (the second step then runs on the power-of-two
 * register kernels), else among all; ties go to the larger n2.  Returns
 * n1, or 0 if there is no split. */
static inline int64_t pfk_long_split(int64_t nf) {
  int64_t best[2] = {0, 0};
  double ratio[2] = {0, 0};
  for (int64_t n2 = PFK_MAXLINE; n2 >= 2; n2--) {
    const int64_t n1 = nf / n2;
    if (nf % n2 || n1 > PFK_MAXLINE || n1 < 2) continue;
    const double r = n1 > n2 ? (double)n1 / (double)n2 : (double)n2 / (double)n1;
    const int c = (n2 >= 64 && (n2 & (n2 - 1)) == 0) ? 0 : 1;
    if (!best[c] || r < ratio[c]) { best[c] = n1; ratio[c] = r; }
  }
  return best[0] ? best[0] : best[1];
}

/* One task of a batched execution: one pass of one plan, with resolved
 * device pointers.  `dep` is the index of the task that must have finished
 * all its tiles before any tile of this task starts (-1: none). */
typedef struct pfk_task {
  pfk_pass pass;
  const void *src;
  void *dst;
  int32_t C, pitch;       /* filled by pfk_batch_create (C < 0: short-length kernel) */
  int64_t ntiles;         /* filled by pfk_batch_create */
  int64_t first_item;     /* filled by pfk_batch_create */
  int32_t dep;
  int32_t pad;
} pfk_task;

/* Batched execution of many passes: tasks grouped by pass level, one
 * persistent launch per level and kernel class walking a device work-list,
 * all levels optionally captured into one hipGraph and replayed.  All tasks
 * must share one precision. */
typedef struct pfk_batch pfk_batch;
pfk_batch *pfk_batch_create(pfk_task *tasks, int ntasks, int use_graph);
int  pfk_batch_run(pfk_batch *b);
void pfk_batch_destroy(pfk_batch *b);
/* The same launches without any allocation: the device work-list goes
 * into a fixed device array of the library (at most 2048 passes; tasks is
 * host memory, in level order as for pfk_batch_create, and is reordered).
 * Direct launches, no graph.  Blocking; calls that overlap on one GPU
 * share the array and race. */
int  pfk_batch_run_noalloc(pfk_task *tasks, int ntasks);

/* Ghost cells of a 3-d local block on one process (sizes in the plan's
 * memory order, tuple reals innermost).  The block n[0] x n[1] x n[2] is
 * periodic with period n[t] along dimension t; the box extends it by
 * gb[t] cells below and ngc[t] - gb[t] - n[t] above.
 *   exchange: the block at the front of `data` becomes the box, every
 *             ghost cell holding its periodic image;
 *   reduce:   every block element receives the sum of its images in the
 *             box `data`, the block is stored at the front of `data` and
 *             the rest of the box is zeroed.
 * `scratch` holds tuple*n0*n1*n2 reals.  `ev` (may be NULL) holds three
 * events recorded before, between and after the two phases. */
typedef struct pfk_gc {
  int32_t prec;
  int32_t pad;
  int64_t n[3], gb[3], ngc[3];
  int64_t tuple;
} pfk_gc;
int pfk_gc_exchange(const pfk_gc *g, void *data, void *scratch, void *const *ev);
int pfk_gc_reduce(const pfk_gc *g, void *data, void *scratch, void *const *ev);

/* Device/runtime plumbing.  All return 0 on success, nonzero on error. */
int  pfk_init(void);
int  pfk_device_count(void);
int  pfk_malloc(void **p, size_t bytes);
int  pfk_malloc_host(void **p, size_t bytes);
int  pfk_free(void *p);
int  pfk_free_host(void *p);
/* 1 if p is device (or managed) memory the GPU can address directly */
int  pfk_is_device_ptr(const void *p);
int  pfk_memcpy_h2d(void *dst, const void *src, size_t bytes);
int  pfk_memcpy_d2h(void *dst, const void *src, size_t bytes);
int  pfk_memcpy_d2d(void *dst, const void *src, size_t bytes);
int  pfk_memset(void *dst, int v, size_t bytes);
int  pfk_sync(void);
const char *pfk_last_error(void);

/* Launch one pass on the library stream (asynchronous). */
int  pfk_launch_pass(const pfk_pass *p, const void *src, void *dst);
/* r2r kinds computed through an inner FFT of about the line's own length
 * (r2r_path 1 and 2), between elementwise pre and post steps:
 *   mode 0  DHT, REDFT10, RODFT10, REDFT00, RODFT00: the real FFT of a real
 *           line of length L = n, n, n, 2(n-1), 2(n+1) (Makhoul's
 *           permutation for the 10 kinds, the short symmetric extension for
 *           the 00 kinds);
 *   mode 1  REDFT01, RODFT01: the inverse real FFT of length n (inverse
 *           Makhoul);
 *   mode 2  REDFT11, RODFT11: a complex FFT of length n/2 with pre and post
 *           twiddles (n even).
 * Path 1 (even L, even n): the inner transform is always the complex FFT
 * of length M = L/2 (or n/2); the real packing z_j = v_2j + i v_2j+1 and
 * the half-spectrum combine are done in the pre/post steps.  Each work
 * array is stored line-major when the pass's lines are contiguous on that
 * side, else with the lines innermost, and the inner pass transposes
 * between the two.  Path 2 (odd lengths, modes 0 and 1): the inner r2c /
 * c2r of length L or n, line-major.
 * Tables: tw = the inner complex FFT; rtw = W_L^m (mode 0) or W_n^m
 * (mode 1), m <= M, path 1 only; tw2 = W_4n^k, k <= n (10/01 kinds) or
 * W_8n^m, m < 2n (11 kinds). */
PFK_HD static inline int pfk_r2r_mode(int kind) {
  return (kind == 4 || kind == 8) ? 1 : (kind == 6 || kind == 10) ? 2 : 0;   /* 01, 11 kinds */
}
PFK_HD static inline int64_t pfk_r2r_inner_len(int kind, int64_t n) {
  return kind == 3 ? 2 * (n - 1) : kind == 7 ? 2 * (n + 1) : n;   /* REDFT00, RODFT00 */
}
/* The inner pass of an r2r pass, its path (1 or 2) and the work arrays'
 * line pitches in reals; nonzero if the kind / length has no such path. */
int  pfk_r2r_inner(const pfk_pass *p, pfk_pass *inner, int64_t *ld1, int64_t *ld2, int *path);

/* 0 if a kernel can run the (one-workgroup-per-line) pass: one of the
 * specialised families covers it or the general kernel's line fits the
 * device's LDS; else nonzero with the reason in pfk_last_error(). */
int  pfk_pass_check(const pfk_pass *p);
/* Kernel families with a kernel for the pass shape (bit mask):
 * PFK_HAS_POW2 register kernels for power-of-two lengths 64..4096,
 * PFK_HAS_SMALL one-line-per-lane kernels for short c2c lengths,
 * PFK_HAS_MFMA MFMA kernels (complex64 arithmetic forms).  The general
 * kernel covers every pass.  With PFFT_AMD_GENERIC_ONLY=1 the mask is 0. */
#define PFK_HAS_POW2  1
#define PFK_HAS_SMALL 2
#define PFK_HAS_MFMA  4
int  pfk_pass_kernels(const pfk_pass *p);
/* 1 if the pass is plain (every input and output position present, no
 * modulation or weights): the instantiations without masks run it */
int  pfk_pass_plain(const pfk_pass *p);

/* Timing helpers on the library stream (HIP events). */
int  pfk_event_create(void **ev);
int  pfk_event_record(void *ev);
int  pfk_event_elapsed_ms(void *ev0, void *ev1, float *ms);
int  pfk_event_destroy(void *ev);

/* Complex conjugation / scaling utility kernels on device arrays. */
int  pfk_scale(void *data, int prec, int64_t nreal, double factor);

/* Largest grid (workgroups) of one launch of `threads`-thread workgroups:
 * the work-item count per grid dimension must fit in 32 bits, so passes
 * with more tiles are launched in chunks of this many workgroups. */
int64_t pfk_max_grid(int threads);

/* 1 if an even r2c / c2r pass of half length nf (f64 nonzero: complex
 * double) has the short-length split kernel by length (the dispatcher's
 * rule, before the device's LDS limit). */
int pfk_small_real_length(int nf, int f64);

/* Streaming bandwidth probe (GB/s of bytes read plus written), best of
 * `iters` launches: mode 0 read, 1 write, 2 copy; width 16/32/64 bytes per
 * lane and access; nt nontemporal; grid workgroups (0: one pass). */
double pfk_bw_probe(int mode, int width, int nt, int64_t grid, size_t bytes, int iters);

#ifdef __cplusplus
}
#endif
#endif
