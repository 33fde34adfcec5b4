// kernels_generic.hip -- runtime plumbing and the general pass kernel.
//
// The general pass kernel handles every pass the planner can produce:
// any transform length whose computed FFT fits in one workgroup's LDS,
// any radix factorisation (output-centric Stockham stages, so radices
// 2..nf are all the same code), arbitrary line/batch strides, embed
// (zero-pad) and trunc offsets, +-1 index modulations, and the r2c/c2r
// half-length packing.  The fast path for the large power-of-two sizes
// lives in kernels_pow2.hip; this kernel is what keeps every size and
// flag combination correct.
//
// Layout of a workgroup's work: one tile = C lines that are consecutive
// along the tile dimension.  The tile is staged in LDS as C rows of
// `pitch` complex slots; global loads/stores map consecutive lanes to
// consecutive addresses (along the tile dimension when it has unit
// stride, else along the line).
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "pfk.h"
#include "pfk_internal.hpp"
#include "pass_tile.hpp"

namespace pfk {

hipStream_t g_stream = nullptr;
int g_generic_only = 0;
int g_max_lds = 65536;
static char g_err[512] = "";

void set_error(const char *what, hipError_t e) {
  snprintf(g_err, sizeof(g_err), "%s: %s", what, hipGetErrorString(e));
}
void set_error_msg(const char *msg) { snprintf(g_err, sizeof(g_err), "%s", msg); }

template <typename T, int KIND, int EPT>
__global__ void __launch_bounds__(256)
k_pass_generic(const pfk_pass P, const int C, const int pitch, const int64_t tile_base,
               const T *__restrict__ src, T *__restrict__ dst) {
  extern __shared__ unsigned char smem_raw[];
  pass_tile<T, KIND, EPT>(P, C, pitch, tile_base + blockIdx.x, src, dst, smem_raw);
}

// number of complex LDS slots one line needs
static int64_t slots_per_line(const pfk_pass *p) {
  if (p->kind == PFK_R2R) return p->nf;   // the extended line
  if (p->kind == PFK_R2C) return (p->n % 2 == 0) ? p->n / 2 : p->n;
  if (p->kind == PFK_C2R) return (p->n % 2 == 0) ? p->n / 2 + 1 : p->n;
  return p->n;
}

int generic_tiling(const pfk_pass *p, int *Cout, int *pitch_out, size_t *lds_out) {
  constexpr int NT = 256, EPT = 16;
  const int64_t ns = slots_per_line(p);
  const int64_t nf = (p->kind == PFK_COPY) ? 1 : p->nf;
  const size_t esz = (p->prec == PFK_F64) ? 16 : 8;
  int C = 64;
  while (C > 1 && (C * nf > (int64_t)NT * EPT || C * (ns + 1) * (int64_t)esz > g_max_lds ||
                   C > 2 * p->text))
    C >>= 1;
  int64_t pitch = (C > 1) ? ns + 1 : ns;
  if (C * nf > (int64_t)NT * EPT || C * pitch * (int64_t)esz > g_max_lds) {
    set_error_msg("pfk: transform length too large for the general LDS kernel");
    return 1;
  }
  *Cout = C;
  *pitch_out = (int)pitch;
  *lds_out = (size_t)(C * pitch) * esz;
  return 0;
}

template <typename T>
static int launch_generic_t(const pfk_pass *p, const void *src, void *dst) {
  constexpr int NT = 256, EPT = 16;
  int C, pitch;
  size_t lds;
  if (generic_tiling(p, &C, &pitch, &lds)) return 1;
  int64_t ntiles = (p->text + C - 1) / C;
  for (int i = 0; i < p->nb; ++i) ntiles *= p->bext[i];
  if (ntiles <= 0) return 0;
  const T *s = (const T *)src;
  T *d = (T *)dst;
  const int64_t chunk = pfk_max_grid(NT);
  for (int64_t base = 0; base < ntiles; base += chunk) {
    const int64_t nblk = (ntiles - base) < chunk ? (ntiles - base) : chunk;
    dim3 grid((unsigned)nblk), block(NT);
    switch (p->kind) {
      case PFK_C2C:
        hipLaunchKernelGGL((k_pass_generic<T, PFK_C2C, EPT>), grid, block, lds, g_stream, *p, C,
                           (int)pitch, base, s, d);
        break;
      case PFK_R2C:
        hipLaunchKernelGGL((k_pass_generic<T, PFK_R2C, EPT>), grid, block, lds, g_stream, *p, C,
                           (int)pitch, base, s, d);
        break;
      case PFK_C2R:
        hipLaunchKernelGGL((k_pass_generic<T, PFK_C2R, EPT>), grid, block, lds, g_stream, *p, C,
                           (int)pitch, base, s, d);
        break;
      case PFK_R2R:
        hipLaunchKernelGGL((k_pass_generic<T, PFK_R2R, EPT>), grid, block, lds, g_stream, *p, C,
                           (int)pitch, base, s, d);
        break;
      default:
        hipLaunchKernelGGL((k_pass_generic<T, PFK_COPY, EPT>), grid, block, lds, g_stream, *p, C,
                           (int)pitch, base, s, d);
        break;
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) { set_error("pfk_launch_pass(generic)", e); return 1; }
  }
  return 0;
}

int launch_generic(const pfk_pass *p, const void *src, void *dst) {
  return (p->prec == PFK_F64) ? launch_generic_t<double>(p, src, dst)
                              : launch_generic_t<float>(p, src, dst);
}

// A long c2c line (nf = n1 * n2 > PFK_MAXLINE) in two steps through the
// plan's work buffer W, laid out [line][k1][j2] (line = tile index, then the
// batch indices, innermost first):
//   step A: for every line and j2, the length-n1 DFT of the decimated line
//           x[j2], x[j2 + n2], ... (embed and input modulation applied to
//           the original index), times W_nf^(j2*k1), stored at W[line][k1][j2];
//   step B: for every line and k1, the length-n2 DFT over j2, stored at
//           original output index k1 + n1*k2 (trunc and output modulation).
// Both steps take j2 resp. k1 as the tile dimension (consecutive lines are
// adjacent in memory when the line stride is 1).  Step A runs on the
// power-of-two register kernels (twiddled-store instantiation) unless the
// line is embedded or modulated; step B on any kernel unless it truncates
// or modulates.
static int long_c2c(const pfk_pass *p, const void *src, void *dst) {
  const int64_t N = p->nf, N1 = p->n1, N2 = N / N1;
  if (p->kind != PFK_C2C || N1 < 2 || N1 * N2 != N || N1 > PFK_MAXLINE || N2 > PFK_MAXLINE ||
      !p->work) {
    set_error_msg("pfk: bad long-line pass");
    return 1;
  }
  const int nbo = p->nb + (p->text > 1 ? 1 : 0);   // original line dims
  if (nbo > PFK_MAXB) { set_error_msg("pfk: too many batch dimensions for a long line"); return 1; }
  int64_t ext[PFK_MAXB + 1], osin[PFK_MAXB + 1], osout[PFK_MAXB + 1];
  int m = 0;
  if (p->text > 1) { ext[m] = p->text; osin[m] = p->tsin; osout[m] = p->tsout; m++; }
  for (int i = 0; i < p->nb; i++) {
    ext[m] = p->bext[i]; osin[m] = p->bsin[i]; osout[m] = p->bsout[i]; m++;
  }
  int64_t lines = 1;
  for (int i = 0; i < m; i++) lines *= ext[i];
  const size_t ce = (p->prec == PFK_F64) ? 16 : 8;
  const char *tw = (const char *)p->tw;

  pfk_pass a = *p;
  a.n1 = 0; a.work = NULL; a.hw = 0; a.mfma = 0; a.dec_out = 0;
  a.n = a.nf = N1;
  a.nstages = pfk_factorize(N1, a.radix);
  a.tw = tw;
  a.twn = N;
  a.tw2 = tw + ce * (size_t)(N1 + N2);
  // embed / modulation of the original line: decimated loads (general
  // kernel); otherwise plain strided lines (any kernel with the twiddle)
  const int embed = p->nin != N || p->zin != 0 || p->imod.on;
  a.dec_in = embed ? N2 : 0;
  a.variant = embed ? 1 : 0;
  if (!embed) { a.nin = N1; a.zin = 0; }
  a.sin = N2 * p->sin;
  a.nout = N1; a.zout = 0; a.omod.on = 0;
  a.sout = N2;
  a.text = N2; a.tsin = p->sin; a.tsout = 1;
  a.nb = m;
  int64_t wst = N;
  for (int i = 0; i < m; i++) {
    a.bext[i] = ext[i]; a.bsin[i] = osin[i]; a.bsout[i] = wst;
    wst *= ext[i];
  }
  a.dst_elems = lines * N;

  pfk_pass b = *p;
  b.n1 = 0; b.work = NULL; b.hw = 0; b.dec_in = 0; b.twn = 0; b.tw2 = NULL;
  b.n = b.nf = N2;
  b.nstages = pfk_factorize(N2, b.radix);
  b.tw = tw + ce * (size_t)N1;
  b.nin = N2; b.zin = 0; b.imod.on = 0;
  b.sin = 1;
  const int trunc = p->nout != N || p->zout != 0 || p->omod.on;
  b.dec_out = trunc ? N1 : 0;           // trunc / modulation of the original line
  if (!trunc) { b.nout = N2; b.zout = 0; }
  b.sout = N1 * p->sout;
  b.text = N1; b.tsin = N2; b.tsout = p->sout;
  b.nb = m;
  wst = N;
  for (int i = 0; i < m; i++) {
    b.bext[i] = ext[i]; b.bsin[i] = wst; b.bsout[i] = osout[i];
    wst *= ext[i];
  }
  b.src_elems = lines * N;
  if (a.nstages < 0 || b.nstages < 0) { set_error_msg("pfk: long-line factors"); return 1; }
  if (pfk_launch_pass(&a, src, p->work)) return 1;
  return pfk_launch_pass(&b, p->work, dst);
}

// ---- long real lines (n even, M = n/2 > PFK_MAXLINE) ----
// Line index l (tile index innermost, then the batch indices) -> offset.
// (32-bit divisions when the line index fits: several times cheaper)
__device__ __forceinline__ int64_t line_offset(const pfk_pass &P, int64_t l, bool out) {
  if (l < INT32_MAX && P.text < INT32_MAX) {
    uint32_t u = (uint32_t)l;
    const uint32_t t = u % (uint32_t)P.text;
    u /= (uint32_t)P.text;
    int64_t o = (int64_t)t * (out ? P.tsout : P.tsin);
    for (int i = 0; i < P.nb; ++i) {
      const uint32_t e = (uint32_t)P.bext[i];   // <= the line count
      const uint32_t d = u % e;
      u /= e;
      o += (int64_t)d * (out ? P.bsout[i] : P.bsin[i]);
    }
    return o;
  }
  const int64_t t = l % P.text;
  l /= P.text;
  int64_t o = t * (out ? P.tsout : P.tsin);
  for (int i = 0; i < P.nb; ++i) {
    const int64_t d = l % P.bext[i];
    l /= P.bext[i];
    o += d * (out ? P.bsout[i] : P.bsin[i]);
  }
  return o;
}

// g -> (fast index, slow index) for g < fast_n * slow_n, 32-bit when it fits
__device__ __forceinline__ void split_index(int64_t g, int64_t fast_n, int64_t *f, int64_t *sl) {
  if (g < INT32_MAX && fast_n < INT32_MAX) {
    const uint32_t q = (uint32_t)g / (uint32_t)fast_n;
    *sl = q;
    *f = (uint32_t)g - q * (uint32_t)fast_n;
  } else {
    *sl = g / fast_n;
    *f = g - *sl * fast_n;
  }
}

// r2c: Z = DFT_M(z), z_j = x_2j + i x_2j+1, lines of M in `Z`.  Output k
// (kept range [zout, zout+nout) of 0..M) is X_k = E_k + W_n^k O_k with
// E_k = (Z_k + conj Z_(M-k))/2, O_k = (Z_k - conj Z_(M-k))/(2i); times 2
// off DC/Nyquist for the adjoint weights (hw 2) and the output modulation.
template <typename T>
__global__ void __launch_bounds__(256)
k_r2c_long_post(const pfk_pass P, const int64_t total, const cpx<T> *__restrict__ Z,
                cpx<T> *__restrict__ dst) {
  const int64_t g = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (g >= total) return;
  const int64_t M = P.nf, k = g % P.nout, l = g / P.nout;
  const int64_t kk = P.zout + k;
  const cpx<T> *z = Z + l * M;
  const cpx<T> a = z[kk % M], b0 = z[(M - kk) % M];
  const cpx<T> b{b0.x, -b0.y};
  const cpx<T> e{(a.x + b.x) * (T)0.5, (a.y + b.y) * (T)0.5};
  const cpx<T> o{(a.y - b.y) * (T)0.5, -(a.x - b.x) * (T)0.5};   // (a - b) / (2i)
  const cpx<T> w = reinterpret_cast<const cpx<T> *>(P.rtw)[kk];
  cpx<T> x = cmul(w, o);
  x.x += e.x;
  x.y += e.y;
  T f = (T)mod_sign(P.omod, k);
  if (P.hw == 2 && kk > 0 && kk < M) f *= (T)2;
  dst[line_offset(P, l, true) + k * P.sout] = cpx<T>{x.x * f, x.y * f};
}

// c2r: Z'_k = (X_k + conj X_(M-k)) + i W_n^k (X_k - conj X_(M-k)), k < M,
// from the half spectrum (kept input range [zin, zin+nin) of 0..M, input
// modulation, x1/2 off DC/Nyquist for the adjoint weights (hw 1), imaginary
// parts of DC and Nyquist ignored); then x_2j + i x_2j+1 = DFT_M(Z')_j.
template <typename T>
__device__ __forceinline__ cpx<T> c2r_in(const pfk_pass &P, const cpx<T> *s, int64_t k) {
  const int64_t j = k - P.zin;
  if (j < 0 || j >= P.nin) return cpx<T>{0, 0};
  cpx<T> v = s[j * P.sin];
  T f = (T)mod_sign(P.imod, j);
  if (P.hw == 1 && k > 0 && k < P.nf) f *= (T)0.5;
  if (k == 0 || k == P.nf) v.y = 0;
  return cpx<T>{v.x * f, v.y * f};
}

template <typename T>
__global__ void __launch_bounds__(256)
k_c2r_long_pre(const pfk_pass P, const int64_t total, const cpx<T> *__restrict__ src,
               cpx<T> *__restrict__ Z) {
  const int64_t g = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (g >= total) return;
  const int64_t M = P.nf, k = g % M, l = g / M;
  const cpx<T> *s = src + line_offset(P, l, false);
  const cpx<T> a = c2r_in(P, s, k), b0 = c2r_in(P, s, M - k);
  const cpx<T> b{b0.x, -b0.y};
  const cpx<T> d{a.x - b.x, a.y - b.y};
  const cpx<T> w = reinterpret_cast<const cpx<T> *>(P.rtw)[k];
  const cpx<T> wd = cmul(w, d);
  Z[l * M + k] = cpx<T>{a.x + b.x - wd.y, a.y + b.y + wd.x};   // + i * wd
}

// ---- long r2r lines (extended length nf > PFK_MAXLINE) ----
// u_q of the extension (layout in pfk.h) from the embedded real line: the
// line index p and sign that land on extension slot q, or p = -1 (zero)
__device__ __forceinline__ int64_t r2r_source(int kind, int64_t n, int64_t nf, int64_t q,
                                              int *sgn) {
  *sgn = 1;
  switch (kind) {
    case PFK_DHT: return q;
    case PFK_REDFT00: return q < n ? q : nf - q;
    case PFK_REDFT10:
      if (!(q & 1)) return -1;
      return q < 2 * n ? (q - 1) / 2 : (nf - q - 1) / 2;
    case PFK_REDFT01:
      if (q < n) return q;
      return q > 3 * n ? nf - q : -1;
    case PFK_REDFT11:
      if (!(q & 1)) return -1;
      if (q < 2 * n) return (q - 1) / 2;
      return q > 6 * n ? (nf - q - 1) / 2 : -1;
    case PFK_RODFT00:
      if (q >= 1 && q <= n) return q - 1;
      *sgn = -1;
      return (q >= n + 2) ? nf - q - 1 : -1;
    case PFK_RODFT10:
      if (!(q & 1)) return -1;
      if (q < 2 * n) return (q - 1) / 2;
      *sgn = -1;
      return (nf - q - 1) / 2;
    case PFK_RODFT01:
      if (q >= 1 && q <= n) return q - 1;
      *sgn = -1;
      return q > 3 * n ? nf - q - 1 : -1;
    default:   // RODFT11
      if (!(q & 1)) return -1;
      if (q < 2 * n) return (q - 1) / 2;
      *sgn = -1;
      return q > 6 * n ? (nf - q - 1) / 2 : -1;
  }
}

template <typename T>
__global__ void __launch_bounds__(256)
k_r2r_long_pre(const pfk_pass P, const int64_t total, const T *__restrict__ src,
               cpx<T> *__restrict__ U) {
  const int64_t g = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (g >= total) return;
  const int64_t nf = P.nf, q = g % nf, l = g / nf;
  int sgn;
  const int64_t p = r2r_source(P.r2r, P.n, nf, q, &sgn);
  const int64_t j = p - P.zin;   // index into the embedded input line
  T v = 0;
  if (p >= 0 && j >= 0 && j < P.nin)
    v = src[line_offset(P, l, false) + j * P.sin] * (T)(sgn * mod_sign(P.imod, j));
  U[g] = cpx<T>{v, (T)0};
}

template <typename T>
__global__ void __launch_bounds__(256)
k_r2r_long_post(const pfk_pass P, const int64_t total, const cpx<T> *__restrict__ F,
                T *__restrict__ dst) {
  const int64_t g = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (g >= total) return;
  const int64_t k = g % P.nout, l = g / P.nout, kk = P.zout + k;
  const cpx<T> *f = F + l * P.nf;
  T y;
  switch (P.r2r) {
    case PFK_DHT: y = f[kk].x - f[kk].y; break;
    case PFK_REDFT00: case PFK_REDFT10: y = f[kk].x; break;
    case PFK_REDFT01: case PFK_REDFT11: y = f[2 * kk + 1].x; break;
    case PFK_RODFT00: case PFK_RODFT10: y = -f[kk + 1].y; break;
    default: y = -f[2 * kk + 1].y; break;   // RODFT01, RODFT11
  }
  dst[line_offset(P, l, true) + k * P.sout] = y * (T)mod_sign(P.omod, k);
}

static int64_t pass_lines(const pfk_pass *p) {
  int64_t lines = p->text;
  for (int i = 0; i < p->nb; i++) lines *= p->bext[i];
  return lines;
}

// dense line-major layout of `lines` lines of M complex elements in a work
// buffer, in the pass's line order (tile index innermost)
static void dense_lines(const pfk_pass *p, int64_t M, int64_t *tst, int64_t *bst) {
  *tst = M;
  int64_t w = M * p->text;
  for (int i = 0; i < p->nb; i++) { bst[i] = w; w *= p->bext[i]; }
}

// ---- r2r through an inner FFT (r2r_path 1, 2; modes in pfk.h) ----
// element j of the embedded, modulated input line (0 outside the embed)
template <typename T>
__device__ __forceinline__ T r2r_x(const pfk_pass &P, const T *s, int64_t j) {
  const int64_t i = j - P.zin;
  if (j < 0 || i < 0 || i >= P.nin) return (T)0;
  return s[i * P.sin] * (T)mod_sign(P.imod, i);
}

// mode 0: element m of the real line whose real FFT the kind needs
template <typename T>
__device__ __forceinline__ T r2r_v(const pfk_pass &P, const T *s, int64_t m) {
  const int64_t n = P.n;
  switch (P.r2r) {
    case PFK_DHT: return r2r_x(P, s, m);
    case PFK_REDFT10: case PFK_RODFT10: {   // Makhoul: v_m = x_2m, v_(n-1-m') = x_(2m'+1)
      const int64_t j = m < (n + 1) / 2 ? 2 * m : 2 * (n - 1 - m) + 1;
      const T v = r2r_x(P, s, j);
      return (P.r2r == PFK_RODFT10 && (j & 1)) ? -v : v;   // DST-II: x_j (-1)^j
    }
    case PFK_REDFT00: return r2r_x(P, s, m < n ? m : 2 * (n - 1) - m);
    default:   // RODFT00
      if (m >= 1 && m <= n) return r2r_x(P, s, m - 1);
      return (m >= n + 2) ? -r2r_x(P, s, 2 * (n + 1) - m - 1) : (T)0;
  }
}

// mode 1: V_j = e^(i pi j/(2n)) (x'_j - i x'_(n-j)), x'_n = 0, j <= n/2
// (x' = x, reversed for DST-III); Im of V_0 and V_(n/2) dropped as a c2r
// drops them
template <typename T>
__device__ __forceinline__ cpx<T> r2r_V(const pfk_pass &P, const T *s, int64_t j) {
  const int64_t n = P.n;
  const bool rev = P.r2r == PFK_RODFT01;
  const T a = r2r_x(P, s, rev ? n - 1 - j : j);
  const T b = (j == 0) ? (T)0 : r2r_x(P, s, rev ? j - 1 : n - j);
  const cpx<T> t = reinterpret_cast<const cpx<T> *>(P.tw2)[j];   // e^(-i pi j/(2n))
  cpx<T> v = cmul(cpx<T>{t.x, -t.y}, cpx<T>{a, -b});
  if (j == 0 || 2 * j == n) v.y = 0;
  return v;
}

// complex index of element q of line l in a work array
__device__ __forceinline__ int64_t widx(int lm, int64_t l, int64_t q, int64_t lines,
                                        int64_t M) {
  return lm ? l * M + q : q * lines + l;
}

// path 1, pre: element q < M of the inner complex line of line l
template <typename T>
__global__ void __launch_bounds__(256)
k_r2r_pre1(const pfk_pass P, const int64_t lines, const int64_t M, const int lm,
           const T *__restrict__ src, cpx<T> *__restrict__ w1) {
  const int64_t g = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (g >= lines * M) return;
  int64_t q, l;
  if (lm) split_index(g, M, &q, &l);
  else split_index(g, lines, &l, &q);
  const int64_t n = P.n;
  const T *s = src + line_offset(P, l, false);
  const int mode = pfk_r2r_mode(P.r2r);
  cpx<T> z;
  if (mode == 0) {
    z = cpx<T>{r2r_v(P, s, 2 * q), r2r_v(P, s, 2 * q + 1)};
  } else if (mode == 1) {   // c2r combine: (V_q + conj V_(M-q)) + i W_n^q (V_q - conj V_(M-q))
    const cpx<T> a = r2r_V(P, s, q), b0 = r2r_V(P, s, M - q);
    const cpx<T> b{b0.x, -b0.y};
    const cpx<T> wd = cmul(reinterpret_cast<const cpx<T> *>(P.rtw)[q], cpx<T>{a.x - b.x, a.y - b.y});
    z = cpx<T>{a.x + b.x - wd.y, a.y + b.y + wd.x};
  } else {   // z_q = (x'_2q + i x'_(n-1-2q)) W_8n^(4q+1), x' reversed for DST-IV
    const bool rev = P.r2r == PFK_RODFT11;
    const T a = r2r_x(P, s, rev ? n - 1 - 2 * q : 2 * q);
    const T b = r2r_x(P, s, rev ? 2 * q : n - 1 - 2 * q);
    z = cmul(cpx<T>{a, b}, reinterpret_cast<const cpx<T> *>(P.tw2)[4 * q + 1]);
  }
  w1[widx(lm, l, q, lines, M)] = z;
}

// path 1, post: kept output k of line l from the inner result Z
template <typename T>
__global__ void __launch_bounds__(256)
k_r2r_post1(const pfk_pass P, const int64_t lines, const int64_t M, const int lm,
            const cpx<T> *__restrict__ Z, T *__restrict__ dst) {
  const int64_t g = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (g >= lines * P.nout) return;
  int64_t k, l;
  if (lm) split_index(g, P.nout, &k, &l);
  else split_index(g, lines, &l, &k);
  const int64_t n = P.n, kk = P.zout + k;
  const int mode = pfk_r2r_mode(P.r2r);
  const cpx<T> *tw2 = reinterpret_cast<const cpx<T> *>(P.tw2);
  auto z = [&](int64_t m) { return Z[widx(lm, l, m, lines, M)]; };
  T y;
  if (mode == 0) {
    // V_m of the real line (length L = 2M): E_m + W_L^m O_m from Z_m and
    // conj Z_(M-m); for m > M the conjugate of V_(L-m)
    auto V = [&](int64_t m) -> cpx<T> {
      const bool up = m > M;
      const int64_t mm = up ? 2 * M - m : m;
      const cpx<T> a = z(mm % M), b0 = z((M - mm) % M);
      const cpx<T> b{b0.x, -b0.y};
      const cpx<T> e{(a.x + b.x) * (T)0.5, (a.y + b.y) * (T)0.5};
      const cpx<T> o{(a.y - b.y) * (T)0.5, -(a.x - b.x) * (T)0.5};
      cpx<T> v = cmul(reinterpret_cast<const cpx<T> *>(P.rtw)[mm], o);
      v = cpx<T>{v.x + e.x, v.y + e.y};
      return up ? cpx<T>{v.x, -v.y} : v;
    };
    switch (P.r2r) {
      case PFK_DHT: { const cpx<T> v = V(kk); y = v.x - v.y; break; }
      case PFK_REDFT10: case PFK_RODFT10: {
        const int64_t m = P.r2r == PFK_RODFT10 ? n - 1 - kk : kk;   // DST-II: reversed
        y = 2 * cmul(tw2[m], V(m)).x;
        break;
      }
      case PFK_REDFT00: y = V(kk).x; break;
      default: y = -V(kk + 1).y; break;   // RODFT00
    }
  } else if (mode == 1) {   // y_2m = v_m, y_(2m+1) = v_(n-1-m); v_2j + i v_2j+1 = Z_j
    const int64_t idx = (kk & 1) ? n - 1 - (kk - 1) / 2 : kk / 2;
    const cpx<T> c = z(idx / 2);
    y = (idx & 1) ? c.y : c.x;
    if (P.r2r == PFK_RODFT01 && (kk & 1)) y = -y;   // DST-III: (-1)^k
  } else {   // u_j = Z_j W_8n^(4j); y_2j = 2 Re u_j, y_(n-1-2j) = -2 Im u_j
    const int64_t j = (kk & 1) ? (n - 1 - kk) / 2 : kk / 2;
    const cpx<T> u = cmul(z(j), tw2[4 * j]);
    y = (kk & 1) ? -2 * u.y : 2 * u.x;
    if (P.r2r == PFK_RODFT11 && (kk & 1)) y = -y;   // DST-IV: (-1)^k
  }
  dst[line_offset(P, l, true) + k * P.sout] = y * (T)mod_sign(P.omod, k);
}

// path 2 (odd lengths), pre: one element of the inner r2c / c2r input line
template <typename T>
__global__ void __launch_bounds__(256)
k_r2r_pre2(const pfk_pass P, const int64_t total, const int64_t q_per_line, const int64_t ld1,
           const T *__restrict__ src, T *__restrict__ w1) {
  const int64_t g = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (g >= total) return;
  const int64_t q = g % q_per_line, l = g / q_per_line;
  const T *s = src + line_offset(P, l, false);
  T *o = w1 + l * ld1;
  if (pfk_r2r_mode(P.r2r) == 0) {
    o[q] = r2r_v(P, s, q);
  } else {   // the half spectrum V_j, j <= n/2, of the inverse Makhoul step
    const cpx<T> v = r2r_V(P, s, q);
    o[2 * q] = v.x;
    o[2 * q + 1] = v.y;
  }
}

// path 2, post: one kept output from the inner r2c / c2r result
template <typename T>
__global__ void __launch_bounds__(256)
k_r2r_post2(const pfk_pass P, const int64_t total, const int64_t ld2, const T *__restrict__ w2,
            T *__restrict__ dst) {
  const int64_t g = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (g >= total) return;
  const int64_t k = g % P.nout, l = g / P.nout, n = P.n;
  const int64_t kk = P.zout + k;
  const T *w = w2 + l * ld2;
  const cpx<T> *F = reinterpret_cast<const cpx<T> *>(w);
  const cpx<T> *tw2 = reinterpret_cast<const cpx<T> *>(P.tw2);
  const int64_t L = pfk_r2r_inner_len(P.r2r, n);
  auto half = [&](int64_t m) -> cpx<T> {
    return m <= L / 2 ? F[m] : cpx<T>{F[L - m].x, -F[L - m].y};
  };
  T y;
  switch (P.r2r) {
    case PFK_DHT: { const cpx<T> f = half(kk); y = f.x - f.y; break; }
    case PFK_REDFT10: case PFK_RODFT10: {
      const int64_t m = P.r2r == PFK_RODFT10 ? n - 1 - kk : kk;
      y = 2 * cmul(tw2[m], half(m)).x;
      break;
    }
    case PFK_REDFT00: y = F[kk].x; break;
    case PFK_RODFT00: y = -F[kk + 1].y; break;
    default: {   // REDFT01, RODFT01
      const T v = w[(kk & 1) ? n - 1 - (kk - 1) / 2 : kk / 2];
      y = (P.r2r == PFK_RODFT01 && (kk & 1)) ? -v : v;
      break;
    }
  }
  dst[line_offset(P, l, true) + k * P.sout] = y * (T)mod_sign(P.omod, k);
}

template <typename T>
static int r2r_real_path_t(const pfk_pass *p, const void *src, void *dst) {
  pfk_pass in;
  int64_t ld1, ld2;
  int path;
  if (pfk_r2r_inner(p, &in, &ld1, &ld2, &path) || !p->work || !p->work2) {
    set_error_msg("pfk: bad r2r pass");
    return 1;
  }
  const int64_t lines = pass_lines(p), out = lines * p->nout;
  if (lines <= 0 || out <= 0) return 0;
  const dim3 blk(256);
  if (path == 1) {
    const int64_t M = in.nf;
    const int lm1 = p->sin == 1, lm2 = p->sout == 1;
    hipLaunchKernelGGL(k_r2r_pre1<T>, dim3((unsigned)((lines * M + 255) / 256)), blk, 0,
                       g_stream, *p, lines, M, lm1, (const T *)src, (cpx<T> *)p->work);
  } else {
    const int mode = pfk_r2r_mode(p->r2r);
    const int64_t qpl = mode == 0 ? pfk_r2r_inner_len(p->r2r, p->n) : p->n / 2 + 1;
    hipLaunchKernelGGL(k_r2r_pre2<T>, dim3((unsigned)((lines * qpl + 255) / 256)), blk, 0,
                       g_stream, *p, lines * qpl, qpl, ld1, (const T *)src, (T *)p->work);
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) { set_error("pfk r2r pre", e); return 1; }
  if (pfk_launch_pass(&in, p->work, p->work2)) return 1;
  if (path == 1)
    hipLaunchKernelGGL(k_r2r_post1<T>, dim3((unsigned)((out + 255) / 256)), blk, 0, g_stream, *p,
                       lines, (int64_t)in.nf, (int)(p->sout == 1), (const cpx<T> *)p->work2,
                       (T *)dst);
  else
    hipLaunchKernelGGL(k_r2r_post2<T>, dim3((unsigned)((out + 255) / 256)), blk, 0, g_stream, *p,
                       out, ld2, (const T *)p->work2, (T *)dst);
  e = hipGetLastError();
  if (e != hipSuccess) { set_error("pfk r2r post", e); return 1; }
  return 0;
}

static int r2r_real_path(const pfk_pass *p, const void *src, void *dst) {
  return p->prec == PFK_F64 ? r2r_real_path_t<double>(p, src, dst)
                            : r2r_real_path_t<float>(p, src, dst);
}

// r2r: extension into work2, the long c2c in place on work2, output picks
static int long_r2r(const pfk_pass *p, const void *src, void *dst) {
  const int64_t nf = p->nf, lines = pass_lines(p);
  if (!p->work2) { set_error_msg("pfk: bad long r2r pass"); return 1; }
  const int64_t tot = lines * nf;
  if (tot <= 0) return 0;
  if (p->prec == PFK_F64)
    hipLaunchKernelGGL(k_r2r_long_pre<double>, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0,
                       g_stream, *p, tot, (const double *)src, (cpx<double> *)p->work2);
  else
    hipLaunchKernelGGL(k_r2r_long_pre<float>, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0,
                       g_stream, *p, tot, (const float *)src, (cpx<float> *)p->work2);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) { set_error("pfk r2r long extension", e); return 1; }
  pfk_pass c = *p;
  c.kind = PFK_C2C;
  c.n = nf;
  c.nin = c.nout = nf;
  c.zin = c.zout = 0;
  c.imod.on = c.omod.on = 0;
  c.sin = c.sout = 1;
  dense_lines(p, nf, &c.tsin, c.bsin);
  dense_lines(p, nf, &c.tsout, c.bsout);
  c.src_elems = c.dst_elems = tot;
  if (long_c2c(&c, p->work2, p->work2)) return 1;   // step A reads work2 before step B writes it
  const int64_t out = lines * p->nout;
  if (out <= 0) return 0;
  if (p->prec == PFK_F64)
    hipLaunchKernelGGL(k_r2r_long_post<double>, dim3((unsigned)((out + 255) / 256)), dim3(256), 0,
                       g_stream, *p, out, (const cpx<double> *)p->work2, (double *)dst);
  else
    hipLaunchKernelGGL(k_r2r_long_post<float>, dim3((unsigned)((out + 255) / 256)), dim3(256), 0,
                       g_stream, *p, out, (const cpx<float> *)p->work2, (float *)dst);
  e = hipGetLastError();
  if (e != hipSuccess) { set_error("pfk r2r long output", e); return 1; }
  return 0;
}

static int launch_long(const pfk_pass *p, const void *src, void *dst) {
  if (p->kind == PFK_C2C) return long_c2c(p, src, dst);
  if (p->kind == PFK_R2R) return long_r2r(p, src, dst);
  if ((p->kind != PFK_R2C && p->kind != PFK_C2R) || !p->work2 || p->n != 2 * p->nf) {
    set_error_msg("pfk: bad long real-line pass");
    return 1;
  }
  const int64_t M = p->nf, lines = pass_lines(p);
  pfk_pass c = *p;   // the half-length complex line
  c.kind = PFK_C2C;
  c.n = M;
  c.hw = 0;
  c.rtw = NULL;
  c.nin = c.nout = M;
  c.zin = c.zout = 0;
  c.imod.on = c.omod.on = 0;
  hipError_t e;
  if (p->kind == PFK_R2C) {
    // real input viewed as complex: contiguous lines, offsets halved
    c.sin = 1;
    c.tsin = p->tsin / 2;
    for (int i = 0; i < p->nb; i++) c.bsin[i] = p->bsin[i] / 2;
    c.src_elems = p->src_elems / 2;
    c.sout = 1;
    dense_lines(p, M, &c.tsout, c.bsout);
    c.dst_elems = lines * M;
    if (long_c2c(&c, src, p->work2)) return 1;
    const int64_t total = lines * p->nout;
    if (total <= 0) return 0;
    if (p->prec == PFK_F64)
      hipLaunchKernelGGL(k_r2c_long_post<double>, dim3((unsigned)((total + 255) / 256)), dim3(256),
                         0, g_stream, *p, total, (const cpx<double> *)p->work2,
                         (cpx<double> *)dst);
    else
      hipLaunchKernelGGL(k_r2c_long_post<float>, dim3((unsigned)((total + 255) / 256)), dim3(256),
                         0, g_stream, *p, total, (const cpx<float> *)p->work2, (cpx<float> *)dst);
    e = hipGetLastError();
    if (e != hipSuccess) { set_error("pfk r2c long combine", e); return 1; }
    return 0;
  }
  const int64_t total = lines * M;
  if (total <= 0) return 0;
  if (p->prec == PFK_F64)
    hipLaunchKernelGGL(k_c2r_long_pre<double>, dim3((unsigned)((total + 255) / 256)), dim3(256), 0,
                       g_stream, *p, total, (const cpx<double> *)src, (cpx<double> *)p->work2);
  else
    hipLaunchKernelGGL(k_c2r_long_pre<float>, dim3((unsigned)((total + 255) / 256)), dim3(256), 0,
                       g_stream, *p, total, (const cpx<float> *)src, (cpx<float> *)p->work2);
  e = hipGetLastError();
  if (e != hipSuccess) { set_error("pfk c2r long combine", e); return 1; }
  // real output viewed as complex: contiguous lines, offsets halved
  c.sin = 1;
  dense_lines(p, M, &c.tsin, c.bsin);
  c.src_elems = lines * M;
  c.sout = 1;
  c.tsout = p->tsout / 2;
  for (int i = 0; i < p->nb; i++) c.bsout[i] = p->bsout[i] / 2;
  c.dst_elems = p->dst_elems / 2;
  return long_c2c(&c, p->work2, dst);
}

template <typename T>
__global__ void k_scale(T *d, int64_t n, T f) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    d[i] *= f;
}

}  // namespace pfk

using namespace pfk;

extern "C" {

int pfk_init(void) {
  if (g_stream) return 0;
  hipError_t e = hipFree(0);
  if (e != hipSuccess) { set_error("pfk_init(hipFree(0))", e); return 1; }
  int dev = 0;
  (void)hipGetDevice(&dev);
  int v = 0;
  if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess &&
      v > 0)
    g_max_lds = v < 65536 ? v : 65536;
  const char *go = getenv("PFFT_AMD_GENERIC_ONLY");
  g_generic_only = (go && go[0] == '1') ? 1 : 0;
  e = hipStreamCreateWithFlags(&g_stream, hipStreamNonBlocking);
  if (e != hipSuccess) { set_error("pfk_init(hipStreamCreate)", e); g_stream = nullptr; return 1; }
  return 0;
}

int pfk_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

const char *pfk_last_error(void) { return g_err; }

int pfk_malloc(void **p, size_t bytes) {
  if (pfk_init()) return 1;
  hipError_t e = hipMalloc(p, bytes ? bytes : 1);
  if (e != hipSuccess) { set_error("pfk_malloc", e); *p = nullptr; return 1; }
  return 0;
}
int pfk_malloc_host(void **p, size_t bytes) {
  if (pfk_init()) return 1;
  hipError_t e = hipHostMalloc(p, bytes ? bytes : 1, hipHostMallocDefault);
  if (e != hipSuccess) { set_error("pfk_malloc_host", e); *p = nullptr; return 1; }
  return 0;
}
int pfk_free(void *p) {
  if (!p) return 0;
  hipError_t e = hipFree(p);
  if (e != hipSuccess) { set_error("pfk_free", e); return 1; }
  return 0;
}
int pfk_free_host(void *p) {
  if (!p) return 0;
  hipError_t e = hipHostFree(p);
  if (e != hipSuccess) { set_error("pfk_free_host", e); return 1; }
  return 0;
}

int pfk_is_device_ptr(const void *p) {
  if (!p || pfk_init()) return 0;
  hipPointerAttribute_t a;
  memset(&a, 0, sizeof(a));
  hipError_t e = hipPointerGetAttributes(&a, p);
  if (e != hipSuccess) { (void)hipGetLastError(); return 0; }
  return (a.type == hipMemoryTypeDevice || a.type == hipMemoryTypeManaged) ? 1 : 0;
}

static int do_copy(void *dst, const void *src, size_t bytes, hipMemcpyKind k, const char *w) {
  if (!bytes) return 0;
  if (pfk_init()) return 1;
  hipError_t e = hipMemcpyAsync(dst, src, bytes, k, g_stream);
  if (e == hipSuccess) e = hipStreamSynchronize(g_stream);
  if (e != hipSuccess) { set_error(w, e); return 1; }
  return 0;
}
int pfk_memcpy_h2d(void *d, const void *s, size_t b) { return do_copy(d, s, b, hipMemcpyHostToDevice, "pfk_memcpy_h2d"); }
int pfk_memcpy_d2h(void *d, const void *s, size_t b) { return do_copy(d, s, b, hipMemcpyDeviceToHost, "pfk_memcpy_d2h"); }
int pfk_memcpy_d2d(void *d, const void *s, size_t b) { return do_copy(d, s, b, hipMemcpyDeviceToDevice, "pfk_memcpy_d2d"); }

int pfk_memset(void *d, int v, size_t bytes) {
  if (!bytes) return 0;
  if (pfk_init()) return 1;
  hipError_t e = hipMemsetAsync(d, v, bytes, g_stream);
  if (e == hipSuccess) e = hipStreamSynchronize(g_stream);
  if (e != hipSuccess) { set_error("pfk_memset", e); return 1; }
  return 0;
}

int pfk_sync(void) {
  if (pfk_init()) return 1;
  hipError_t e = hipStreamSynchronize(g_stream);
  if (e != hipSuccess) { set_error("pfk_sync", e); return 1; }
  return 0;
}

int pfk_r2r_inner(const pfk_pass *p, pfk_pass *in, int64_t *ld1, int64_t *ld2, int *path) {
  if (p->kind != PFK_R2R) return 1;
  const int mode = pfk_r2r_mode(p->r2r);
  const int64_t n = p->n, L = pfk_r2r_inner_len(p->r2r, n);
  if (p->r2r == PFK_REDFT00 && n < 2) return 1;
  const int even = mode == 0 ? L % 2 == 0 : n % 2 == 0;
  if (mode == 2 && !even) return 1;   // the n/2-point form needs even n
  int64_t lines = p->text;
  for (int i = 0; i < p->nb; i++) lines *= p->bext[i];
  memset(in, 0, sizeof(*in));
  in->prec = p->prec;
  in->xcd_remap = p->xcd_remap;
  in->tw = p->tw;
  in->text = lines;
  in->nb = 0;
  if (even) {
    // path 1: complex FFT of length M; each work array line-major where the
    // pass's lines are contiguous on that side, else lines innermost
    const int64_t M = mode == 0 ? L / 2 : n / 2;
    *path = 1;
    in->kind = PFK_C2C;
    in->sign = mode == 1 ? +1 : -1;
    in->n = in->nf = in->nin = in->nout = M;
    const int lm1 = p->sin == 1, lm2 = p->sout == 1;
    in->sin = lm1 ? 1 : lines;
    in->tsin = lm1 ? M : 1;
    in->sout = lm2 ? 1 : lines;
    in->tsout = lm2 ? M : 1;
    *ld1 = *ld2 = 2 * M;
    in->src_elems = in->dst_elems = lines * M;
  } else {
    // path 2: r2c / c2r of the odd length, line-major
    *path = 2;
    in->sin = in->sout = 1;
    if (mode == 0) {
      in->kind = PFK_R2C;
      in->sign = -1;
      in->n = in->nin = in->nf = L;
      in->nout = L / 2 + 1;
      *ld1 = L + 1;                     // even pitch: complex pairs stay aligned
      *ld2 = 2 * (L / 2 + 1);
      in->tsin = *ld1;
      in->tsout = *ld2 / 2;
    } else {
      in->kind = PFK_C2R;
      in->sign = +1;
      in->n = in->nout = in->nf = n;
      in->nin = n / 2 + 1;
      *ld1 = 2 * (n / 2 + 1);
      *ld2 = n + 1;
      in->tsin = *ld1 / 2;
      in->tsout = *ld2;
    }
    in->src_elems = lines * (in->kind == PFK_R2C ? *ld1 : *ld1 / 2);
    in->dst_elems = lines * (in->kind == PFK_C2R ? *ld2 : *ld2 / 2);
  }
  in->nstages = pfk_factorize(in->nf, in->radix);
  return in->nstages < 0;
}

int pfk_pass_check(const pfk_pass *p) {
  if (pfk_init()) return 1;
  if (!g_generic_only && (fast_covers(p) || small_covers(p) || mfma_covers(p))) return 0;
  int C, pitch;
  size_t lds;
  return generic_tiling(p, &C, &pitch, &lds);
}

int pfk_launch_pass(const pfk_pass *p, const void *src, void *dst) {
  if (pfk_init()) return 1;
  if (p->kind == PFK_R2R && p->r2r_path) return r2r_real_path(p, src, dst);
  if (p->n1) return launch_long(p, src, dst);
  if (!g_generic_only && p->mfma) {
    const int m = launch_mfma(p, src, dst);
    if (m == 0) return 0;
    if (m < 0) return 1;
  }
  // variant 1: general kernel; 3: short-length kernel first where it has
  // one; else the power-of-two kernels, then the short-length kernels
  int r = 1;
  if (!g_generic_only && p->variant == 3) r = launch_small(p, src, dst);
  if (r > 0) r = launch_fast(p, src, dst);
  if (r > 0 && !g_generic_only && p->variant != 1) r = launch_small(p, src, dst);
  if (r == 0) return 0;
  if (r < 0) return 1;
  return launch_generic(p, src, dst);
}

int pfk_event_create(void **ev) {
  if (pfk_init()) return 1;
  hipEvent_t e;
  if (hipEventCreate(&e) != hipSuccess) return 1;
  *ev = (void *)e;
  return 0;
}
int pfk_event_record(void *ev) {
  return hipEventRecord((hipEvent_t)ev, g_stream) == hipSuccess ? 0 : 1;
}
int pfk_event_elapsed_ms(void *a, void *b, float *ms) {
  if (hipEventSynchronize((hipEvent_t)b) != hipSuccess) return 1;
  return hipEventElapsedTime(ms, (hipEvent_t)a, (hipEvent_t)b) == hipSuccess ? 0 : 1;
}
int pfk_event_destroy(void *ev) { return hipEventDestroy((hipEvent_t)ev) == hipSuccess ? 0 : 1; }

int pfk_scale(void *data, int prec, int64_t nreal, double factor) {
  if (pfk_init()) return 1;
  if (nreal <= 0) return 0;
  int64_t blocks = (nreal + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  if (prec == PFK_F64)
    hipLaunchKernelGGL(k_scale<double>, dim3((unsigned)blocks), dim3(256), 0, g_stream,
                       (double *)data, nreal, factor);
  else
    hipLaunchKernelGGL(k_scale<float>, dim3((unsigned)blocks), dim3(256), 0, g_stream,
                       (float *)data, nreal, (float)factor);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) { set_error("pfk_scale", e); return 1; }
  return 0;
}

}  // extern "C"

