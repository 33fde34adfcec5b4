// kernels_pow2.hip -- register-resident Stockham pass kernels for
// power-of-two lengths 64..4096 on gfx950.
//
// Work split.  One workgroup transforms C lines of length NF (the
// computed complex FFT length).  Each line is handled by T = NF/8
// threads; thread (c, t) holds the 8 elements t + T*m (m = 0..7) of line c
// in registers.  Thread index = t*C + c, so the C lanes with equal t are
// adjacent: every global access by a wave covers 64/C rows of C elements
// (tile dimension contiguous) or C runs of 64/C elements (line contiguous),
// i.e. 128-byte segments for C = 8 and 16-byte complex doubles, whatever the
// load/store layout.
//
// Stages.  NF = 8^a * 2^b: a radix-8 stages and, if b > 0, one final
// radix-2^b stage in which each thread does 8/R butterflies.  With T
// threads per line, the inputs of every stage for thread t are exactly
// positions t + T*m, so only the outputs move: between stages each thread
// writes its outputs to LDS and reads back positions t + T*m.  LDS is laid
// out [position][line] (index p*C + c): the 8 lanes that issue one 16-byte
// access together touch 8 consecutive 16-byte words, so every exchange is
// bank-conflict free without padding.  After the last stage thread t holds
// X[t + T*m] in natural order.
//
// Real transforms (even n, NF = n/2) use the half-length packing:
// r2c loads z_p = x_2p + i x_2p+1, transforms, and unpacks
//   X_k = (Z_k + conj Z_{h-k})/2 - i W^k (Z_k - conj Z_{h-k})/2,  W = e^{sign 2 pi i/n};
// c2r packs Z_k = E_k + i O_k with E_k = X_k + conj X_{h-k},
// O_k = (X_k - conj X_{h-k}) W^k before the transform and stores
// x_2p = Re z_p, x_2p+1 = Im z_p.  The mirror partner Z_{h-k} lives in
// another thread, so both use one extra LDS exchange.
#include <stdlib.h>
#include "pfk_internal.hpp"

#ifndef PFK_NONTEMPORAL
#define PFK_NONTEMPORAL 1
#endif

namespace pfk {

template <typename T>
using C2 = cpx<T>;

template <typename T>
__device__ __forceinline__ C2<T> cadd(C2<T> a, C2<T> b) { return C2<T>{a.x + b.x, a.y + b.y}; }
template <typename T>
__device__ __forceinline__ C2<T> csub(C2<T> a, C2<T> b) { return C2<T>{a.x - b.x, a.y - b.y}; }
// multiply by SIGN*i
template <int SIGN, typename T>
__device__ __forceinline__ C2<T> mul_si(C2<T> a) {
  return SIGN < 0 ? C2<T>{a.y, -a.x} : C2<T>{-a.y, a.x};
}

template <int SIGN, typename T>
__device__ __forceinline__ void dft2(C2<T> &a, C2<T> &b) {
  const C2<T> s = cadd(a, b), d = csub(a, b);
  a = s;
  b = d;
}

// in-place DFT-4 of (a0,a1,a2,a3) with W4 = SIGN*i
template <int SIGN, typename T>
__device__ __forceinline__ void dft4(C2<T> &a0, C2<T> &a1, C2<T> &a2, C2<T> &a3) {
  const C2<T> c0 = cadd(a0, a2), c1 = csub(a0, a2);
  const C2<T> c2 = cadd(a1, a3), c3 = mul_si<SIGN>(csub(a1, a3));
  a0 = cadd(c0, c2);
  a2 = csub(c0, c2);
  a1 = cadd(c1, c3);
  a3 = csub(c1, c3);
}

// in-place DFT-8 of v[0..7] (natural order in and out)
template <int SIGN, typename T>
__device__ __forceinline__ void dft8(C2<T> *v) {
  C2<T> e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6];
  C2<T> o0 = v[1], o1 = v[3], o2 = v[5], o3 = v[7];
  dft4<SIGN>(e0, e1, e2, e3);
  dft4<SIGN>(o0, o1, o2, o3);
  const T s = (T)0.70710678118654752440084436210484903928L;
  // W8^1 = s(1 + SIGN i), W8^2 = SIGN i, W8^3 = s(-1 + SIGN i)
  const C2<T> t1{s * (o1.x - (SIGN) * o1.y), s * (o1.y + (SIGN) * o1.x)};
  const C2<T> t2 = mul_si<SIGN>(o2);
  const C2<T> t3{s * (-o3.x - (SIGN) * o3.y), s * (-o3.y + (SIGN) * o3.x)};
  v[0] = cadd(e0, o0);
  v[4] = csub(e0, o0);
  v[1] = cadd(e1, t1);
  v[5] = csub(e1, t1);
  v[2] = cadd(e2, t2);
  v[6] = csub(e2, t2);
  v[3] = cadd(e3, t3);
  v[7] = csub(e3, t3);
}

// Streaming global accesses: every element is read once and written once
// per pass, so loads and stores bypass cache retention (nontemporal).
template <typename T>
__device__ __forceinline__ C2<T> ld_stream(const C2<T> *p) {
#if PFK_NONTEMPORAL
  typedef T v2 __attribute__((ext_vector_type(2)));
  const v2 v = __builtin_nontemporal_load(reinterpret_cast<const v2 *>(p));
  return C2<T>{v.x, v.y};
#else
  return *p;
#endif
}
template <typename T>
__device__ __forceinline__ void st_stream(C2<T> *p, C2<T> v) {
#if PFK_NONTEMPORAL
  typedef T v2 __attribute__((ext_vector_type(2)));
  v2 w;
  w.x = v.x;
  w.y = v.y;
  __builtin_nontemporal_store(w, reinterpret_cast<v2 *>(p));
#else
  *p = v;
#endif
}

__device__ __forceinline__ int msign(const pfk_mod &m, int64_t j) {
  if (!m.on) return 1;
  const int64_t t = j + m.start;
  if (t < m.half) return ((t & 1) ? -1 : 1) * m.extra;
  return 1;
}

constexpr int ilog2c(int n) { return n <= 1 ? 0 : 1 + ilog2c(n / 2); }

// One Stockham stage of radix R on the thread's 8 registers.
// b = t + i*T for i < 8/R; inputs m = i + r*(8/R); twiddle W_NF^(r*k*NF/(L*R)).
// Returns the outputs in `v` at register m' = i + q*(8/R) and, unless
// LAST, the LDS position of each output in `pos`.
template <typename T, int NF, int R, int L, int SIGN, bool LAST>
__device__ __forceinline__ void stage(C2<T> *v, int *pos, const int t,
                                      const C2<T> *__restrict__ tw) {
  constexpr int TT = NF / 8;
  constexpr int NB = 8 / R;
  constexpr int M = NF / (L * R);
  C2<T> o[8];
#pragma unroll
  for (int i = 0; i < NB; ++i) {
    const int b = t + i * TT;
    const int k = b & (L - 1);
    C2<T> a[R];
#pragma unroll
    for (int r = 0; r < R; ++r) a[r] = v[i + r * NB];
    if (L > 1) {
#pragma unroll
      for (int r = 1; r < R; ++r) a[r] = cmul(a[r], tw[r * k * M]);
    }
    if (R == 8) dft8<SIGN>(a);
    else if (R == 4) dft4<SIGN>(a[0], a[1], a[2], a[3]);
    else dft2<SIGN>(a[0], a[1]);
#pragma unroll
    for (int q = 0; q < R; ++q) {
      o[i + q * NB] = a[q];
      if (!LAST) pos[i + q * NB] = (b - k) * R + k + q * L;
    }
  }
#pragma unroll
  for (int m = 0; m < 8; ++m) v[m] = o[m];
}

// Exchange through LDS: write v[m] at position pos[m], read back t + T*m.
template <typename T, int NF, int C>
__device__ __forceinline__ void exchange(C2<T> *v, const int *pos, C2<T> *lds, const int t,
                                         const int c) {
  constexpr int TT = NF / 8;
#pragma unroll
  for (int m = 0; m < 8; ++m) lds[pos[m] * C + c] = v[m];
  __syncthreads();
#pragma unroll
  for (int m = 0; m < 8; ++m) v[m] = lds[(t + TT * m) * C + c];
  __syncthreads();
}

// Full length-NF FFT on registers; first stage input = v (positions t+T*m),
// output X[t + T*m] in v[m].
template <typename T, int NF, int C, int SIGN>
__device__ __forceinline__ void fft_regs(C2<T> *v, C2<T> *lds, const int t, const int c,
                                         const C2<T> *__restrict__ tw) {
  constexpr int LG = ilog2c(NF);
  constexpr int A = LG / 3;   // radix-8 stages
  constexpr int B = LG % 3;   // final radix-2^B stage
  int pos[8];
  // unrolled stage list (NF <= 4096: A <= 4)
  if constexpr (A >= 1) {
    stage<T, NF, 8, 1, SIGN, (A == 1 && B == 0)>(v, pos, t, tw);
    if constexpr (!(A == 1 && B == 0)) exchange<T, NF, C>(v, pos, lds, t, c);
  }
  if constexpr (A >= 2) {
    stage<T, NF, 8, 8, SIGN, (A == 2 && B == 0)>(v, pos, t, tw);
    if constexpr (!(A == 2 && B == 0)) exchange<T, NF, C>(v, pos, lds, t, c);
  }
  if constexpr (A >= 3) {
    stage<T, NF, 8, 64, SIGN, (A == 3 && B == 0)>(v, pos, t, tw);
    if constexpr (!(A == 3 && B == 0)) exchange<T, NF, C>(v, pos, lds, t, c);
  }
  if constexpr (A >= 4) {
    stage<T, NF, 8, 512, SIGN, (A == 4 && B == 0)>(v, pos, t, tw);
    if constexpr (!(A == 4 && B == 0)) exchange<T, NF, C>(v, pos, lds, t, c);
  }
  constexpr int LA = (A == 0) ? 1 : (A == 1 ? 8 : (A == 2 ? 64 : (A == 3 ? 512 : 4096)));
  if constexpr (B == 1) stage<T, NF, 2, LA, SIGN, true>(v, pos, t, tw);
  if constexpr (B == 2) stage<T, NF, 4, LA, SIGN, true>(v, pos, t, tw);
}

struct TileCtx {
  int64_t sb, db;
  int nl;
};

// Workgroups are dispatched round-robin over the 8 XCDs (each with its own
// L2).  With remap on, XCD x runs a contiguous range of tiles in order, so
// tiles that share cache lines at their edges run on the same L2 close in
// time.  The map is a bijection on [0, nblk) for any nblk.
__device__ __forceinline__ int64_t xcd_remap(int64_t b, int64_t nblk) {
  const int64_t q = nblk >> 3, r = nblk & 7;
  const int64_t x = b & 7, j = b >> 3;
  return x * q + (x < r ? x : r) + j;
}

__device__ __forceinline__ TileCtx tile_of(const pfk_pass &P, int C, int64_t tile_base) {
  const int64_t ntt = (P.text + C - 1) / C;
  const int64_t tileid =
      tile_base + (P.xcd_remap ? xcd_remap(blockIdx.x, gridDim.x) : (int64_t)blockIdx.x);
  const int64_t tt = tileid % ntt;
  int64_t bflat = tileid / ntt;
  const int64_t c0 = tt * C;
  TileCtx r;
  r.nl = (int)((P.text - c0) < C ? (P.text - c0) : C);
  r.sb = c0 * P.tsin;
  r.db = c0 * P.tsout;
  for (int i = 0; i < P.nb; ++i) {
    const int64_t e = P.bext[i];
    const int64_t d = bflat % e;
    bflat /= e;
    r.sb += d * P.bsin[i];
    r.db += d * P.bsout[i];
  }
  return r;
}

// ------------------------------------------------------------------ c2c
template <typename T, int NF, int C, int SIGN>
__global__ void __launch_bounds__(C *NF / 8)
k_c2c_pow2(const pfk_pass P, const int64_t tile_base, const T *__restrict__ src,
           T *__restrict__ dst) {
  constexpr int TT = NF / 8;
  __shared__ C2<T> lds[NF * C];
  const int tid = threadIdx.x;
  const int c = tid % C, t = tid / C;
  const TileCtx tc = tile_of(P, C, tile_base);
  const bool valid = c < tc.nl;
  const C2<T> *s = reinterpret_cast<const C2<T> *>(src) + tc.sb + (int64_t)c * P.tsin;
  C2<T> *d = reinterpret_cast<C2<T> *>(dst) + tc.db + (int64_t)c * P.tsout;
  const C2<T> *tw = reinterpret_cast<const C2<T> *>(P.tw);
  const int zin = (int)P.zin, nin = (int)P.nin, zout = (int)P.zout, nout = (int)P.nout;
  C2<T> v[8];
#pragma unroll
  for (int m = 0; m < 8; ++m) {
    const int j = t + TT * m - zin;
    C2<T> x{0, 0};
    if (valid && j >= 0 && j < nin) {
      x = ld_stream(s + (int64_t)j * P.sin);
      const T f = (T)msign(P.imod, j);
      x.x *= f;
      x.y *= f;
    }
    v[m] = x;
  }
  fft_regs<T, NF, C, SIGN>(v, lds, t, c, tw);
  if (valid) {
#pragma unroll
    for (int m = 0; m < 8; ++m) {
      const int k = t + TT * m - zout;
      if (k >= 0 && k < nout) {
        const T f = (T)msign(P.omod, k);
        st_stream(d + (int64_t)k * P.sout, C2<T>{v[m].x * f, v[m].y * f});
      }
    }
  }
}

// ------------------------------------------------------------------ r2c
// real line of n = 2*NF contiguous samples (sin == 1, no embed) ->
// kept half-spectrum outputs [zout, zout+nout) of the n/2+1.
template <typename T, int NF, int C, int SIGN>
__global__ void __launch_bounds__(C *NF / 8)
k_r2c_pow2(const pfk_pass P, const int64_t tile_base, const T *__restrict__ src,
           T *__restrict__ dst) {
  constexpr int TT = NF / 8;
  constexpr int H = NF;  // h = n/2
  __shared__ C2<T> lds[NF * C];
  const int tid = threadIdx.x;
  const int c = tid % C, t = tid / C;
  const TileCtx tc = tile_of(P, C, tile_base);
  const bool valid = c < tc.nl;
  const C2<T> *s = reinterpret_cast<const C2<T> *>(src + tc.sb + (int64_t)c * P.tsin);
  C2<T> *d = reinterpret_cast<C2<T> *>(dst) + tc.db + (int64_t)c * P.tsout;
  const C2<T> *tw = reinterpret_cast<const C2<T> *>(P.tw);
  const C2<T> *rtw = reinterpret_cast<const C2<T> *>(P.rtw);
  C2<T> v[8];
#pragma unroll
  for (int m = 0; m < 8; ++m) {
    const int p = t + TT * m;
    C2<T> x{0, 0};
    if (valid) {
      x = ld_stream(s + p);
      if (P.imod.on) {
        x.x *= (T)msign(P.imod, 2 * p);
        x.y *= (T)msign(P.imod, 2 * p + 1);
      }
    }
    v[m] = x;
  }
  fft_regs<T, NF, C, SIGN>(v, lds, t, c, tw);
  // mirror exchange: Z_{h-k}
#pragma unroll
  for (int m = 0; m < 8; ++m) lds[(t + TT * m) * C + c] = v[m];
  __syncthreads();
  const int zout = (int)P.zout, nout = (int)P.nout;
  if (valid) {
#pragma unroll
    for (int m = 0; m < 8; ++m) {
      const int k = t + TT * m;
      const C2<T> a = v[m];
      const C2<T> b = lds[((H - k) & (H - 1)) * C + c];
      // E = (a + conj b)/2, O = -i (a - conj b)/2, X = E + W^k O
      const C2<T> E{(a.x + b.x) * (T)0.5, (a.y - b.y) * (T)0.5};
      const C2<T> O{(a.y + b.y) * (T)0.5, -(a.x - b.x) * (T)0.5};
      const int ko = k - zout;
      if (ko >= 0 && ko < nout) {
        T f = (T)msign(P.omod, ko);
        if (P.hw == 2 && k != 0) f *= (T)2;
        const C2<T> WO = cmul(O, rtw[k]);
        st_stream(d + (int64_t)ko * P.sout, C2<T>{(E.x + WO.x) * f, (E.y + WO.y) * f});
      }
      if (k == 0) {  // Nyquist X_h = E_0 - O_0 (W^h = -1)
        const int kn = H - zout;
        if (kn >= 0 && kn < nout) {
          const T f = (T)msign(P.omod, kn);
          d[(int64_t)kn * P.sout] = C2<T>{(E.x - O.x) * f, (E.y - O.y) * f};
        }
      }
    }
  }
}

// ------------------------------------------------------------------ c2r
// half spectrum (embedded at zin) -> real line of n = 2*NF contiguous
// samples (sout == 1, no trunc).
template <typename T, int NF, int C, int SIGN>
__global__ void __launch_bounds__(C *NF / 8)
k_c2r_pow2(const pfk_pass P, const int64_t tile_base, const T *__restrict__ src,
           T *__restrict__ dst) {
  constexpr int TT = NF / 8;
  constexpr int H = NF;
  __shared__ C2<T> lds[(NF + 1) * C];
  const int tid = threadIdx.x;
  const int c = tid % C, t = tid / C;
  const TileCtx tc = tile_of(P, C, tile_base);
  const bool valid = c < tc.nl;
  const C2<T> *s = reinterpret_cast<const C2<T> *>(src) + tc.sb + (int64_t)c * P.tsin;
  C2<T> *d = reinterpret_cast<C2<T> *>(dst + tc.db + (int64_t)c * P.tsout);
  const C2<T> *tw = reinterpret_cast<const C2<T> *>(P.tw);
  const C2<T> *rtw = reinterpret_cast<const C2<T> *>(P.rtw);
  const int zin = (int)P.zin, nin = (int)P.nin;
  // load X_k, k = t + T*m, and X_h (thread t == 0) into LDS
  C2<T> v[8];
#pragma unroll
  for (int m = 0; m < 8; ++m) {
    const int k = t + TT * m;
    const int j = k - zin;
    C2<T> x{0, 0};
    if (valid && j >= 0 && j < nin) {
      x = ld_stream(s + (int64_t)j * P.sin);
      T f = (T)msign(P.imod, j);
      if (P.hw == 1 && k != 0) f *= (T)0.5;
      x.x *= f;
      x.y *= f;
    }
    v[m] = x;
    lds[k * C + c] = x;
  }
  if (t == 0) {
    const int j = H - zin;
    C2<T> x{0, 0};
    if (valid && j >= 0 && j < nin) {
      x = s[(int64_t)j * P.sin];
      x.x *= (T)msign(P.imod, j);
    }
    lds[H * C + c] = x;
  }
  __syncthreads();
  // Z_k = E_k + i O_k
#pragma unroll
  for (int m = 0; m < 8; ++m) {
    const int k = t + TT * m;
    const C2<T> a = v[m];
    C2<T> z;
    if (k == 0) {
      const T xh = lds[H * C + c].x;
      z = C2<T>{a.x + xh, a.x - xh};
    } else {
      const C2<T> b = lds[(H - k) * C + c];
      const C2<T> E{a.x + b.x, a.y - b.y};
      const C2<T> O = cmul(C2<T>{a.x - b.x, a.y + b.y}, rtw[k]);
      z = C2<T>{E.x - O.y, E.y + O.x};
    }
    v[m] = z;
  }
  __syncthreads();
  fft_regs<T, NF, C, SIGN>(v, lds, t, c, tw);
  if (valid) {
#pragma unroll
    for (int m = 0; m < 8; ++m) {
      const int p = t + TT * m;
      C2<T> x = v[m];
      if (P.omod.on) {
        x.x *= (T)msign(P.omod, 2 * p);
        x.y *= (T)msign(P.omod, 2 * p + 1);
      }
      st_stream(d + p, x);
    }
  }
}

// ------------------------------------------------------------------ dispatch
template <typename T, int NF, int C>
static int launch_nf(const pfk_pass *p, const void *src, void *dst) {
  int64_t ntiles = (p->text + C - 1) / C;
  for (int i = 0; i < p->nb; ++i) ntiles *= p->bext[i];
  if (ntiles <= 0) return 0;
  const T *s = (const T *)src;
  T *d = (T *)dst;
  const int64_t chunk = (int64_t)1 << 30;
  for (int64_t base = 0; base < ntiles; base += chunk) {
    const int64_t nblk = (ntiles - base) < chunk ? (ntiles - base) : chunk;
    dim3 grid((unsigned)nblk), block(C * NF / 8);
    if (p->kind == PFK_C2C) {
      if (p->sign < 0)
        hipLaunchKernelGGL((k_c2c_pow2<T, NF, C, -1>), grid, block, 0, g_stream, *p, base, s, d);
      else
        hipLaunchKernelGGL((k_c2c_pow2<T, NF, C, +1>), grid, block, 0, g_stream, *p, base, s, d);
    } else if (p->kind == PFK_R2C) {
      if (p->sign < 0)
        hipLaunchKernelGGL((k_r2c_pow2<T, NF, C, -1>), grid, block, 0, g_stream, *p, base, s, d);
      else
        hipLaunchKernelGGL((k_r2c_pow2<T, NF, C, +1>), grid, block, 0, g_stream, *p, base, s, d);
    } else {
      if (p->sign < 0)
        hipLaunchKernelGGL((k_c2r_pow2<T, NF, C, -1>), grid, block, 0, g_stream, *p, base, s, d);
      else
        hipLaunchKernelGGL((k_c2r_pow2<T, NF, C, +1>), grid, block, 0, g_stream, *p, base, s, d);
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) { set_error("pfk_launch_pass(pow2)", e); return -1; }
  }
  return 0;
}

static int g_cols512 = 8;

template <typename T>
static int launch_pow2(const pfk_pass *p, const void *src, void *dst) {
  switch (p->nf) {
    case 64: return launch_nf<T, 64, 8>(p, src, dst);
    case 128: return launch_nf<T, 128, 8>(p, src, dst);
    case 256: return launch_nf<T, 256, 8>(p, src, dst);
    case 512:
      if (p->kind == PFK_C2C && g_cols512 == 4) return launch_nf<T, 512, 4>(p, src, dst);
      if (p->kind == PFK_C2C && g_cols512 == 16) return launch_nf<T, 512, 16>(p, src, dst);
      return launch_nf<T, 512, 8>(p, src, dst);
    case 1024: return launch_nf<T, 1024, 4>(p, src, dst);
    case 2048: return launch_nf<T, 2048, 2>(p, src, dst);
    default: return 1;
  }
}

int launch_fast(const pfk_pass *p, const void *src, void *dst) {
  if (g_generic_only) return 1;
  {
    static int init = 0;
    if (!init) {
      const char *e = getenv("PFFT_AMD_COLS512");
      if (e) g_cols512 = atoi(e);
      init = 1;
    }
  }
  if (p->kind == PFK_COPY) return 1;
  const int64_t nf = p->nf;
  if (nf < 64 || nf > 2048 || (nf & (nf - 1))) return 1;
  if (p->kind == PFK_R2C) {
    // real input: contiguous, not embedded
    if (p->n != 2 * nf || p->sin != 1 || p->zin != 0 || p->nin != p->n) return 1;
    if (p->hw == 1) return 1;
  }
  if (p->kind == PFK_C2R) {
    if (p->n != 2 * nf || p->sout != 1 || p->zout != 0 || p->nout != p->n) return 1;
    if (p->hw == 2) return 1;
  }
  if (p->kind == PFK_C2C && p->hw) return 1;
  const int r = (p->prec == PFK_F64) ? launch_pow2<double>(p, src, dst)
                                     : launch_pow2<float>(p, src, dst);
  return r;   // 0 launched, 1 not covered, -1 launch error
}

}  // namespace pfk
