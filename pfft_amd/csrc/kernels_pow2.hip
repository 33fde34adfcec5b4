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

// in-place DFT-16 of v[0..15] as 4x4: DFT-4 over stride-4 groups, twiddle
// W16^(r*q), DFT-4 across groups
template <int SIGN, typename T>
__device__ __forceinline__ void dft16(C2<T> *v) {
#pragma unroll
  for (int r = 0; r < 4; ++r) dft4<SIGN>(v[r], v[r + 4], v[r + 8], v[r + 12]);
  // v[r + 4q] now holds sub-DFT output q of group r; multiply by W16^(r*q)
  const T c1 = (T)0.92387953251128675612818318939678828682L;  // cos(pi/8)
  const T s1 = (T)0.38268343236508977172845998403039886676L;  // sin(pi/8)
  const T h = (T)0.70710678118654752440084436210484903928L;
  const T S = (T)SIGN;
  // W16^1 = c1 + S i s1, W16^2 = h + S i h, W16^3 = s1 + S i c1,
  // W16^4 = S i, W16^6 = -h + S i h, W16^9 = -c1 - S i s1
  const C2<T> w1{c1, S * s1}, w2{h, S * h}, w3{s1, S * c1};
  const C2<T> w6{-h, S * h}, w9{-c1, -S * s1};
  v[5] = cmul(v[5], w1);
  v[9] = cmul(v[9], w2);
  v[13] = cmul(v[13], w3);
  v[6] = cmul(v[6], w2);
  v[10] = mul_si<SIGN>(v[10]);
  v[14] = cmul(v[14], w6);
  v[7] = cmul(v[7], w3);
  v[11] = cmul(v[11], w6);
  v[15] = cmul(v[15], w9);
#pragma unroll
  for (int q = 0; q < 4; ++q) dft4<SIGN>(v[4 * q], v[4 * q + 1], v[4 * q + 2], v[4 * q + 3]);
  // output k = q + 4*s sits in v[4q + s]: transpose the 4x4 index
  C2<T> o[16];
#pragma unroll
  for (int q = 0; q < 4; ++q)
#pragma unroll
    for (int t = 0; t < 4; ++t) o[q + 4 * t] = v[4 * q + t];
#pragma unroll
  for (int k = 0; k < 16; ++k) v[k] = o[k];
}

// +-1 modulation factor of index j, computed with selects only: a branch
// here would sit between a load and its use and make the compiler wait for
// every load before issuing the next one.
template <typename T>
__device__ __forceinline__ T msign(const pfk_mod &m, int64_t j) {
  const int64_t t = j + m.start;
  const int s = (m.on && t < m.half) ? ((t & 1) ? -m.extra : m.extra) : 1;
  return (T)s;
}

constexpr int ilog2c(int n) { return n <= 1 ? 0 : 1 + ilog2c(n / 2); }

// One Stockham stage of radix R on the thread's E registers (E = 8 or 16,
// T = NF/E threads per line).  Butterflies b = t + i*T for i < E/R; inputs
// m = i + r*(E/R); twiddle W_NF^(r*k*NF/(L*R)), k = b mod L.  Outputs go to
// register m' = i + q*(E/R) and, unless LAST, their LDS positions to `pos`.
template <typename T, int NF, int E, int R, int L, int SIGN, bool LAST>
__device__ __forceinline__ void stage(C2<T> *v, int *pos, const int t,
                                      const C2<T> *__restrict__ tw) {
  constexpr int TT = NF / E;
  constexpr int NB = E / R;
  constexpr int M = NF / (L * R);
  C2<T> o[E];
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
    if (R == 16) dft16<SIGN>(a);
    else if (R == 8) dft8<SIGN>(a);
    else if (R == 4) dft4<SIGN>(a[0], a[1], a[2], a[3]);
    else dft2<SIGN>(a[0], a[1]);
#pragma unroll
    for (int q = 0; q < R; ++q) {
      o[i + q * NB] = a[q];
      if (!LAST) pos[i + q * NB] = (b - k) * R + k + q * L;
    }
  }
#pragma unroll
  for (int m = 0; m < E; ++m) v[m] = o[m];
}

// LDS slot of position p of line c.  Tile-major lane order (consecutive
// lanes = consecutive lines): [position][line], conflict-free without
// padding.  Line-major lane order (consecutive lanes = consecutive
// positions of one line): [line][position] with one slot of padding per 16
// positions, so the radix-8/16 output scatters (position strides 8 and 16)
// spread over the banks, and an odd row pitch, so that the tile-major
// accesses of a staged load or store (consecutive lanes = consecutive
// lines at one position) do too.  The row has room for position NF (the
// Nyquist bin of c2r).
template <int NF>
constexpr int lm_pitch() { return NF + NF / 16 + 1; }
template <int NF, int C, bool LM>
__device__ __forceinline__ int lidx(int p, int c) {
  if constexpr (LM) return c * lm_pitch<NF>() + p + (p >> 4);
  else return p * C + c;
}
template <int NF, int C, bool LM>
constexpr int lds_slots() { return LM ? C * lm_pitch<NF>() : NF * C; }

// In line-major order a line's NF/E lanes are consecutive threads, so with
// at most 64 of them every line lies inside one wavefront, and an exchange
// within lines only reorders data among the lanes of one wave.  A wave's own
// LDS accesses execute in order, so such an exchange needs only the
// compiler to keep them in order (wavefront-scope fences), not a workgroup
// barrier that would hold every wave of the tile.
template <int NF, int E, bool LM>
constexpr bool wave_lines() { return LM && NF / E <= 64; }

template <bool WAVE>
__device__ __forceinline__ void line_sync() {
  if constexpr (WAVE) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront", "local");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront", "local");
  } else {
    lds_barrier();
  }
}

// Exchange through LDS: write v[m] at position pos[m], read back t + T*m.
template <typename T, int NF, int C, int E, bool LM>
__device__ __forceinline__ void exchange(C2<T> *v, const int *pos, C2<T> *lds, const int t,
                                         const int c) {
  constexpr int TT = NF / E;
#pragma unroll
  for (int m = 0; m < E; ++m) lds[lidx<NF, C, LM>(pos[m], c)] = v[m];
  line_sync<wave_lines<NF, E, LM>()>();
#pragma unroll
  for (int m = 0; m < E; ++m) v[m] = lds[lidx<NF, C, LM>(t + TT * m, c)];
  line_sync<wave_lines<NF, E, LM>()>();
}

// Full length-NF FFT on registers; first stage input = v (positions t+T*m),
// output X[t + T*m] in v[m].  Stages of radix E, then one of radix
// NF / E^A if NF is not a power of E.
template <typename T, int NF, int C, int E, int SIGN, int L, bool LM>
__device__ __forceinline__ void fft_stages(C2<T> *v, int *pos, C2<T> *lds, const int t,
                                           const int c, const C2<T> *__restrict__ tw) {
  constexpr int REM = NF / L;  // remaining length to transform
  if constexpr (REM >= E) {
    constexpr bool LAST = (REM == E);
    stage<T, NF, E, E, L, SIGN, LAST>(v, pos, t, tw);
    if constexpr (!LAST) {
      exchange<T, NF, C, E, LM>(v, pos, lds, t, c);
      fft_stages<T, NF, C, E, SIGN, L * E, LM>(v, pos, lds, t, c, tw);
    }
  } else if constexpr (REM > 1) {
    stage<T, NF, E, REM, L, SIGN, true>(v, pos, t, tw);
  }
}

template <typename T, int NF, int C, int E, int SIGN, bool LM = false>
__device__ __forceinline__ void fft_regs(C2<T> *v, C2<T> *lds, const int t, const int c,
                                         const C2<T> *__restrict__ tw) {
  int pos[E];
  fft_stages<T, NF, C, E, SIGN, 1, LM>(v, pos, lds, t, c, tw);
}

// Exchanges of a length-NF FFT in radix-E stages (stages minus one).
__host__ __device__ constexpr int n_exchanges(int nf, int e) {
  int s = 0, l = 1;
  while (nf / l >= e) { l *= e; ++s; }
  if (nf / l > 1) ++s;
  return s - 1;
}

// Exchange through the line-major LDS layout with a change of lane order:
// v[m] is written at position pos[m] of line cw and read back as position
// tr + T*m of line cr.  WAVE: both sides in line-major order with every line
// inside one wave (no change of order), so wavefront-scope ordering suffices.
template <typename T, int NF, int C, int E, bool WAVE = false>
__device__ __forceinline__ void exchange_x(C2<T> *v, const int *pos, C2<T> *lds, const int cw,
                                           const int tr, const int cr) {
  constexpr int TT = NF / E;
#pragma unroll
  for (int m = 0; m < E; ++m) lds[lidx<NF, C, true>(pos[m], cw)] = v[m];
  line_sync<WAVE>();
#pragma unroll
  for (int m = 0; m < E; ++m) v[m] = lds[lidx<NF, C, true>(tr + TT * m, cr)];
  line_sync<WAVE>();
}

// The FFT of fft_regs with a change of lane order at exchange SW (1-based):
// stages before it run as thread (ta, ca) (position index, line), the
// exchange writes in that order and reads in order (tb, cb), and the later
// stages run as (tb, cb).  K counts the exchanges done.  One of the two
// orders is line-major (the kernels stage tile-major loads or stores), so
// every exchange but the switching one stays within line-major lines.
template <typename T, int NF, int C, int E, int SIGN, int L, int K, int SW>
__device__ __forceinline__ void fft_stages_sw(C2<T> *v, int *pos, C2<T> *lds, const int ta,
                                              const int ca, const int tb, const int cb,
                                              const C2<T> *__restrict__ tw) {
  constexpr int REM = NF / L;
  constexpr bool BEFORE = K < SW;
  const int t = BEFORE ? ta : tb;
  const int c = BEFORE ? ca : cb;
  if constexpr (REM >= E) {
    constexpr bool LAST = (REM == E);
    stage<T, NF, E, E, L, SIGN, LAST>(v, pos, t, tw);
    if constexpr (!LAST) {
      constexpr bool AFTER = K + 1 >= SW;   // the next stage runs as (tb, cb)
      constexpr bool SWITCH = K + 1 == SW;
      exchange_x<T, NF, C, E, !SWITCH && wave_lines<NF, E, true>()>(v, pos, lds, c,
                                                                     AFTER ? tb : ta,
                                                                     AFTER ? cb : ca);
      fft_stages_sw<T, NF, C, E, SIGN, L * E, K + 1, SW>(v, pos, lds, ta, ca, tb, cb, tw);
    }
  } else if constexpr (REM > 1) {
    stage<T, NF, E, REM, L, SIGN, true>(v, pos, t, tw);
  }
}

struct TileCtx {
  int64_t sb, db;
  int64_t c0;   // first line of the tile along the tile dimension
  int nl;
};

// Workgroups are dispatched round-robin over the 8 XCDs (each with its own
// L2).  With remap on, XCD x runs a contiguous range of tiles in order, so
// tiles that share cache lines at their edges run on the same L2 close in
// time.  The map is a bijection on [0, n) for any n; item i runs on XCD
// i mod 8 because the persistent grid is a multiple of 8.
__device__ __forceinline__ int64_t xcd_remap(int64_t b, int64_t n) {
  const int64_t q = n >> 3, r = n & 7;
  const int64_t x = b & 7, j = b >> 3;
  return x * q + (x < r ? x : r) + j;
}

// Tile of work item `item` (0 <= item < ntiles).  Tile counts below 2^32
// (every practical pass) decode with 32-bit divisions, several times
// shorter than the 64-bit chain.
__device__ __forceinline__ TileCtx tile_at(const pfk_pass &P, int C, int64_t item,
                                           int64_t ntiles) {
  const int64_t ntt = (P.text + C - 1) / C;
  const int64_t tileid = P.xcd_remap ? xcd_remap(item, ntiles) : item;
  int64_t tt, sb = 0, db = 0;
  if (ntiles <= (int64_t)0xFFFFFFFFu) {
    uint32_t id = (uint32_t)tileid;
    const uint32_t n32 = (uint32_t)ntt;
    tt = id % n32;
    id /= n32;
    for (int i = 0; i < P.nb; ++i) {
      const uint32_t e = (uint32_t)P.bext[i];
      const uint32_t d = id % e;
      id /= e;
      sb += (int64_t)d * P.bsin[i];
      db += (int64_t)d * P.bsout[i];
    }
  } else {
    tt = tileid % ntt;
    int64_t bflat = tileid / ntt;
    for (int i = 0; i < P.nb; ++i) {
      const int64_t e = P.bext[i];
      const int64_t d = bflat % e;
      bflat /= e;
      sb += d * P.bsin[i];
      db += d * P.bsout[i];
    }
  }
  const int64_t c0 = tt * C;
  TileCtx r;
  r.c0 = c0;
  r.nl = (int)((P.text - c0) < C ? (P.text - c0) : C);
  r.sb = sb + c0 * P.tsin;
  r.db = db + c0 * P.tsout;
  return r;
}

// Persistent pass kernels.  A fixed grid of resident workgroups walks the
// tiles (workgroup b takes items item0 + b, + G, + 2G, ... below item_end);
// launched with one workgroup per tile instead (G = item_end - item0) the
// same kernel does one tile per workgroup.  While a tile is
// transformed and stored, the loads of the workgroup's next tile are
// already in flight in registers (software pipelining across tiles), so
// HBM stays busy through the compute and LDS phases.  On the last item the
// prefetch degenerates to every lane reading the tile's first element (one
// cached line), which keeps the load sequence free of branches.
//
// Loads are issued unconditionally from clamped in-bounds addresses and
// masked afterwards: a branch between a load and its use would make the
// compiler wait for each load in turn.

// Twiddle tables live in LDS when they fit next to the exchange buffer
// (limit 156 KiB per workgroup): LDS reads count on lgkmcnt, so a twiddle
// read does not wait for the prefetched tile the way a global load would
// (vmcnt completes in order).
template <typename T>
constexpr bool tw_in_lds(int slots, int nf, int extra) {
  return ((size_t)slots + nf + extra) * sizeof(C2<T>) <= 156 * 1024;
}
template <typename T>
__device__ __forceinline__ void stage_table(C2<T> *dst, const C2<T> *src, int n) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) dst[i] = src[i];
}

// c2c: raw loads of elements t + T*m (minus the embed offset) of line c
template <typename T, int NF, int C, int E>
__device__ __forceinline__ void load_c2c(const pfk_pass &P, const T *src, const TileCtx &tc,
                                         const bool live, const int c, const int t,
                                         C2<T> *v) {
  constexpr int TT = NF / E;
  const int zin = (int)P.zin, nin = (int)P.nin;
  const C2<T> *s0 = reinterpret_cast<const C2<T> *>(src) + tc.sb;
  const C2<T> *s = (live && c < tc.nl) ? s0 + (int64_t)c * P.tsin : s0;
  const int64_t sin = live ? P.sin : 0;
#pragma unroll
  for (int m = 0; m < E; ++m) {
    const int j = t + TT * m - zin;
    const bool ok = j >= 0 && j < nin;
    v[m] = ld_stream(s + (int64_t)(ok ? j : 0) * sin);
  }
}

// c2c loads of a plain pass (no embed, no modulation): one 64-bit address
// per thread, then constant strides (TT*sin, uniform) between its elements
template <typename T, int NF, int C, int E>
__device__ __forceinline__ void load_c2c_pl(const pfk_pass &P, const T *src, const TileCtx &tc,
                                            const bool live, const int c, const int t,
                                            C2<T> *v) {
  constexpr int TT = NF / E;
  const C2<T> *s0 = reinterpret_cast<const C2<T> *>(src) + tc.sb;
  const C2<T> *s = (live && c < tc.nl) ? s0 + (int64_t)c * P.tsin + (int64_t)t * P.sin : s0;
  const int64_t st = live ? (int64_t)TT * P.sin : 0;
#pragma unroll
  for (int m = 0; m < E; ++m) v[m] = ld_stream(s + m * st);
}

// Paired-line stores for complex64 tiles whose lines are adjacent in
// memory (tile stride 1), tile-major lane order (line c = 2j + e of the
// tile in lane parity e).  Every 16-byte store covers two adjacent lines at
// one position: lane e = 0 of a pair writes the even element slots m = 0,
// 2, 4, ... of lines 2j and 2j+1, lane e = 1 the odd ones, after one DPP
// lane swap (quad_perm 1,0,3,2: lane ^ 1) per value has handed each lane
// its partner's values.  A wave instruction then moves 128-byte segments
// of 8 positions (the 4 of 8-byte lanes otherwise), the access shape of
// the complex128 tiles.  An odd last line is stored with 8-byte stores.
__device__ __forceinline__ float lane_xor1(float v) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0xB1, 0xF, 0xF, true));
}

typedef float pf4 __attribute__((ext_vector_type(4)));

// stores of lines c < tc.nl: pairs as 16-byte stores, an odd last line with
// 8-byte stores
template <int NF, int C, int E>
__device__ __forceinline__ void store_c2c_pair(const pfk_pass &P, float *dst, const TileCtx &tc,
                                               const int c, const int t, const C2<float> *v) {
  constexpr int TT = NF / E;
  const int e = c & 1, j = c >> 1;
  C2<float> send[E / 2], recv[E / 2];
#pragma unroll
  for (int i = 0; i < E / 2; ++i) {
    send[i] = e ? v[2 * i] : v[2 * i + 1];
    recv[i] = C2<float>{lane_xor1(send[i].x), lane_xor1(send[i].y)};
  }
  C2<float> *d0 = reinterpret_cast<C2<float> *>(dst) + tc.db;
  if (2 * j + 1 < tc.nl) {
    C2<float> *d = d0 + 2 * j + (int64_t)(t + TT * e) * P.sout;
    const int64_t st = (int64_t)2 * TT * P.sout;
#pragma unroll
    for (int i = 0; i < E / 2; ++i) {
      const C2<float> lo = e ? recv[i] : v[2 * i], hi = e ? v[2 * i + 1] : recv[i];
      __builtin_nontemporal_store(pf4{lo.x, lo.y, hi.x, hi.y},
                                  reinterpret_cast<pf4 *>(d + i * st));
    }
  } else if (e == 0 && 2 * j < tc.nl) {
    C2<float> *d = d0 + 2 * j + (int64_t)t * P.sout;
#pragma unroll
    for (int m = 0; m < E; ++m) st_stream(d + (int64_t)TT * m * P.sout, v[m]);
  }
}

// LM: line-major lane order (thread = c*T + t: a wave's lanes run along a
// line), for passes whose lines are contiguous in memory; otherwise
// thread = t*C + c (a wave's lanes run across the tile's lines).
// XS (line-major plain passes only): 1 = the loads, 2 = the stores take the
// tile-major lane order; the FFT switches between the two orders at its
// first (1) or last (2) LDS exchange, so the staging costs no extra pass
// through LDS.  This is for a
// transposing pass whose lines are contiguous on one side and adjacent
// along the tile dimension on the other: complex64 has 8-byte elements,
// so tile-major lanes move 128-byte segments across 16 lines, and
// line-major lanes move 512-byte runs along one line.  Each side is
// accessed in the order that coalesces it.
// TW: output k of tile line c0 + c is multiplied by W_twn^((c0+c)*k) from
// the two-level table tw2 (first step of a long line, pfk.h).
// PL: plain pass (every input and output position present, no
// modulation): no masks or selects, and element addresses advance by a
// uniform stride, which halves the vector instructions spent per element
// outside the butterflies.
template <typename T, int NF, int C, int E, int SIGN, bool LM = false, bool TW = false,
          bool PL = false, int XS = 0, int PAIR = 0>
__global__ void __launch_bounds__(C *NF / E)
k_c2c_pow2(const pfk_pass P, const int64_t item0, const int64_t item_end, const int64_t ntiles,
           const T *__restrict__ src, T *__restrict__ dst) {
  constexpr int TT = NF / E;
  constexpr bool TWL = tw_in_lds<T>(lds_slots<NF, C, LM>(), NF, 0);
  __shared__ C2<T> lds[lds_slots<NF, C, LM>()];
  __shared__ C2<T> twl[TWL ? NF : 1];
  static_assert(XS == 0 || (LM && PL && !TW), "staged loads or stores: line-major plain passes");
  // PAIR: paired-line stores (complex64 plain passes, tile-major side)
  static_assert(PAIR == 0 || (sizeof(T) == 4 && PL && !TW && C % 2 == 0), "paired lines");
  static_assert(!PAIR || !LM || XS == 2, "paired stores need tile-major stores");
  const int tid = threadIdx.x;
  const int c0 = LM ? tid / TT : tid % C, t0 = LM ? tid % TT : tid / C;
  const int cx0 = tid % C, tx0 = tid / C;   // tile-major lane order of staged accesses
  const C2<T> *twg = reinterpret_cast<const C2<T> *>(P.tw);
  if constexpr (TWL) stage_table(twl, twg, NF);
  const C2<T> *tw = TWL ? twl : twg;
  const int zin = (int)P.zin, nin = (int)P.nin, zout = (int)P.zout, nout = (int)P.nout;
  int64_t item = item0 + blockIdx.x;
  TileCtx tc = tile_at(P, C, item, ntiles);
  // one step: prefetch the next tile into `nxt`, transform and store `cur`;
  // false after the last item.  Two steps per loop trip alternate the
  // buffers, so no register copy has to wait for the prefetch.
  auto step = [&](C2<T> *cur, C2<T> *nxt) -> bool {
    // lane indices made opaque per step: index arithmetic is recomputed in
    // the loop instead of being hoisted and kept live in registers
    int t = t0, c = c0, tx = tx0, cx = cx0;
    asm volatile("" : "+v"(t), "+v"(c), "+v"(tx), "+v"(cx));
    const int64_t ni = item + gridDim.x;
    const bool more = ni < item_end;
    const TileCtx tn = tile_at(P, C, more ? ni : item, ntiles);
    if constexpr (XS == 1) load_c2c_pl<T, NF, C, E>(P, src, tn, more, cx, tx, nxt);
    else if constexpr (PL) load_c2c_pl<T, NF, C, E>(P, src, tn, more, c, t, nxt);
    else load_c2c<T, NF, C, E>(P, src, tn, more, c, t, nxt);
    const bool valid = c < tc.nl;
    if constexpr (!PL) {
#pragma unroll
      for (int m = 0; m < E; ++m) {
        const int j = t + TT * m - zin;
        const bool ok = valid && j >= 0 && j < nin;
        const T f = msign<T>(P.imod, j);
        cur[m] = ok ? C2<T>{cur[m].x * f, cur[m].y * f} : C2<T>{0, 0};
      }
    }
    if constexpr (XS == 0) {
      fft_regs<T, NF, C, E, SIGN, LM>(cur, lds, t, c, tw);
    } else {
      // the lane order changes at one of the FFT's own exchanges: the first
      // (tile-major loads, line-major stores) or the last (the reverse)
      static_assert(n_exchanges(NF, E) >= 1, "staged passes need an exchange");
      int pos[E];
      if constexpr (XS == 1)
        fft_stages_sw<T, NF, C, E, SIGN, 1, 0, 1>(cur, pos, lds, tx, cx, t, c, tw);
      else
        fft_stages_sw<T, NF, C, E, SIGN, 1, 0, n_exchanges(NF, E)>(cur, pos, lds, t, c, tx, cx,
                                                                    tw);
    }
    if constexpr (PAIR) {   // tile-major order, paired lines (every lane takes part)
      if constexpr (XS == 2) store_c2c_pair<NF, C, E>(P, dst, tc, cx, tx, cur);
      else store_c2c_pair<NF, C, E>(P, dst, tc, c, t, cur);
    } else if constexpr (XS == 2) {   // registers in tile-major order: tile-major stores
      if (cx < tc.nl) {
        C2<T> *d = reinterpret_cast<C2<T> *>(dst) + tc.db + (int64_t)cx * P.tsout +
                   (int64_t)tx * P.sout;
        const int64_t st = (int64_t)TT * P.sout;
#pragma unroll
        for (int m = 0; m < E; ++m) st_stream(d + m * st, cur[m]);
      }
    } else if (PL && valid) {
      // lines are independent, so a ragged tile's unused lines (garbage in,
      // never stored) cannot affect the stored ones
      C2<T> *d = reinterpret_cast<C2<T> *>(dst) + tc.db + (int64_t)c * P.tsout +
                 (int64_t)t * P.sout;
      const int64_t st = (int64_t)TT * P.sout;
#pragma unroll
      for (int m = 0; m < E; ++m) st_stream(d + m * st, cur[m]);
    } else if (valid) {
      C2<T> *d = reinterpret_cast<C2<T> *>(dst) + tc.db + (int64_t)c * P.tsout;
#pragma unroll
      for (int m = 0; m < E; ++m) {
        const int k = t + TT * m - zout;
        if (k >= 0 && k < nout) {
          const T f = msign<T>(P.omod, k);
          C2<T> v{cur[m].x * f, cur[m].y * f};
          if constexpr (TW) {
            const C2<T> *lo = reinterpret_cast<const C2<T> *>(P.tw2);
            const int64_t mm = (tc.c0 + c) * (int64_t)(k + zout);
            v = cmul(v, cmul(lo[mm & 4095], lo[4096 + (mm >> 12)]));
          }
          st_stream(d + (int64_t)k * P.sout, v);
        }
      }
    }
    if (!more) return false;
    // the next tile reuses the LDS exchange buffer (staged lane orders
    // cross waves)
    line_sync<XS == 0 && wave_lines<NF, E, LM>()>();
    item = ni;
    tc = tn;
    return true;
  };
  C2<T> va[E], vb[E];
  if constexpr (XS == 1) load_c2c_pl<T, NF, C, E>(P, src, tc, true, cx0, tx0, va);
  else if constexpr (PL) load_c2c_pl<T, NF, C, E>(P, src, tc, true, c0, t0, va);
  else load_c2c<T, NF, C, E>(P, src, tc, true, c0, t0, va);
  if constexpr (TWL) lds_barrier();
  while (step(va, vb) && step(vb, va)) {
  }
}

// ------------------------------------------------------------------ r2c
// real line of n = 2*NF contiguous samples (sin == 1, no embed) ->
// kept half-spectrum outputs [zout, zout+nout) of the n/2+1.
template <typename T, int NF, int C, int E>
__device__ __forceinline__ void load_r2c(const pfk_pass &P, const T *src, const TileCtx &tc,
                                         const bool live, const int c, const int t,
                                         C2<T> *v) {
  constexpr int TT = NF / E;
  const T *s0 = src + tc.sb;
  const C2<T> *s = reinterpret_cast<const C2<T> *>((live && c < tc.nl) ? s0 + (int64_t)c * P.tsin
                                                                     : s0);
  const int step = live ? TT : 0;
#pragma unroll
  for (int m = 0; m < E; ++m) v[m] = ld_stream(s + (live ? t : 0) + step * m);
}

template <typename T, int NF, int C, int E, int SIGN, bool PL = false, bool LM = false>
__global__ void __launch_bounds__(C *NF / E)
k_r2c_pow2(const pfk_pass P, const int64_t item0, const int64_t item_end, const int64_t ntiles,
           const T *__restrict__ src, T *__restrict__ dst) {
  constexpr int TT = NF / E;
  constexpr int H = NF;  // h = n/2
  constexpr bool TWL = tw_in_lds<T>(lds_slots<NF, C, LM>(), NF, NF);
  __shared__ C2<T> lds[lds_slots<NF, C, LM>()];
  __shared__ C2<T> twl[TWL ? NF : 1], rtwl[TWL ? NF : 1];
  const int tid = threadIdx.x;
  const int c0 = LM ? tid / TT : tid % C, t0 = LM ? tid % TT : tid / C;
  const C2<T> *twg = reinterpret_cast<const C2<T> *>(P.tw);
  const C2<T> *rtwg = reinterpret_cast<const C2<T> *>(P.rtw);
  if constexpr (TWL) {
    stage_table(twl, twg, NF);
    stage_table(rtwl, rtwg, NF);
  }
  const C2<T> *tw = TWL ? twl : twg;
  const C2<T> *rtw = TWL ? rtwl : rtwg;
  const int zout = (int)P.zout, nout = (int)P.nout;
  int64_t item = item0 + blockIdx.x;
  TileCtx tc = tile_at(P, C, item, ntiles);
  auto step = [&](C2<T> *cur, C2<T> *nxt) -> bool {
    int t = t0, c = c0;
    asm volatile("" : "+v"(t), "+v"(c));
    const int64_t ni = item + gridDim.x;
    const bool more = ni < item_end;
    const TileCtx tn = tile_at(P, C, more ? ni : item, ntiles);
    load_r2c<T, NF, C, E>(P, src, tn, more, c, t, nxt);
    const bool valid = c < tc.nl;
    if constexpr (!PL) {
#pragma unroll
      for (int m = 0; m < E; ++m) {
        const int p = t + TT * m;
        const T fx = msign<T>(P.imod, 2 * p), fy = msign<T>(P.imod, 2 * p + 1);
        cur[m] = valid ? C2<T>{cur[m].x * fx, cur[m].y * fy} : C2<T>{0, 0};
      }
    }
    fft_regs<T, NF, C, E, SIGN, LM>(cur, lds, t, c, tw);
    // mirror exchange: Z_{h-k}
#pragma unroll
    for (int m = 0; m < E; ++m) lds[lidx<NF, C, LM>(t + TT * m, c)] = cur[m];
    line_sync<wave_lines<NF, E, LM>()>();
    if (PL && valid) {
      // every output k = t + T*m (and the Nyquist bin, thread t = 0) is kept
      C2<T> *d = reinterpret_cast<C2<T> *>(dst) + tc.db + (int64_t)c * P.tsout;
      C2<T> *dk = d + (int64_t)t * P.sout;
      const int64_t st = (int64_t)TT * P.sout;
#pragma unroll
      for (int m = 0; m < E; ++m) {
        const int k = t + TT * m;
        const C2<T> a = cur[m];
        const C2<T> b = lds[lidx<NF, C, LM>((H - k) & (H - 1), c)];
        const C2<T> Ev{(a.x + b.x) * (T)0.5, (a.y - b.y) * (T)0.5};
        const C2<T> O{(a.y + b.y) * (T)0.5, -(a.x - b.x) * (T)0.5};
        const C2<T> WO = cmul(O, rtw[k]);
        st_stream(dk + m * st, C2<T>{Ev.x + WO.x, Ev.y + WO.y});
        if (m == 0 && t == 0)
          st_stream(d + (int64_t)H * P.sout, C2<T>{Ev.x - O.x, Ev.y - O.y});
      }
    } else if (valid) {
      C2<T> *d = reinterpret_cast<C2<T> *>(dst) + tc.db + (int64_t)c * P.tsout;
#pragma unroll
      for (int m = 0; m < E; ++m) {
        const int k = t + TT * m;
        const C2<T> a = cur[m];
        const C2<T> b = lds[lidx<NF, C, LM>((H - k) & (H - 1), c)];
        // Ev = (a + conj b)/2, O = -i (a - conj b)/2, X = Ev + W^k O
        const C2<T> Ev{(a.x + b.x) * (T)0.5, (a.y - b.y) * (T)0.5};
        const C2<T> O{(a.y + b.y) * (T)0.5, -(a.x - b.x) * (T)0.5};
        const int ko = k - zout;
        if (ko >= 0 && ko < nout) {
          T f = msign<T>(P.omod, ko);
          if (P.hw == 2 && k != 0) f *= (T)2;
          const C2<T> WO = cmul(O, rtw[k]);
          st_stream(d + (int64_t)ko * P.sout, C2<T>{(Ev.x + WO.x) * f, (Ev.y + WO.y) * f});
        }
        if (k == 0) {  // Nyquist X_h = E_0 - O_0 (W^h = -1)
          const int kn = H - zout;
          if (kn >= 0 && kn < nout) {
            const T f = msign<T>(P.omod, kn);
            d[(int64_t)kn * P.sout] = C2<T>{(Ev.x - O.x) * f, (Ev.y - O.y) * f};
          }
        }
      }
    }
    if (!more) return false;
    line_sync<wave_lines<NF, E, LM>()>();
    item = ni;
    tc = tn;
    return true;
  };
  C2<T> va[E], vb[E];
  load_r2c<T, NF, C, E>(P, src, tc, true, c0, t0, va);
  if constexpr (TWL) lds_barrier();
  while (step(va, vb) && step(vb, va)) {
  }
}

// ------------------------------------------------------------------ c2r
// half spectrum (embedded at zin) -> real line of n = 2*NF contiguous
// samples (sout == 1, no trunc).
template <typename T, int NF, int C, int E>
__device__ __forceinline__ void load_c2r(const pfk_pass &P, const T *src, const TileCtx &tc,
                                         const bool live, const int c, const int t, C2<T> *v,
                                         C2<T> &xh) {
  load_c2c<T, NF, C, E>(P, src, tc, live, c, t, v);
  const int jh = NF - (int)P.zin;
  const bool okh = jh >= 0 && jh < (int)P.nin;
  const C2<T> *s0 = reinterpret_cast<const C2<T> *>(src) + tc.sb;
  const C2<T> *s = (live && c < tc.nl) ? s0 + (int64_t)c * P.tsin : s0;
  xh = s[(int64_t)((live && okh) ? jh : 0) * P.sin];
}

// plain half spectrum (all n/2+1 bins, no modulation): loads with uniform
// strides
template <typename T, int NF, int C, int E>
__device__ __forceinline__ void load_c2r_pl(const pfk_pass &P, const T *src, const TileCtx &tc,
                                            const bool live, const int c, const int t, C2<T> *v,
                                            C2<T> &xh) {
  load_c2c_pl<T, NF, C, E>(P, src, tc, live, c, t, v);
  const C2<T> *s0 = reinterpret_cast<const C2<T> *>(src) + tc.sb;
  const C2<T> *s = (live && c < tc.nl) ? s0 + (int64_t)c * P.tsin : s0;
  xh = ld_stream(s + (live ? (int64_t)NF * P.sin : 0));
}

template <typename T, int NF, int C, int E, int SIGN, bool PL = false, bool LM = false>
__global__ void __launch_bounds__(C *NF / E)
k_c2r_pow2(const pfk_pass P, const int64_t item0, const int64_t item_end, const int64_t ntiles,
           const T *__restrict__ src, T *__restrict__ dst) {
  constexpr int TT = NF / E;
  constexpr int H = NF;
  constexpr int SLOTS = LM ? lds_slots<NF, C, true>() : (NF + 1) * C;   // with the Nyquist bin
  constexpr bool TWL = tw_in_lds<T>(SLOTS, NF, NF);
  __shared__ C2<T> lds[SLOTS];
  __shared__ C2<T> twl[TWL ? NF : 1], rtwl[TWL ? NF : 1];
  const int tid = threadIdx.x;
  const int c0 = LM ? tid / TT : tid % C, t0 = LM ? tid % TT : tid / C;
  const C2<T> *twg = reinterpret_cast<const C2<T> *>(P.tw);
  const C2<T> *rtwg = reinterpret_cast<const C2<T> *>(P.rtw);
  if constexpr (TWL) {
    stage_table(twl, twg, NF);
    stage_table(rtwl, rtwg, NF);
  }
  const C2<T> *tw = TWL ? twl : twg;
  const C2<T> *rtw = TWL ? rtwl : rtwg;
  const int zin = (int)P.zin, nin = (int)P.nin;
  int64_t item = item0 + blockIdx.x;
  TileCtx tc = tile_at(P, C, item, ntiles);
  auto step = [&](C2<T> *cur, C2<T> &curh, C2<T> *nxt, C2<T> &nxth) -> bool {
    int t = t0, c = c0;
    asm volatile("" : "+v"(t), "+v"(c));
    const int64_t ni = item + gridDim.x;
    const bool more = ni < item_end;
    const TileCtx tn = tile_at(P, C, more ? ni : item, ntiles);
    if constexpr (PL) load_c2r_pl<T, NF, C, E>(P, src, tn, more, c, t, nxt, nxth);
    else load_c2r<T, NF, C, E>(P, src, tn, more, c, t, nxt, nxth);
    const bool valid = c < tc.nl;
    // X_k, k = t + T*m, and X_h (kept by thread t == 0) into LDS
    if constexpr (PL) {
#pragma unroll
      for (int m = 0; m < E; ++m) lds[lidx<NF, C, LM>(t + TT * m, c)] = cur[m];
      if (t == 0) lds[lidx<NF, C, LM>(H, c)] = C2<T>{curh.x, (T)0};
    } else {
#pragma unroll
      for (int m = 0; m < E; ++m) {
        const int k = t + TT * m;
        const int j = k - zin;
        const bool ok = valid && j >= 0 && j < nin;
        T f = msign<T>(P.imod, j);
        if (P.hw == 1 && k != 0) f *= (T)0.5;
        cur[m] = ok ? C2<T>{cur[m].x * f, cur[m].y * f} : C2<T>{0, 0};
        lds[lidx<NF, C, LM>(k, c)] = cur[m];
      }
      if (t == 0) {
        const int jh = H - zin;
        const bool okh = valid && jh >= 0 && jh < nin;
        lds[lidx<NF, C, LM>(H, c)] = okh ? C2<T>{curh.x * msign<T>(P.imod, jh), (T)0} : C2<T>{0, 0};
      }
    }
    line_sync<wave_lines<NF, E, LM>()>();
    // Z_k = E_k + i O_k
#pragma unroll
    for (int m = 0; m < E; ++m) {
      const int k = t + TT * m;
      const C2<T> a = cur[m];
      C2<T> z;
      if (k == 0) {
        const T xhr = lds[lidx<NF, C, LM>(H, c)].x;
        z = C2<T>{a.x + xhr, a.x - xhr};
      } else {
        const C2<T> b = lds[lidx<NF, C, LM>(H - k, c)];
        const C2<T> Ev{a.x + b.x, a.y - b.y};
        const C2<T> O = cmul(C2<T>{a.x - b.x, a.y + b.y}, rtw[k]);
        z = C2<T>{Ev.x - O.y, Ev.y + O.x};
      }
      cur[m] = z;
    }
    line_sync<wave_lines<NF, E, LM>()>();
    fft_regs<T, NF, C, E, SIGN, LM>(cur, lds, t, c, tw);
    if (valid) {
      C2<T> *d = reinterpret_cast<C2<T> *>(dst + tc.db + (int64_t)c * P.tsout);
#pragma unroll
      for (int m = 0; m < E; ++m) {
        const int p = t + TT * m;
        C2<T> x = cur[m];
        if (!PL && P.omod.on) {
          x.x *= msign<T>(P.omod, 2 * p);
          x.y *= msign<T>(P.omod, 2 * p + 1);
        }
        st_stream(d + p, x);
      }
    }
    if (!more) return false;
    line_sync<wave_lines<NF, E, LM>()>();
    item = ni;
    tc = tn;
    return true;
  };
  C2<T> va[E], vb[E], ha, hb;
  if constexpr (PL) load_c2r_pl<T, NF, C, E>(P, src, tc, true, c0, t0, va, ha);
  else load_c2r<T, NF, C, E>(P, src, tc, true, c0, t0, va, ha);
  if constexpr (TWL) lds_barrier();
  while (step(va, ha, vb, hb) && step(vb, hb, va, ha)) {
  }
}

// ------------------------------------------------------------------ dispatch
// Persistent grid: resident workgroups per CU (occupancy of the kernel at
// its LDS size) times the CU count, a multiple of 8 (the XCD count).
template <typename K>
static int64_t persistent_grid(K kernel, int threads) {
  int dev = 0, ncu = 0, occ = 0;
  (void)hipGetDevice(&dev);
  (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kernel, threads, 0) != hipSuccess ||
      occ < 1)
    occ = 1;
  if (ncu < 8) ncu = 8;
  return (int64_t)occ * (ncu / 8) * 8;
}

// Launches kernel k over `ntiles` items: persistent (grid = resident
// workgroups, `pgrid`) or one workgroup per tile (in chunks of 2^30).
template <typename T>
static int launch_items(void (*k)(pfk_pass, int64_t, int64_t, int64_t, const T *, T *),
                        const pfk_pass *p, int64_t ntiles, int64_t pgrid, bool persistent,
                        int threads, const T *src, T *dst) {
  if (persistent) {
    const int64_t g = ntiles < pgrid ? ntiles : pgrid;
    hipLaunchKernelGGL(k, dim3((unsigned)g), dim3(threads), 0, g_stream, *p, (int64_t)0, ntiles,
                       ntiles, src, dst);
  } else {
    const int64_t chunk = pfk_max_grid(threads);
    for (int64_t base = 0; base < ntiles; base += chunk) {
      const int64_t nblk = (ntiles - base) < chunk ? (ntiles - base) : chunk;
      hipLaunchKernelGGL(k, dim3((unsigned)nblk), dim3(threads), 0, g_stream, *p, base,
                         base + nblk, ntiles, src, dst);
    }
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) { set_error("pfk_launch_pass(pow2)", e); return -1; }
  return 0;
}

// A pass with every input and output position present and no modulation
// or adjoint weights runs the plain (PL) instantiations.
static bool pass_plain(const pfk_pass *p) {
  if (p->zin || p->zout || p->imod.on || p->omod.on || p->hw || p->twn) return false;
  if (p->kind == PFK_C2C) return p->nin == p->nf && p->nout == p->nf;
  if (p->kind == PFK_R2C) return p->nin == p->n && p->nout == p->nf + 1;
  if (p->kind == PFK_C2R) return p->nin == p->nf + 1 && p->nout == p->n;
  return false;
}

// Paired-line stores (the PAIR parameter of k_c2c_pow2) on the store side
// (or loads, `in`): lines adjacent in memory (tile stride 1), even element
// strides and a 16-byte-aligned array, so that every pair of lines starts
// on a 16-byte boundary; at least two lines.
static bool pair_side_ok(const pfk_pass *p, bool in, const void *ptr) {
  const int64_t ts = in ? p->tsin : p->tsout, ls = in ? p->sin : p->sout;
  if (p->prec != PFK_F32 || ts != 1 || p->text < 2 || (ls & 1) || ((uintptr_t)ptr & 15))
    return false;
  for (int i = 0; i < p->nb; ++i)
    if ((in ? p->bsin[i] : p->bsout[i]) & 1) return false;
  return true;
}

// Line-major staging of a complex64 c2c pass (the XS parameter of
// k_c2c_pow2): 1 when the lines are contiguous only on the store side and
// adjacent along the tile dimension on the load side, 2 for the reverse.
static int lm_staging(const pfk_pass *p) {
  if (p->sin != 1 && p->tsin == 1 && p->sout == 1) return 1;
  if (p->sin == 1 && p->sout != 1 && p->tsout == 1) return 2;
  return 0;
}

template <typename T, int NF, int C, int E, int SIGN, bool TWOK, bool PL>
static int launch_kind_pl(const pfk_pass *p, int64_t ntiles, const void *src, void *dst) {
  constexpr int NT = C * NF / E;
  static int64_t grid_c2c = 0, grid_lm = 0, grid_lm1 = 0, grid_lm2 = 0, grid_r2c = 0,
                 grid_c2r = 0, grid_tw = 0, grid_r2c_lm = 0, grid_c2r_lm = 0, grid_lm2p = 0;
  const bool persistent = !(p->variant & 4);
  const T *s = (const T *)src;
  T *d = (T *)dst;
  if (p->kind == PFK_C2C && p->twn) {
    if constexpr (TWOK && !PL) {
      if (!grid_tw) grid_tw = persistent_grid(k_c2c_pow2<T, NF, C, E, SIGN, false, true>, NT);
      return launch_items(k_c2c_pow2<T, NF, C, E, SIGN, false, true>, p, ntiles, grid_tw,
                          persistent, NT, s, d);
    }
    return 1;
  }
  // line-major lane order (variant bit 8): complex64, and complex128 on
  // plain passes
  constexpr bool LMOK = sizeof(T) == 4 || PL;
  const bool lm = LMOK && (p->variant & 8);
  if (p->kind == PFK_C2C) {
    if constexpr (LMOK) {
      if (lm) {
        const int xs = PL ? lm_staging(p) : 0;
        if constexpr (PL && sizeof(T) == 4) {
          // paired-line stores on the tile side of a transposing pass (512-point
          // complex64 lines: 5.10 -> 5.27 TB/s); paired loads measured no
          // faster there and slower on passes that are tile-adjacent on both
          // sides (5.0 -> 4.7 TB/s), so those keep 8-byte lanes
          if (xs == 2 && pair_side_ok(p, false, dst)) {
            if (!grid_lm2p)
              grid_lm2p =
                  persistent_grid(k_c2c_pow2<T, NF, C, E, SIGN, true, false, true, 2, 1>, NT);
            return launch_items(k_c2c_pow2<T, NF, C, E, SIGN, true, false, true, 2, 1>, p, ntiles,
                                grid_lm2p, persistent, NT, s, d);
          }
        }
        if constexpr (PL) {
          if (xs == 1) {
            if (!grid_lm1)
              grid_lm1 = persistent_grid(k_c2c_pow2<T, NF, C, E, SIGN, true, false, true, 1>, NT);
            return launch_items(k_c2c_pow2<T, NF, C, E, SIGN, true, false, true, 1>, p, ntiles,
                                grid_lm1, persistent, NT, s, d);
          }
          if (xs == 2) {
            if (!grid_lm2)
              grid_lm2 = persistent_grid(k_c2c_pow2<T, NF, C, E, SIGN, true, false, true, 2>, NT);
            return launch_items(k_c2c_pow2<T, NF, C, E, SIGN, true, false, true, 2>, p, ntiles,
                                grid_lm2, persistent, NT, s, d);
          }
        }
        if (!grid_lm) grid_lm = persistent_grid(k_c2c_pow2<T, NF, C, E, SIGN, true, false, PL>, NT);
        return launch_items(k_c2c_pow2<T, NF, C, E, SIGN, true, false, PL>, p, ntiles, grid_lm,
                            persistent, NT, s, d);
      }
    }
    if (!grid_c2c) grid_c2c = persistent_grid(k_c2c_pow2<T, NF, C, E, SIGN, false, false, PL>, NT);
    return launch_items(k_c2c_pow2<T, NF, C, E, SIGN, false, false, PL>, p, ntiles, grid_c2c,
                        persistent, NT, s, d);
  } else if (p->kind == PFK_R2C) {
    if constexpr (LMOK) {
      if (lm) {
        if (!grid_r2c_lm) grid_r2c_lm = persistent_grid(k_r2c_pow2<T, NF, C, E, SIGN, PL, true>, NT);
        return launch_items(k_r2c_pow2<T, NF, C, E, SIGN, PL, true>, p, ntiles, grid_r2c_lm,
                            persistent, NT, s, d);
      }
    }
    if (!grid_r2c) grid_r2c = persistent_grid(k_r2c_pow2<T, NF, C, E, SIGN, PL>, NT);
    return launch_items(k_r2c_pow2<T, NF, C, E, SIGN, PL>, p, ntiles, grid_r2c, persistent, NT,
                        s, d);
  }
  if constexpr (LMOK) {
    if (lm) {
      if (!grid_c2r_lm) grid_c2r_lm = persistent_grid(k_c2r_pow2<T, NF, C, E, SIGN, PL, true>, NT);
      return launch_items(k_c2r_pow2<T, NF, C, E, SIGN, PL, true>, p, ntiles, grid_c2r_lm,
                          persistent, NT, s, d);
    }
  }
  if (!grid_c2r) grid_c2r = persistent_grid(k_c2r_pow2<T, NF, C, E, SIGN, PL>, NT);
  return launch_items(k_c2r_pow2<T, NF, C, E, SIGN, PL>, p, ntiles, grid_c2r, persistent, NT, s,
                      d);
}

// the plain instantiations exist for the default tiles only
template <typename T, int NF, int C, int E, int SIGN, bool TWOK>
static int launch_kind(const pfk_pass *p, int64_t ntiles, const void *src, void *dst) {
  if constexpr (TWOK)
    if (pass_plain(p)) return launch_kind_pl<T, NF, C, E, SIGN, true, true>(p, ntiles, src, dst);
  return launch_kind_pl<T, NF, C, E, SIGN, TWOK, false>(p, ntiles, src, dst);
}

// TWOK: the table also instantiates the twiddled first step of long lines
template <typename T, int NF, int C, int E, bool TWOK = true>
static int launch_nf(const pfk_pass *p, const void *src, void *dst) {
  int64_t ntiles = (p->text + C - 1) / C;
  for (int i = 0; i < p->nb; ++i) ntiles *= p->bext[i];
  if (ntiles <= 0) return 0;
  return p->sign < 0 ? launch_kind<T, NF, C, E, -1, TWOK>(p, ntiles, src, dst)
                     : launch_kind<T, NF, C, E, +1, TWOK>(p, ntiles, src, dst);
}


// Tile width C (consecutive lines per workgroup) gives 128-byte segments:
// C = 8 for 16-byte FP64 complex, C = 16 for 8-byte FP32 complex (C = 4 / 8
// at the longest lengths, where the tile would exceed 128 KiB of LDS).
// E = 16 elements per thread from NF = 1024 on keeps the workgroup at
// <= 1024 threads; complex64 takes E = 16 from NF = 512 (512 threads in
// two radix-16 stages and a radix-2 one, 95-117 VGPRs without spills, two
// workgroups per CU: its column passes of the 512^3 pair 207-212 -> 200 us
// against E = 8 at 1024 threads).
static int launch_pow2_f64(const pfk_pass *p, const void *src, void *dst) {
  switch (p->nf) {
    case 64: return launch_nf<double, 64, 8, 8>(p, src, dst);
    case 128: return launch_nf<double, 128, 8, 8>(p, src, dst);
    case 256: return launch_nf<double, 256, 8, 8>(p, src, dst);
    case 512: return launch_nf<double, 512, 8, 8>(p, src, dst);
    case 1024: return launch_nf<double, 1024, 8, 16>(p, src, dst);
    case 2048: return launch_nf<double, 2048, 4, 16>(p, src, dst);
    case 4096: return launch_nf<double, 4096, 2, 16>(p, src, dst);
    default: return 1;
  }
}

// alternative (narrower) tiles: half the lines per workgroup, half the LDS
static int launch_pow2_f64_alt(const pfk_pass *p, const void *src, void *dst) {
  switch (p->nf) {
    case 64: return launch_nf<double, 64, 4, 8, false>(p, src, dst);
    case 128: return launch_nf<double, 128, 4, 8, false>(p, src, dst);
    case 256: return launch_nf<double, 256, 4, 8, false>(p, src, dst);
    case 512: return launch_nf<double, 512, 4, 8, false>(p, src, dst);
    case 1024: return launch_nf<double, 1024, 4, 16, false>(p, src, dst);
    case 2048: return launch_nf<double, 2048, 2, 16, false>(p, src, dst);
    case 4096: return launch_nf<double, 4096, 1, 16, false>(p, src, dst);
    default: return 1;
  }
}

static int launch_pow2_f32(const pfk_pass *p, const void *src, void *dst) {
  switch (p->nf) {
    case 64: return launch_nf<float, 64, 16, 8>(p, src, dst);
    case 128: return launch_nf<float, 128, 16, 8>(p, src, dst);
    case 256: return launch_nf<float, 256, 16, 8>(p, src, dst);
    case 512: return launch_nf<float, 512, 16, 16>(p, src, dst);
    case 1024: return launch_nf<float, 1024, 16, 16>(p, src, dst);
    case 2048: return launch_nf<float, 2048, 8, 16>(p, src, dst);
    case 4096: return launch_nf<float, 4096, 4, 16>(p, src, dst);
    default: return 1;
  }
}

static int launch_pow2_f32_alt(const pfk_pass *p, const void *src, void *dst) {
  switch (p->nf) {
    case 64: return launch_nf<float, 64, 8, 8, false>(p, src, dst);
    case 128: return launch_nf<float, 128, 8, 8, false>(p, src, dst);
    case 256: return launch_nf<float, 256, 8, 8, false>(p, src, dst);
    case 512: return launch_nf<float, 512, 8, 8, false>(p, src, dst);
    case 1024: return launch_nf<float, 1024, 8, 16, false>(p, src, dst);
    case 2048: return launch_nf<float, 2048, 4, 16, false>(p, src, dst);
    case 4096: return launch_nf<float, 4096, 2, 16, false>(p, src, dst);
    default: return 1;
  }
}

int fast_covers(const pfk_pass *p) {
  if (p->n1 || p->dec_in || p->dec_out) return 0;   // long lines, decimated steps
  if (p->twn && (p->kind != PFK_C2C || (p->variant & ~4))) return 0;   // default tiles only
  if (p->kind != PFK_C2C && p->kind != PFK_R2C && p->kind != PFK_C2R) return 0;
  const int64_t nf = p->nf;
  if (nf < 64 || nf > 4096 || (nf & (nf - 1))) return 0;
  if (p->kind == PFK_R2C) {
    // real input: contiguous, not embedded
    if (p->n != 2 * nf || p->sin != 1 || p->zin != 0 || p->nin != p->n) return 0;
    if (p->hw == 1) return 0;
  }
  if (p->kind == PFK_C2R) {
    if (p->n != 2 * nf || p->sout != 1 || p->zout != 0 || p->nout != p->n) return 0;
    if (p->hw == 2) return 0;
  }
  if (p->kind == PFK_C2C && p->hw) return 0;
  return 1;
}

int launch_fast(const pfk_pass *p, const void *src, void *dst) {
  if (g_generic_only || p->variant == 1 || !fast_covers(p)) return 1;
  if ((p->variant & 3) == 2)
    return (p->prec == PFK_F64) ? launch_pow2_f64_alt(p, src, dst)
                                : launch_pow2_f32_alt(p, src, dst);
  return (p->prec == PFK_F64) ? launch_pow2_f64(p, src, dst) : launch_pow2_f32(p, src, dst);
}

}  // namespace pfk

extern "C" int pfk_pass_plain(const pfk_pass *p) { return pfk::pass_plain(p) ? 1 : 0; }

extern "C" int pfk_pass_kernels(const pfk_pass *p) {
  if (pfk::g_generic_only) return 0;
  int m = 0;
  if (pfk::mfma_covers(p)) m |= PFK_HAS_MFMA;
  if (pfk::fast_covers(p)) m |= PFK_HAS_POW2;
  if (pfk::small_covers(p)) m |= PFK_HAS_SMALL;
  return m;
}
