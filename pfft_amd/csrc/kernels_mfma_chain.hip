// kernels_mfma_chain.hip -- plain complex64 c2c passes of 256, 512 and
// 1024 points whose radix-16 stages run on the matrix cores (gfx950 MFMA),
// for plans created with PFFT_AMD_MFMA_F32 or PFFT_AMD_MFMA_BF16X2.
//
// Radix-16 stage as a K = 16 product.  Y_k = sum_p x_p W16^(pk) with
// W16^(pk) = c_pk + i s_pk.  Points p and 16-p share the cosine and have
// opposite sines, so with u = Re x_p + Re x_16-p, v = Im x_p - Im x_16-p,
// w = Im x_p + Im x_16-p, z = Re x_p - Re x_16-p (p = 1..7) and the
// self-paired points 0 and 8 taken as e = x_0 + x_8 (even k) and
// d = x_0 - x_8 (odd k):
//   Re Y_k = sum_u c u - sum_v s v + [k even] Re e + [k odd] Re d
//   Im Y_k = sum_w c w + sum_z s z + [k even] Im e + [k odd] Im d,
// two real 16x16 products with K = 16 instead of the K = 32 of the plain
// complex product: 8 v_mfma_f32_16x16x4f32 per 16 butterflies, or 6
// v_mfma_f32_16x16x16bf16_1k for split BF16 (hi + lo operands, three
// products, FP32 accumulation).
//
// Slot layout.  Lane l = 16 q + j of a wave supplies, in K step s, the
// operand of slot (q, s); the lane holds the four points P(q, 0..3) of its
// row group q: (4, 0, 12, 8), (1, 2, 15, 14), (3, 5, 13, 11), (6, 7, 10, 9),
// i.e. two points p, p' and their mirrors, so every slot is formed from the
// lane's own registers.  Group 0 holds the pair (4, 12) and the two
// self-paired points; its fourth slot is Re d (for Re Y) / Im d (for Im Y)
// where the other groups have the v / z difference of the second pair,
// one select.  The coefficient lane (q, k) holds the matching entries for
// output frequency k.  Outputs: lane (q, j) receives rows 4q..4q+3.
//
// Tile-major form (strided lines, k_mfma_tile).  A workgroup transforms
// 16 lines (the 16 MFMA columns; 128-byte segments of complex64) of
// NF = 16 Q points, n = n1 + Q n2.  Stage 1: wave w, group g takes
// n1 = 4w + g and the DFT-16 over n2 on the matrix cores; the outputs
// k2 = 4q + v are multiplied by W_NF^(n1 k2) and written to LDS
// [n1][k2][line] (the one exchange of the pass).  Stage 2: the Q-point DFT
// over n1 = a + R b (R = Q/16): a DFT-16 over b on the matrix cores for
// each (a, k2), the twiddle W_Q^(a kb), and a radix-R DFT over a inside
// the lane; X[k2 + 16 kb + 256 ka] is stored straight from the registers
// (16 lines x 8 bytes per row of the wave).
//
// Line-major form (contiguous lines, k_mfma_line).  A wave transforms
// lines of NF = R 256 points without an LDS exchange: for each of the R
// decimated sub-transforms of 256 points two chained radix-16 products,
// then a radix-R step (DIT) inside the lanes.  The data is the A operand
// (rows = butterflies): the product's output rows 4q..4q+3 of lane (q, k2)
// are the butterflies P(q, 0..3) of the first stage (rows are loaded in
// that order), so they are directly the four slot points of the second
// stage.  Y_j[k2 + 16 k1] sits in lane (q, k1), row k2 = 4q + v; a
// per-wave LDS row turns the outputs into stores of 64 consecutive
// elements per instruction.
//
// Per-stage scaling of the split form: the DFT coefficients are multiplied
// by 1/4 (exact) so magnitudes stay O(1); the store multiplies by 16.
#include <stdint.h>
#include "pfk_internal.hpp"

namespace pfk {
namespace {

typedef float v4f __attribute__((ext_vector_type(4)));
typedef float v2f __attribute__((ext_vector_type(2)));
typedef short v4s __attribute__((ext_vector_type(4)));
typedef __bf16 v4bf __attribute__((ext_vector_type(4)));

struct cf {
  float x, y;
};

__device__ __forceinline__ cf cmf(cf a, cf b) {
  return cf{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
__device__ __forceinline__ cf ldnt(const cf *p) {
  const v2f v = __builtin_nontemporal_load(reinterpret_cast<const v2f *>(p));
  return cf{v.x, v.y};
}
__device__ __forceinline__ void stnt(cf *p, cf a) {
  v2f v;
  v.x = a.x;
  v.y = a.y;
  __builtin_nontemporal_store(v, reinterpret_cast<v2f *>(p));
}

// point of slot (q, s): nibble q of one constant per s
__device__ __forceinline__ int ptab(int q, int s) {
  const int t = s == 0 ? 0x6314 : s == 1 ? 0x7520 : s == 2 ? 0xADFC : 0x9BE8;
  return (t >> (4 * q)) & 15;
}

// cos(2 pi m / 16)
__constant__ float kcos16[16] = {1.f, 0.92387953251128675613f, 0.70710678118654752440f,
                                 0.38268343236508977173f, 0.f, -0.38268343236508977173f,
                                 -0.70710678118654752440f, -0.92387953251128675613f, -1.f,
                                 -0.92387953251128675613f, -0.70710678118654752440f,
                                 -0.38268343236508977173f, 0.f, 0.38268343236508977173f,
                                 0.70710678118654752440f, 0.92387953251128675613f};

// hi + lo split of four floats into BF16 operands: hi = bf16(x), lo =
// bf16(x - hi), both rounded to nearest even (the conversion is the
// packed v_cvt_pk_bf16_f32 of gfx950: 3 instructions per float)
__device__ __forceinline__ void split_bf16(v4f x, v4s &hi, v4s &lo) {
  const v4bf h = __builtin_convertvector(x, v4bf);
  const v4f back = __builtin_convertvector(h, v4f);
  hi = __builtin_bit_cast(v4s, h);
  lo = __builtin_bit_cast(v4s, __builtin_convertvector(x - back, v4bf));
}

// DFT-16 coefficients of this lane (row group q, output frequency k):
// FP32 form cr/ci[s]; split-BF16 form hi/lo vectors (scaled by 1/4)
template <int MODE, int SIGN>
struct Coef {
  float cr[4], ci[4];
  v4s rh, rl, ih, il;
  __device__ __forceinline__ void init(int q, int k) {
    float r[4], i[4];
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      if (s & 1 && q == 0) {
        const float e = (k & 1) ? 0.f : 1.f;
        r[s] = i[s] = (s == 1) ? e : 1.f - e;
      } else {
        const int m = (ptab(q, s & 1) * k) & 15;
        const float c = kcos16[m], sn = (float)SIGN * kcos16[(m + 12) & 15];
        r[s] = s < 2 ? c : -sn;
        i[s] = s < 2 ? c : sn;
      }
    }
    if (MODE == 0) {
#pragma unroll
      for (int s = 0; s < 4; ++s) { cr[s] = r[s]; ci[s] = i[s]; }
    } else {
      split_bf16(v4f{0.25f * r[0], 0.25f * r[1], 0.25f * r[2], 0.25f * r[3]}, rh, rl);
      split_bf16(v4f{0.25f * i[0], 0.25f * i[1], 0.25f * i[2], 0.25f * i[3]}, ih, il);
    }
  }
};

// N radix-16 products of 16 butterflies each, issued interleaved (2N
// independent accumulation chains keep the matrix pipeline busy while
// each chain waits for its previous step).  R[4n..4n+3]: this lane's
// points P(q, 0..3) of product n; q0: row group 0.  DATA_A: the data is
// the A operand (rows = butterflies, outputs by frequency in the columns),
// else the B operand (columns = butterflies, outputs by frequency in the
// rows).
template <int N, int MODE, int SIGN, bool DATA_A>
__device__ __forceinline__ void dft16_mfma(const Coef<MODE, SIGN> &K, const cf *R, bool q0,
                                           v4f *accr, v4f *acci) {
  float yr[N][4], yi[N][4];
#pragma unroll
  for (int n = 0; n < N; ++n) {
    const cf *r = R + 4 * n;
    yr[n][0] = r[0].x + r[2].x;
    yr[n][1] = r[1].x + r[3].x;
    yr[n][2] = r[0].y - r[2].y;
    yi[n][0] = r[0].y + r[2].y;
    yi[n][1] = r[1].y + r[3].y;
    yi[n][2] = r[0].x - r[2].x;
    const float drr = r[1].x - r[3].x, dii = r[1].y - r[3].y;
    yr[n][3] = q0 ? drr : dii;
    yi[n][3] = q0 ? dii : drr;
    accr[n] = v4f{0.f, 0.f, 0.f, 0.f};
    acci[n] = v4f{0.f, 0.f, 0.f, 0.f};
  }
  if constexpr (MODE == 0) {
#pragma unroll
    for (int s = 0; s < 4; ++s)
#pragma unroll
      for (int n = 0; n < N; ++n) {
        if constexpr (DATA_A) {
          accr[n] = __builtin_amdgcn_mfma_f32_16x16x4f32(yr[n][s], K.cr[s], accr[n], 0, 0, 0);
          acci[n] = __builtin_amdgcn_mfma_f32_16x16x4f32(yi[n][s], K.ci[s], acci[n], 0, 0, 0);
        } else {
          accr[n] = __builtin_amdgcn_mfma_f32_16x16x4f32(K.cr[s], yr[n][s], accr[n], 0, 0, 0);
          acci[n] = __builtin_amdgcn_mfma_f32_16x16x4f32(K.ci[s], yi[n][s], acci[n], 0, 0, 0);
        }
      }
  } else {
    v4s rh[N], rl[N], ih[N], il[N];
#pragma unroll
    for (int n = 0; n < N; ++n) {
      split_bf16(v4f{yr[n][0], yr[n][1], yr[n][2], yr[n][3]}, rh[n], rl[n]);
      split_bf16(v4f{yi[n][0], yi[n][1], yi[n][2], yi[n][3]}, ih[n], il[n]);
    }
#define PFK_BF(x, y, acc) acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(x, y, acc, 0, 0, 0)
#pragma unroll
    for (int n = 0; n < N; ++n) {
      if constexpr (DATA_A) { PFK_BF(rh[n], K.rh, accr[n]); PFK_BF(ih[n], K.ih, acci[n]); }
      else { PFK_BF(K.rh, rh[n], accr[n]); PFK_BF(K.ih, ih[n], acci[n]); }
    }
#pragma unroll
    for (int n = 0; n < N; ++n) {
      if constexpr (DATA_A) { PFK_BF(rl[n], K.rh, accr[n]); PFK_BF(il[n], K.ih, acci[n]); }
      else { PFK_BF(K.rh, rl[n], accr[n]); PFK_BF(K.ih, il[n], acci[n]); }
    }
#pragma unroll
    for (int n = 0; n < N; ++n) {
      if constexpr (DATA_A) { PFK_BF(rh[n], K.rl, accr[n]); PFK_BF(ih[n], K.il, acci[n]); }
      else { PFK_BF(K.rl, rh[n], accr[n]); PFK_BF(K.il, ih[n], acci[n]); }
    }
#undef PFK_BF
  }
}

template <int SIGN>
__device__ __forceinline__ cf mul_si(cf a) {   // a * SIGN i
  return SIGN < 0 ? cf{a.y, -a.x} : cf{-a.y, a.x};
}

struct Tile16 {
  int64_t sb, db;
  int nl;
};

// Tile `item` of CL consecutive lines along the tile dimension (the other
// pass kernels' decode).  Indices below 2^32 (every practical pass) take
// the 32-bit division chain, which is several times shorter than the
// 64-bit one.
__device__ __forceinline__ Tile16 tile_at_cl(const pfk_pass &P, int CL, int64_t item,
                                             int64_t ntiles) {
  int64_t tileid = item;
  if (P.xcd_remap) {
    const int64_t qn = ntiles >> 3, rn = ntiles & 7, x = item & 7, jx = item >> 3;
    tileid = x * qn + (x < rn ? x : rn) + jx;
  }
  const int64_t ntt = (P.text + CL - 1) / CL;
  Tile16 r;
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
  const int64_t c0 = tt * CL;
  r.nl = (int)((P.text - c0) < CL ? (P.text - c0) : CL);
  r.sb = sb + c0 * P.tsin;
  r.db = db + c0 * P.tsout;
  return r;
}

// ---------------------------------------------------------------- tile-major
template <int NF, int MODE, int SIGN>
__global__ void __launch_bounds__(NF)
k_mfma_tile(const pfk_pass P, const int64_t item0, const int64_t item_end, const int64_t ntiles,
            const float *__restrict__ src, float *__restrict__ dst) {
  constexpr int Q = NF / 16, R = Q / 16, KK = 4 / R;
  __shared__ cf ex[16 * NF];   // [n1][k2][line]
  __shared__ cf twl[NF];
  const int tid = threadIdx.x;
  twl[tid] = reinterpret_cast<const cf *>(P.tw)[tid];
  Coef<MODE, SIGN> K;
  K.init((tid & 63) >> 4, tid & 15);
  const cf *s0 = reinterpret_cast<const cf *>(src);
  cf *d0 = reinterpret_cast<cf *>(dst);

  int64_t item = item0 + blockIdx.x;
  Tile16 tc = tile_at_cl(P, 16, item, ntiles);
  // loads of tile `t`: lane (q, c) of wave w, group g, slot s holds
  // x[c][4w + g + Q P(q, s)]
  // live == false (no next tile): every lane reads the tile's first
  // element, so the last prefetch costs one segment instead of a tile
  auto load = [&](const Tile16 &t, int lt, cf *x, bool live) {
    const int c = lt & 15, q = (lt & 63) >> 4;
    const int w = __builtin_amdgcn_readfirstlane(lt >> 6);   // wave-uniform
    const int cc = (live && c < t.nl) ? c : 0;
    const int64_t sin = live ? P.sin : 0;
    const cf *b = s0 + t.sb + (int64_t)cc * P.tsin + (int64_t)(4 * w) * sin;
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      const cf *bs = b + (int64_t)(Q * ptab(q, s)) * sin;
#pragma unroll
      for (int g = 0; g < 4; ++g) x[4 * g + s] = ldnt(bs + (int64_t)g * sin);
    }
  };
  auto step = [&](cf *x, cf *nx) -> bool {
    int lt = tid;
    asm volatile("" : "+v"(lt));
    const int c = lt & 15, q = (lt & 63) >> 4;
    const int w = __builtin_amdgcn_readfirstlane(lt >> 6);
    const bool q0 = q == 0;
    const int64_t ni = item + gridDim.x;
    const bool more = ni < item_end;
    const Tile16 tn = tile_at_cl(P, 16, more ? ni : item, ntiles);
    load(tn, lt, nx, more);
    __builtin_amdgcn_sched_barrier(0);   // the prefetch stays ahead of the transform
    // stage 1: DFT-16 over n2, twiddle W_NF^(n1 k2), to LDS
    // (the split form in two batches of two: its operand splits need the
    // registers of a batch of four)
    constexpr int NB = MODE == 1 ? 2 : 4;
#pragma unroll
    for (int g0 = 0; g0 < 4; g0 += NB) {
      v4f ar[NB], ai[NB];
      dft16_mfma<NB, MODE, SIGN, false>(K, x + 4 * g0, q0, ar, ai);
#pragma unroll
      for (int gg = 0; gg < NB; ++gg) {
        const int g = g0 + gg, n1 = 4 * w + g;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int k2 = 4 * q + v;
          const cf y = cmf(cf{ar[gg][v], ai[gg][v]}, twl[n1 * k2]);
          ex[(n1 * 16 + k2) * 16 + c] = y;
        }
      }
    }
    lds_barrier();
    // stage 2: DFT-16 over b (n1 = a + R b) for the wave's (a, k2) pairs
    cf res[4][4];
    {
      cf pts[16];
#pragma unroll
      for (int g = 0; g < 4; ++g) {
        const int k2 = KK * w + g / R, a = g % R;
#pragma unroll
        for (int s = 0; s < 4; ++s) pts[4 * g + s] = ex[((a + R * ptab(q, s)) * 16 + k2) * 16 + c];
      }
#pragma unroll
      for (int g0 = 0; g0 < 4; g0 += NB) {
        v4f ar[NB], ai[NB];
        dft16_mfma<NB, MODE, SIGN, false>(K, pts + 4 * g0, q0, ar, ai);
#pragma unroll
        for (int gg = 0; gg < NB; ++gg) {
          const int g = g0 + gg, a = g % R;
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            cf y{ar[gg][v], ai[gg][v]};
            if (a) y = cmf(y, twl[16 * a * (4 * q + v)]);
            res[g][v] = y;
          }
        }
      }
    }
    // radix-R over a inside the lane; store X[k2 + 16 kb + 256 ka]
    if (c < tc.nl) {
      const float un = MODE == 1 ? 16.f : 1.f;
      cf *db = d0 + tc.db + (int64_t)c * P.tsout + (int64_t)(64 * q) * P.sout;
#pragma unroll
      for (int kk = 0; kk < KK; ++kk) {
        const int k2 = KK * w + kk;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          cf o[R];
          if constexpr (R == 1) {
            o[0] = res[kk][v];
          } else if constexpr (R == 2) {
            const cf e = res[2 * kk][v], f = res[2 * kk + 1][v];
            o[0] = cf{e.x + f.x, e.y + f.y};
            o[1] = cf{e.x - f.x, e.y - f.y};
          } else {
            const cf a0 = res[0][v], a1 = res[1][v], a2 = res[2][v], a3 = res[3][v];
            const cf t0{a0.x + a2.x, a0.y + a2.y}, t1{a0.x - a2.x, a0.y - a2.y};
            const cf t2{a1.x + a3.x, a1.y + a3.y};
            const cf t3 = mul_si<SIGN>(cf{a1.x - a3.x, a1.y - a3.y});
            o[0] = cf{t0.x + t2.x, t0.y + t2.y};
            o[2] = cf{t0.x - t2.x, t0.y - t2.y};
            o[1] = cf{t1.x + t3.x, t1.y + t3.y};
            o[3] = cf{t1.x - t3.x, t1.y - t3.y};
          }
#pragma unroll
          for (int ka = 0; ka < R; ++ka)
            stnt(db + (int64_t)(k2 + 16 * v + 256 * ka) * P.sout,
                 cf{o[ka].x * un, o[ka].y * un});
        }
      }
    }
    if (!more) return false;
    lds_barrier();   // the next tile's stage 1 rewrites the exchange buffer
    item = ni;
    tc = tn;
    return true;
  };
  cf xa[16], xb[16];
  load(tc, tid, xa, true);
  lds_barrier();   // twiddle table
  while (step(xa, xb) && step(xb, xa)) {
  }
}

// ---------------------------------------------------------------- line-major
// Lines with sin = sout = 1.  A workgroup (4 waves) takes tiles of
// CL = 4 * LPW consecutive lines along the tile dimension; wave w
// transforms lines w*LPW .. w*LPW+LPW-1 of the tile, NP = LPW * R radix-16
// products per stage (NP = 2 for 256 and 512 points, 4 for 1024: a lane
// holds 4 NP complex values), and the loads of the workgroup's next tile
// are issued before the current one is transformed (two register buffers
// of 4 NP values; with NP = 4 at 256 or 512 points the pair no longer fits
// the register budget of four waves per SIMD and spills).  No LDS
// exchange besides the twiddle table and the per-wave store row, no
// workgroup barrier in the loop.
//
// Decimation in time: sub-transform j (j < R) is the 256-point DFT of
// x[R n' + j], so the R values a lane needs at one n' are adjacent in
// memory (one 16-byte load for R = 2, two for R = 4), and the radix-R
// step runs after the products: X[k + 256 m] = sum_j W_NF^(jk) Y_j[k]
// W_R^(jm), with Y_j[k] at the same register position of every j.
template <int NF>
struct LineGeo {
  static constexpr int R = NF / 256, NP = NF == 1024 ? 4 : 2, LPW = NP / R, CL = 4 * LPW;
};

template <int R>
__device__ __forceinline__ void load_run(const cf *p, cf *x) {   // x[4 j] = p[j], j < R
  if constexpr (R == 1) {
    x[0] = ldnt(p);
  } else {
#pragma unroll
    for (int h = 0; h < R / 2; ++h) {
      const v4f v = __builtin_nontemporal_load(reinterpret_cast<const v4f *>(p) + h);
      x[4 * (2 * h)] = cf{v.x, v.y};
      x[4 * (2 * h + 1)] = cf{v.z, v.w};
    }
  }
}

template <int NF, int MODE, int SIGN>
__global__ void __launch_bounds__(256, 4)
k_mfma_line(const pfk_pass P, const int64_t item0, const int64_t item_end, const int64_t ntiles,
            const float *__restrict__ src, float *__restrict__ dst) {
  constexpr int R = LineGeo<NF>::R, NP = LineGeo<NF>::NP, LPW = LineGeo<NF>::LPW;
  constexpr int CL = LineGeo<NF>::CL;
  constexpr int SLOTS = NF + NF / 16;   // one line, one pad slot per 16 positions
  __shared__ cf twl[NF];
  __shared__ cf lx[4][SLOTS];           // per-wave transpose of a finished line
  const int tid = threadIdx.x;
  for (int i = tid; i < NF; i += 256) twl[i] = reinterpret_cast<const cf *>(P.tw)[i];
  Coef<MODE, SIGN> K;
  K.init((tid & 63) >> 4, tid & 15);
  const cf *s0 = reinterpret_cast<const cf *>(src);
  cf *d0 = reinterpret_cast<cf *>(dst);
  int64_t item = item0 + blockIdx.x;
  Tile16 tc = tile_at_cl(P, CL, item, ntiles);
  // lane (q, j16): first-stage butterfly rho = P(j16/4, j16%4); slot u of
  // sub-transform j of line i is x[R (rho + 16 P(q, u)) + j], held in
  // x[(i R + j) 4 + u]
  // live == false (no next tile): every lane reads the tile's first elements
  auto load = [&](const Tile16 &t, int lt, cf *x, bool live) {
    const int q = (lt & 63) >> 4, j16 = lt & 15;
    const int w = __builtin_amdgcn_readfirstlane(lt >> 6);
    const int rho = live ? ptab(j16 >> 2, j16 & 3) : 0;
#pragma unroll
    for (int i = 0; i < LPW; ++i) {
      const int ln = w * LPW + i;
      const cf *s = s0 + t.sb + (int64_t)((live && ln < t.nl) ? ln : 0) * P.tsin + R * rho;
#pragma unroll
      for (int u = 0; u < 4; ++u)
        load_run<R>(s + (live ? 16 * R * ptab(q, u) : 0), x + i * R * 4 + u);
    }
  };
  auto step = [&](cf *x, cf *nx) -> bool {
    int lt = tid;
    asm volatile("" : "+v"(lt));
    const int q = (lt & 63) >> 4, j16 = lt & 15;
    const int w = __builtin_amdgcn_readfirstlane(lt >> 6);
    const bool q0 = q == 0;
    const int64_t ni = item + gridDim.x;
    const bool more = ni < item_end;
    const Tile16 tn = tile_at_cl(P, CL, more ? ni : item, ntiles);
    load(tn, lt, nx, more);
    // keep the prefetch ahead of the transform (the scheduler would
    // otherwise sink the loads below it to save registers)
    __builtin_amdgcn_sched_barrier(0);
    const float un = MODE == 1 ? 16.f : 1.f;
    // the NP = LPW * R sub-transforms of 256 points together: first
    // products, twiddle W_256^(n1 k2) (rows 4q + v are butterflies
    // P(q, v), column k2 = j16), second products
    {
      v4f ar[NP], ai[NP];
      dft16_mfma<NP, MODE, SIGN, true>(K, x, q0, ar, ai);
#pragma unroll
      for (int n = 0; n < NP; ++n)
#pragma unroll
        for (int v = 0; v < 4; ++v)
          x[4 * n + v] = cmf(cf{ar[n][v], ai[n][v]}, twl[R * ptab(q, v) * j16]);
      dft16_mfma<NP, MODE, SIGN, true>(K, x, q0, ar, ai);
#pragma unroll
      for (int n = 0; n < NP; ++n)
#pragma unroll
        for (int v = 0; v < 4; ++v) x[4 * n + v] = cf{ar[n][v] * un, ai[n][v] * un};
    }
    // Y_j[k] (k = 4q + v + 16 j16) is x[(i R + j) 4 + v]; the radix-R step
    // leaves X[k + 256 m] in x[(i R + m) 4 + v]
    if constexpr (R > 1) {
#pragma unroll
      for (int i = 0; i < LPW; ++i) {
        cf *xl = x + i * R * 4;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int k = 4 * q + v + 16 * j16;
          if constexpr (R == 2) {
            const cf a = xl[v], b = cmf(xl[4 + v], twl[k]);
            xl[v] = cf{a.x + b.x, a.y + b.y};
            xl[4 + v] = cf{a.x - b.x, a.y - b.y};
          } else {
            const cf a0 = xl[v], a1 = cmf(xl[4 + v], twl[k]);
            const cf a2 = cmf(xl[8 + v], twl[2 * k]), a3 = cmf(xl[12 + v], twl[3 * k]);
            const cf t0{a0.x + a2.x, a0.y + a2.y}, t1{a0.x - a2.x, a0.y - a2.y};
            const cf t2{a1.x + a3.x, a1.y + a3.y};
            const cf t3 = mul_si<SIGN>(cf{a1.x - a3.x, a1.y - a3.y});
            xl[v] = cf{t0.x + t2.x, t0.y + t2.y};
            xl[8 + v] = cf{t0.x - t2.x, t0.y - t2.y};
            xl[4 + v] = cf{t1.x + t3.x, t1.y + t3.y};
            xl[12 + v] = cf{t1.x - t3.x, t1.y - t3.y};
          }
        }
      }
    }
    // each line goes through the wave's LDS row (slot p + p/16) and is
    // stored as 128 consecutive elements per instruction (two per lane)
    cf *row = lx[w];
    const int lane = lt & 63;
#pragma unroll
    for (int i = 0; i < LPW; ++i) {
      const cf *xl = x + i * R * 4;
#pragma unroll
      for (int m = 0; m < R; ++m)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int pp = 4 * q + v + 16 * j16 + 256 * m;
          row[pp + (pp >> 4)] = xl[4 * m + v];
        }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      const int ln = w * LPW + i;
      cf *d = d0 + tc.db + (int64_t)(ln < tc.nl ? ln : 0) * P.tsout + 2 * lane;
      v4f o[NF / 128];   // positions 2 lane + 128 m and the one after
#pragma unroll
      for (int m = 0; m < NF / 128; ++m) {
        const int pp = 2 * lane + 128 * m;
        const cf a = row[pp + (pp >> 4)], b = row[pp + 1 + ((pp + 1) >> 4)];
        o[m] = v4f{a.x, a.y, b.x, b.y};
      }
      if (ln < tc.nl) {
#pragma unroll
        for (int m = 0; m < NF / 128; ++m)
          __builtin_nontemporal_store(o[m], reinterpret_cast<v4f *>(d + 128 * m));
      }
      // the row is rewritten by the next line only after every lane read it
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    if (!more) return false;
    item = ni;
    tc = tn;
    return true;
  };
  cf xa[4 * NP], xb[4 * NP];
  load(tc, tid, xa, true);
  lds_barrier();   // twiddle table
  while (step(xa, xb) && step(xb, xa)) {
  }
}

template <typename K>
static int64_t resident_grid(K kernel, int threads) {
  int dev = 0, ncu = 0, occ = 0;
  (void)hipGetDevice(&dev);
  (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kernel, threads, 0) != hipSuccess ||
      occ < 1)
    occ = 1;
  if (ncu < 8) ncu = 8;
  return (int64_t)occ * (ncu / 8) * 8;
}

template <int NF, int MODE, int SIGN>
static int launch_tile(const pfk_pass *p, const void *src, void *dst) {
  int64_t ntiles = (p->text + 15) / 16;
  for (int i = 0; i < p->nb; ++i) ntiles *= p->bext[i];
  if (ntiles <= 0) return 0;
  static int64_t pgrid = 0;
  const float *s = (const float *)src;
  float *d = (float *)dst;
  if (!(p->variant & 4)) {
    if (!pgrid) pgrid = resident_grid(k_mfma_tile<NF, MODE, SIGN>, NF);
    const int64_t g = ntiles < pgrid ? ntiles : pgrid;
    hipLaunchKernelGGL((k_mfma_tile<NF, MODE, SIGN>), dim3((unsigned)g), dim3(NF), 0, g_stream,
                       *p, (int64_t)0, ntiles, ntiles, s, d);
  } else {
    const int64_t chunk = pfk_max_grid(NF);
    for (int64_t base = 0; base < ntiles; base += chunk) {
      const int64_t nblk = (ntiles - base) < chunk ? (ntiles - base) : chunk;
      hipLaunchKernelGGL((k_mfma_tile<NF, MODE, SIGN>), dim3((unsigned)nblk), dim3(NF), 0,
                         g_stream, *p, base, base + nblk, ntiles, s, d);
    }
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) { set_error("pfk_launch_pass(mfma tile)", e); return -1; }
  return 0;
}

template <int NF, int MODE, int SIGN>
static int launch_line(const pfk_pass *p, const void *src, void *dst) {
  constexpr int CL = LineGeo<NF>::CL;
  int64_t ntiles = (p->text + CL - 1) / CL;
  for (int i = 0; i < p->nb; ++i) ntiles *= p->bext[i];
  if (ntiles <= 0) return 0;
  static int64_t pgrid = 0;
  const float *s = (const float *)src;
  float *d = (float *)dst;
  if (!(p->variant & 4)) {
    if (!pgrid) pgrid = resident_grid(k_mfma_line<NF, MODE, SIGN>, 256);
    const int64_t g = ntiles < pgrid ? ntiles : pgrid;
    hipLaunchKernelGGL((k_mfma_line<NF, MODE, SIGN>), dim3((unsigned)g), dim3(256), 0, g_stream,
                       *p, (int64_t)0, ntiles, ntiles, s, d);
  } else {
    const int64_t chunk = pfk_max_grid(256);
    for (int64_t base = 0; base < ntiles; base += chunk) {
      const int64_t nblk = (ntiles - base) < chunk ? (ntiles - base) : chunk;
      hipLaunchKernelGGL((k_mfma_line<NF, MODE, SIGN>), dim3((unsigned)nblk), dim3(256), 0,
                         g_stream, *p, base, base + nblk, ntiles, s, d);
    }
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) { set_error("pfk_launch_pass(mfma line)", e); return -1; }
  return 0;
}

template <int NF, int MODE>
static int launch_nf_mode(const pfk_pass *p, bool line, const void *src, void *dst) {
  if (line)
    return p->sign < 0 ? launch_line<NF, MODE, -1>(p, src, dst)
                       : launch_line<NF, MODE, +1>(p, src, dst);
  return p->sign < 0 ? launch_tile<NF, MODE, -1>(p, src, dst)
                     : launch_tile<NF, MODE, +1>(p, src, dst);
}

}  // namespace

// Plain complex64 c2c passes of 256, 512 and 1024 points in an MFMA form.
int mfma_chain_covers(const pfk_pass *p) {
  if (p->n1 || p->dec_in || p->dec_out || p->twn || p->hw) return 0;
  if (p->mfma == 0 || p->prec != PFK_F32 || p->kind != PFK_C2C) return 0;
  if (p->nf != 256 && p->nf != 512 && p->nf != 1024) return 0;
  if (p->n != p->nf || p->zin || p->zout || p->nin != p->nf || p->nout != p->nf) return 0;
  return !p->imod.on && !p->omod.on;
}

// line-major form: lines contiguous on both sides; the stores (and for
// 512 and 1024 points the loads) take 16-byte vectors, so every line must
// start on a 16-byte boundary (even element offsets)
static bool line_form_ok(const pfk_pass *p, const void *src, const void *dst) {
  if (p->sin != 1 || p->sout != 1) return false;
  if (((uintptr_t)dst & 15) || (p->tsout & 1)) return false;
  for (int i = 0; i < p->nb; ++i)
    if (p->bsout[i] & 1) return false;
  if (p->nf == 256) return true;
  if (((uintptr_t)src & 15) || (p->tsin & 1)) return false;
  for (int i = 0; i < p->nb; ++i)
    if (p->bsin[i] & 1) return false;
  return true;
}

// 0 launched, 1 not covered, -1 error
int launch_mfma_chain(const pfk_pass *p, const void *src, void *dst) {
  if (!mfma_chain_covers(p)) return 1;
  const bool line = (p->variant & 8) && line_form_ok(p, src, dst);
  const int mode = p->mfma - 1;
  switch (p->nf) {
    case 256:
      return mode ? launch_nf_mode<256, 1>(p, line, src, dst) : launch_nf_mode<256, 0>(p, line, src, dst);
    case 512:
      return mode ? launch_nf_mode<512, 1>(p, line, src, dst) : launch_nf_mode<512, 0>(p, line, src, dst);
    default:
      return mode ? launch_nf_mode<1024, 1>(p, line, src, dst)
                  : launch_nf_mode<1024, 0>(p, line, src, dst);
  }
}

}  // namespace pfk
