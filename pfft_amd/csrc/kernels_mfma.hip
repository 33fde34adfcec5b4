// kernels_mfma.hip -- complex64 c2c passes whose radix-16 stages run on the
// matrix cores (gfx950 MFMA), for plans created with PFFT_AMD_MFMA_F32 or
// PFFT_AMD_MFMA_BF16X2.
//
// A radix-16 Stockham stage applies the 16x16 DFT matrix F (F_qr = W16^(q r),
// W16 = exp(sign 2 pi i/16)) to 16 twiddled inputs per butterfly.  Sixteen
// butterflies form the 16 columns of X, so one stage tile is the complex
// product Y = F X.  In real arithmetic
//   Yr = [Fr | -Fi] [Xr; Xi],   Yi = [Fi | Fr] [Xr; Xi]        (K = 32),
// computed per wave as 16x16 MFMA tiles:
//   FP32 : v_mfma_f32_16x16x4f32, 8 K-steps per output half (16 per tile).
//          Lane (j = l%16, q = l/16) supplies rows q, q+4, q+8, q+12 of
//          column j (re and im) and receives rows 4q..4q+3 of column j.
//   BF16X2: every operand v is split v = hi + lo (hi = bf16(v),
//          lo = bf16(v - hi)) and Y = Fhi Xhi + Fhi Xlo + Flo Xhi is
//          accumulated in FP32 with v_mfma_f32_16x16x16bf16_1k (2 K-blocks
//          per product, 12 MFMAs per tile).  Lane (j, q) supplies rows
//          4q..4q+3 of column j.  The DFT matrix is pre-scaled by 1/4 per
//          stage (exact in both formats) so magnitudes stay O(1) through the
//          stages; the final store multiplies the 4^stages back (exact).
// Stages with a remaining factor < 16 (the last radix-2/4/8 stage of
// 512/1024/2048-point lines) run on the vector ALUs.  The tile of C lines
// lives in LDS for the whole pass (layout [position][line], C = 16 lines =
// 128-byte global segments of complex64); every stage reads its inputs from
// LDS and writes its outputs back, with a barrier between the read and the
// write phases.
#include <stdint.h>
#include "pfk_internal.hpp"

namespace pfk {

typedef float v4f __attribute__((ext_vector_type(4)));
typedef short v4s __attribute__((ext_vector_type(4)));

struct f2 {
  float x, y;
};

__device__ __forceinline__ f2 f2mul(f2 a, f2 b) { return f2{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }

// +-1 modulation factor of index j, with selects only (no branch between a
// load and its use)
__device__ __forceinline__ float msign_m(const pfk_mod &m, int64_t j) {
  const int64_t t = j + m.start;
  const int s = (m.on && t < m.half) ? ((t & 1) ? -m.extra : m.extra) : 1;
  return (float)s;
}

// round-to-nearest-even float -> bf16 bits
__device__ __forceinline__ uint16_t bf16_bits(float f) {
  const uint32_t u = __float_as_uint(f);
  const uint32_t r = u + 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(r >> 16);
}
__device__ __forceinline__ float bf16_float(uint16_t b) { return __uint_as_float(((uint32_t)b) << 16); }

// cos/sin of 2*pi*m/16, m in [0,16)
__constant__ float k_cos16[16] = {1.f, 0.92387953251128675613f, 0.70710678118654752440f,
                                  0.38268343236508977173f, 0.f, -0.38268343236508977173f,
                                  -0.70710678118654752440f, -0.92387953251128675613f, -1.f,
                                  -0.92387953251128675613f, -0.70710678118654752440f,
                                  -0.38268343236508977173f, 0.f, 0.38268343236508977173f,
                                  0.70710678118654752440f, 0.92387953251128675613f};
__device__ __forceinline__ f2 w16(int m, int sign) {
  m &= 15;
  return f2{k_cos16[m], sign * k_cos16[(m + 12) & 15]};   // sin(x) = cos(x - pi/2)
}

template <int NF, int C, int MODE, int SIGN>
__global__ void __launch_bounds__(512)
k_c2c_mfma(const pfk_pass P, const int64_t tile_base, const float *__restrict__ src,
           float *__restrict__ dst) {
  constexpr int NT = 512;
  constexpr int LG = (NF == 256) ? 8 : (NF == 512) ? 9 : (NF == 1024) ? 10 : (NF == 2048) ? 11 : 12;
  constexpr int A = LG / 4;          // radix-16 MFMA stages
  constexpr int RL = 1 << (LG % 4);  // last VALU stage radix (1: none)
  __shared__ f2 lds[NF * C];
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int jj = lane & 15, qq = lane >> 4;

  // ---- tile decode (as in the other pass kernels) ----
  const int64_t ntt = (P.text + C - 1) / C;
  const int64_t tileid = tile_base + blockIdx.x;
  const int64_t tt = tileid % ntt;
  int64_t bflat = tileid / ntt;
  const int64_t c0 = tt * C;
  const int nl = (int)((P.text - c0) < C ? (P.text - c0) : C);
  int64_t sb = c0 * P.tsin, db = c0 * P.tsout;
  for (int i = 0; i < P.nb; ++i) {
    const int64_t e = P.bext[i];
    const int64_t d = bflat % e;
    bflat /= e;
    sb += d * P.bsin[i];
    db += d * P.bsout[i];
  }
  const f2 *s = reinterpret_cast<const f2 *>(src);
  f2 *d = reinterpret_cast<f2 *>(dst);
  const f2 *tw = reinterpret_cast<const f2 *>(P.tw);
  const int zin = (int)P.zin, nin = (int)P.nin, zout = (int)P.zout, nout = (int)P.nout;
  const bool ld_tile = (P.tsin == 1 && P.sin != 1);
  const bool st_tile = (P.tsout == 1 && P.sout != 1);

  // ---- load the tile into LDS (embed + input modulation) ----
  {
    // all loads of the thread are issued before any is used; positions
    // outside the data read the tile's first element and are dropped
    constexpr int PERL = NF * C / NT;
    f2 raw[PERL];
#pragma unroll
    for (int i = 0; i < PERL; ++i) {
      const int e = tid + i * NT;
      const int c = ld_tile ? e % C : e / NF, p = ld_tile ? e / C : e % NF;
      const int j = p - zin;
      const bool ok = c < nl && j >= 0 && j < nin;
      raw[i] = s[ok ? sb + (int64_t)c * P.tsin + (int64_t)j * P.sin : sb];
    }
#pragma unroll
    for (int i = 0; i < PERL; ++i) {
      const int e = tid + i * NT;
      const int c = ld_tile ? e % C : e / NF, p = ld_tile ? e / C : e % NF;
      const int j = p - zin;
      const bool ok = c < nl && j >= 0 && j < nin;
      const float f = msign_m(P.imod, j);
      lds[p * C + c] = ok ? f2{raw[i].x * f, raw[i].y * f} : f2{0.f, 0.f};
    }
  }

  // ---- DFT-16 matrix operands held by this lane ----
  // FP32: A[i = jj][kk = 4s + qq] for s = 0..7 of [Fr|-Fi] (Yr) and [Fi|Fr] (Yi)
  // BF16: A[i = jj][k = 4qq + v], v = 0..3, per K-block (hi and lo parts)
  float ar[8], ai[8];
  v4s ahr[2], alr[2], ahi[2], ali[2];
  if (MODE == 0) {
#pragma unroll
    for (int sidx = 0; sidx < 4; ++sidx) {
      const f2 w = w16(jj * (qq + 4 * sidx), SIGN);
      ar[sidx] = w.x;        // Fr
      ar[sidx + 4] = -w.y;   // -Fi
      ai[sidx] = w.y;        // Fi
      ai[sidx + 4] = w.x;    // Fr
    }
  } else {
#pragma unroll
    for (int blk = 0; blk < 2; ++blk) {
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const f2 w = w16(jj * (4 * qq + v), SIGN);
        const float fr = 0.25f * (blk == 0 ? w.x : -w.y);  // Yr: [Fr | -Fi]
        const float fi = 0.25f * (blk == 0 ? w.y : w.x);   // Yi: [Fi |  Fr]
        const uint16_t hr = bf16_bits(fr), hi_ = bf16_bits(fi);
        ahr[blk][v] = (short)hr;
        alr[blk][v] = (short)bf16_bits(fr - bf16_float(hr));
        ahi[blk][v] = (short)hi_;
        ali[blk][v] = (short)bf16_bits(fi - bf16_float(hi_));
      }
    }
  }
  lds_barrier();

  // ---- radix-16 MFMA stages ----
  constexpr int NBF = NF / 16;               // butterflies per line per stage
  constexpr int TILES = (NBF * C) / 16;      // 16-butterfly tiles per stage
  constexpr int TPW = TILES / (NT / 64);     // tiles per wave
  static_assert(TILES % (NT / 64) == 0, "tile count must divide among the waves");
  int L = 1;
#pragma unroll
  for (int st = 0; st < A; ++st) {
    const int M = NF / (L * 16);
    float yr[TPW][4], yi[TPW][4];
#pragma unroll
    for (int w = 0; w < TPW; ++w) {
      const int tile = wave * TPW + w;
      // the 16 columns of a tile are consecutive lines (bank-conflict-free
      // LDS rows) at one butterfly index when C = 16
      const int bglob = tile * 16 + jj;
      const int c = bglob % C, b = bglob / C;
      const int k = b & (L - 1);
      // gather this lane's 4 rows of column jj (twiddled)
      f2 x[4];
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int r = (MODE == 0) ? (qq + 4 * v) : (4 * qq + v);
        f2 a = lds[(b + r * NBF) * C + c];
        if (st > 0 && r > 0) a = f2mul(a, tw[r * k * M]);
        x[v] = a;
      }
      v4f accr = {0.f, 0.f, 0.f, 0.f}, acci = {0.f, 0.f, 0.f, 0.f};
      if (MODE == 0) {
#pragma unroll
        for (int sidx = 0; sidx < 8; ++sidx) {
          const float bv = (sidx < 4) ? x[sidx].x : x[sidx - 4].y;
          accr = __builtin_amdgcn_mfma_f32_16x16x4f32(ar[sidx], bv, accr, 0, 0, 0);
          acci = __builtin_amdgcn_mfma_f32_16x16x4f32(ai[sidx], bv, acci, 0, 0, 0);
        }
      } else {
        v4s xh[2], xl[2];
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const uint16_t hr = bf16_bits(x[v].x), hi_ = bf16_bits(x[v].y);
          xh[0][v] = (short)hr;
          xl[0][v] = (short)bf16_bits(x[v].x - bf16_float(hr));
          xh[1][v] = (short)hi_;
          xl[1][v] = (short)bf16_bits(x[v].y - bf16_float(hi_));
        }
#pragma unroll
        for (int blk = 0; blk < 2; ++blk) {
          accr = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ahr[blk], xh[blk], accr, 0, 0, 0);
          accr = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ahr[blk], xl[blk], accr, 0, 0, 0);
          accr = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(alr[blk], xh[blk], accr, 0, 0, 0);
          acci = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ahi[blk], xh[blk], acci, 0, 0, 0);
          acci = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ahi[blk], xl[blk], acci, 0, 0, 0);
          acci = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ali[blk], xh[blk], acci, 0, 0, 0);
        }
      }
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        yr[w][v] = accr[v];
        yi[w][v] = acci[v];
      }
    }
    lds_barrier();
#pragma unroll
    for (int w = 0; w < TPW; ++w) {
      const int tile = wave * TPW + w;
      const int bglob = tile * 16 + jj;
      const int c = bglob % C, b = bglob / C;
      const int k = b & (L - 1);
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int q = 4 * qq + v;   // output row held in accumulator slot v
        lds[((b - k) * 16 + k + q * L) * C + c] = f2{yr[w][v], yi[w][v]};
      }
    }
    lds_barrier();
    L *= 16;
  }

  // ---- last radix-RL stage on the vector ALUs ----
  if (RL > 1) {
    constexpr int NB2 = NF / RL;              // butterflies per line
    constexpr int PER = (NB2 * C + NT - 1) / NT;
    f2 res[PER][RL > 1 ? RL : 1];
    const int M = NF / (L * RL);              // == 1 (last stage)
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      const int g = tid + u * NT;
      if (g < NB2 * C) {
        const int c = g % C, b = g / C;
        const int k = b & (L - 1);
        f2 a[RL > 1 ? RL : 1];
#pragma unroll
        for (int r = 0; r < RL; ++r) {
          a[r] = lds[(b + r * NB2) * C + c];
          if (r > 0) a[r] = f2mul(a[r], tw[r * k * M]);
        }
#pragma unroll
        for (int q = 0; q < RL; ++q) {
          f2 acc{0.f, 0.f};
#pragma unroll
          for (int r = 0; r < RL; ++r) {
            const f2 w = w16((16 / RL) * ((q * r) % RL), SIGN);
            const f2 t = f2mul(a[r], w);
            acc.x += t.x;
            acc.y += t.y;
          }
          res[u][q] = acc;
        }
      }
    }
    lds_barrier();
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      const int g = tid + u * NT;
      if (g < NB2 * C) {
        const int c = g % C, b = g / C;
        const int k = b & (L - 1);
#pragma unroll
        for (int q = 0; q < RL; ++q) lds[((b - k) * RL + k + q * L) * C + c] = res[u][q];
      }
    }
    lds_barrier();
  }

  // ---- store (trunc + output modulation + undo the per-stage scaling) ----
  const float unscale = (MODE == 1) ? (float)(1 << (2 * A)) : 1.0f;
  for (int e = tid; e < nout * C; e += NT) {
    int c, k;
    if (st_tile) { c = e % C; k = e / C; }
    else { k = e % nout; c = e / nout; }
    if (c >= nl) continue;
    f2 v = lds[(k + zout) * C + c];
    const float f = msign_m(P.omod, k) * unscale;
    d[db + (int64_t)c * P.tsout + (int64_t)k * P.sout] = f2{v.x * f, v.y * f};
  }
}

template <int NF, int MODE>
static int launch_mfma_nf(const pfk_pass *p, const void *src, void *dst) {
  constexpr int C = (NF <= 1024) ? 16 : (NF == 2048 ? 8 : 4);
  int64_t ntiles = (p->text + C - 1) / C;
  for (int i = 0; i < p->nb; ++i) ntiles *= p->bext[i];
  if (ntiles <= 0) return 0;
  const int64_t chunk = (int64_t)1 << 30;
  for (int64_t base = 0; base < ntiles; base += chunk) {
    const int64_t nblk = (ntiles - base) < chunk ? (ntiles - base) : chunk;
    if (p->sign < 0)
      hipLaunchKernelGGL((k_c2c_mfma<NF, C, MODE, -1>), dim3((unsigned)nblk), dim3(512), 0,
                         g_stream, *p, base, (const float *)src, (float *)dst);
    else
      hipLaunchKernelGGL((k_c2c_mfma<NF, C, MODE, +1>), dim3((unsigned)nblk), dim3(512), 0,
                         g_stream, *p, base, (const float *)src, (float *)dst);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) { set_error("pfk_launch_pass(mfma)", e); return -1; }
  }
  return 0;
}

int mfma_covers(const pfk_pass *p) {
  if (p->mfma == 0 || p->prec != PFK_F32 || p->kind != PFK_C2C || p->hw) return 0;
  return p->nf >= 256 && p->nf <= 4096 && (p->nf & (p->nf - 1)) == 0;
}

// 0 launched, 1 not covered, -1 error
int launch_mfma(const pfk_pass *p, const void *src, void *dst) {
  if (!mfma_covers(p)) return 1;
  const int mode = p->mfma - 1;
  switch (p->nf) {
    case 256: return mode ? launch_mfma_nf<256, 1>(p, src, dst) : launch_mfma_nf<256, 0>(p, src, dst);
    case 512: return mode ? launch_mfma_nf<512, 1>(p, src, dst) : launch_mfma_nf<512, 0>(p, src, dst);
    case 1024: return mode ? launch_mfma_nf<1024, 1>(p, src, dst) : launch_mfma_nf<1024, 0>(p, src, dst);
    case 2048: return mode ? launch_mfma_nf<2048, 1>(p, src, dst) : launch_mfma_nf<2048, 0>(p, src, dst);
    case 4096: return mode ? launch_mfma_nf<4096, 1>(p, src, dst) : launch_mfma_nf<4096, 0>(p, src, dst);
    default: return 1;
  }
}

}  // namespace pfk
