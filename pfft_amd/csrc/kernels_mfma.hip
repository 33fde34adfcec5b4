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

// ---------------------------------------------------------------------------
// Register-resident MFMA pass (256, 512, 1024 points).  The workgroup
// transforms C = 16 lines with NF threads, 16 complex values per thread, as
// the power-of-two VALU kernels do; only the data layout of the radix-16
// stages differs.  Lane l of wave w works on line c = l % 16, which is also
// the column index n = l % 16 of the MFMA B operand, and on row group
// q = l / 16.  For a radix-16 stage the wave owns butterflies
// b = 4w + g (g = 0..3, one 16x16 MFMA tile each: 16 lines x 16 inputs);
// lane (c, q) holds rows r(q, v), v = 0..3, of butterfly b of line c:
// r = q + 4v for the FP32 product (K = 4 per instruction), r = 4q + v for
// BF16 (K = 16 per instruction).  These operands are gathered straight
// from global memory (first stage) or from the LDS exchange between stages,
// so the MFMA layout costs no extra data movement; outputs (rows 4q + v of
// the accumulator) are scattered to their Stockham positions in the next
// exchange.  The remaining radix-2/4 stage runs on the vector ALUs in the
// usual layout (thread t = 4w + q holds positions t + T*m).
template <int MODE>
__device__ __forceinline__ int mrow(int q, int v) { return MODE == 0 ? q + 4 * v : 4 * q + v; }

struct MTile {
  int64_t sb, db;
  int nl;
};

__device__ __forceinline__ MTile mtile_at(const pfk_pass &P, int64_t item, int64_t ntiles) {
  constexpr int C = 16;
  int64_t tileid = item;
  if (P.xcd_remap) {
    const int64_t qn = ntiles >> 3, rn = ntiles & 7, x = item & 7, jx = item >> 3;
    tileid = x * qn + (x < rn ? x : rn) + jx;
  }
  const int64_t ntt = (P.text + C - 1) / C;
  const int64_t tt = tileid % ntt;
  int64_t bflat = tileid / ntt;
  const int64_t c0 = tt * C;
  MTile r;
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

// Line and butterfly of MFMA tile g of `wave` for lane column n.
// Tile-major (LM false): the 16 columns are the workgroup's 16 lines at
// butterfly 4*wave + g.  Line-major (LM true, for passes whose lines are
// contiguous in memory): the columns are 16 consecutive butterflies of one
// line, so a wave's loads and stores run along the line.
template <int NF, bool LM>
__device__ __forceinline__ void mtile(int wave, int g, int n, int &line, int &b) {
  constexpr int NBLK = NF / 256;   // 16-butterfly blocks per line
  if constexpr (LM) {
    const int T = 4 * wave + g;
    line = T / NBLK;
    b = 16 * (T % NBLK) + n;
  } else {
    line = n;
    b = 4 * wave + g;
  }
}
// LDS slot of position p of line c: [position][line], or for line-major
// order [line][position] with one slot of padding per 16 positions
template <int NF, bool LM>
__device__ __forceinline__ int mlidx(int p, int c) {
  if constexpr (LM) return c * (NF + NF / 16) + p + (p >> 4);
  else return p * 16 + c;
}

// raw loads of the first stage's operands of tile `tc` (all issued before
// use; `live` false reads the tile's first element only)
template <int NF, int MODE, bool LM>
__device__ __forceinline__ void mload(const pfk_pass &P, const float *src, const MTile &tc,
                                      bool live, int wave, int n, int q, f2 *x) {
  constexpr int TT = NF / 16;
  const int zin = (int)P.zin, nin = (int)P.nin;
  const f2 *s0 = reinterpret_cast<const f2 *>(src) + tc.sb;
  const int64_t sin = live ? P.sin : 0;
#pragma unroll
  for (int g = 0; g < 4; ++g) {
    int line, b;
    mtile<NF, LM>(wave, g, n, line, b);
    const f2 *s = (live && line < tc.nl) ? s0 + (int64_t)line * P.tsin : s0;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int j = b + TT * mrow<MODE>(q, v) - zin;
      const bool ok = j >= 0 && j < nin;
      x[4 * g + v] = s[(int64_t)(ok ? j : 0) * sin];
    }
  }
}

// Persistent: a resident grid walks the tiles (items item0 + b, + G, ...
// below item_end); the next tile's operands are loaded into a second
// register buffer while the current tile is transformed.  Launched with
// one workgroup per tile it does one tile per workgroup.
template <int NF, int MODE, int SIGN, bool LM>
__global__ void __launch_bounds__(NF)
k_c2c_mfma_reg(const pfk_pass P, const int64_t item0, const int64_t item_end,
               const int64_t ntiles, const float *__restrict__ src, float *__restrict__ dst) {
  constexpr int C = 16, TT = NF / 16;
  constexpr int LG = (NF == 256) ? 8 : (NF == 512) ? 9 : 10;
  constexpr int A = LG / 4;          // radix-16 MFMA stages
  constexpr int RL = 1 << (LG % 4);  // last VALU stage radix (1: none)
  __shared__ f2 lds[LM ? C * (NF + NF / 16) : NF * C];
  __shared__ f2 twl[NF];
  const int tid = threadIdx.x;
  twl[tid] = reinterpret_cast<const f2 *>(P.tw)[tid];
  const int zin = (int)P.zin, nin = (int)P.nin, zout = (int)P.zout, nout = (int)P.nout;

  // ---- DFT-16 matrix operands held by this lane (row i = l%16, k group l/16) ----
  float ar[8], ai[8];
  v4s ahr[2], alr[2], ahi[2], ali[2];
  {
    const int c = tid & 15, q = (tid & 63) >> 4;
    if (MODE == 0) {
#pragma unroll
      for (int sidx = 0; sidx < 4; ++sidx) {
        const f2 w = w16(c * (q + 4 * sidx), SIGN);
        ar[sidx] = w.x;
        ar[sidx + 4] = -w.y;
        ai[sidx] = w.y;
        ai[sidx + 4] = w.x;
      }
    } else {
#pragma unroll
      for (int blk = 0; blk < 2; ++blk)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const f2 w = w16(c * (4 * q + v), SIGN);
          const float fr = 0.25f * (blk == 0 ? w.x : -w.y);
          const float fi = 0.25f * (blk == 0 ? w.y : w.x);
          const uint16_t hr = bf16_bits(fr), hi_ = bf16_bits(fi);
          ahr[blk][v] = (short)hr;
          alr[blk][v] = (short)bf16_bits(fr - bf16_float(hr));
          ahi[blk][v] = (short)hi_;
          ali[blk][v] = (short)bf16_bits(fi - bf16_float(hi_));
        }
    }
  }

  int64_t item = item0 + blockIdx.x;
  MTile tc = mtile_at(P, item, ntiles);
  // one step: prefetch the next tile into `nx`, transform and store `x`
  auto step = [&](f2 *x, f2 *nx) -> bool {
    // lane indices made opaque per step so index arithmetic is recomputed
    // instead of hoisted into live registers
    int lt = tid;
    asm volatile("" : "+v"(lt));
    const int lane = lt & 63, wave = lt >> 6;
    const int n = lane & 15, q = lane >> 4;
    // VALU layout of the last stage and the store: line c, butterfly t
    const int c = LM ? lt / TT : n;
    const int t = LM ? lt % TT : (lt >> 4);
    const int64_t ni = item + gridDim.x;
    const bool more = ni < item_end;
    const MTile tn = mtile_at(P, more ? ni : item, ntiles);
    mload<NF, MODE, LM>(P, src, tn, more, wave, n, q, nx);
#pragma unroll
    for (int g = 0; g < 4; ++g) {
      int line, b;
      mtile<NF, LM>(wave, g, n, line, b);
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int j = b + TT * mrow<MODE>(q, v) - zin;
        const bool ok = line < tc.nl && j >= 0 && j < nin;
        const float f = msign_m(P.imod, j);
        x[4 * g + v] = ok ? f2{x[4 * g + v].x * f, x[4 * g + v].y * f} : f2{0.f, 0.f};
      }
    }

    // ---- radix-16 stages on the matrix cores ----
#pragma unroll
    for (int st = 0; st < A; ++st) {
      const int L = 1 << (4 * st), M = NF / (L * 16);
#pragma unroll
      for (int g = 0; g < 4; ++g) {
        int line, b;
        mtile<NF, LM>(wave, g, n, line, b);
        const int k = b & (L - 1);
        if (st > 0) {
#pragma unroll
          for (int v = 0; v < 4; ++v)
            x[4 * g + v] = f2mul(x[4 * g + v], twl[mrow<MODE>(q, v) * k * M]);
        }
        v4f accr = {0.f, 0.f, 0.f, 0.f}, acci = {0.f, 0.f, 0.f, 0.f};
        if (MODE == 0) {
#pragma unroll
          for (int sidx = 0; sidx < 8; ++sidx) {
            const float bv = (sidx < 4) ? x[4 * g + sidx].x : x[4 * g + sidx - 4].y;
            accr = __builtin_amdgcn_mfma_f32_16x16x4f32(ar[sidx], bv, accr, 0, 0, 0);
            acci = __builtin_amdgcn_mfma_f32_16x16x4f32(ai[sidx], bv, acci, 0, 0, 0);
          }
        } else {
          v4s xh[2], xl[2];
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            const uint16_t hr = bf16_bits(x[4 * g + v].x), hi_ = bf16_bits(x[4 * g + v].y);
            xh[0][v] = (short)hr;
            xl[0][v] = (short)bf16_bits(x[4 * g + v].x - bf16_float(hr));
            xh[1][v] = (short)hi_;
            xl[1][v] = (short)bf16_bits(x[4 * g + v].y - bf16_float(hi_));
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
        // output rows 4q + v of this column, scattered to their Stockham positions
#pragma unroll
        for (int v = 0; v < 4; ++v)
          lds[mlidx<NF, LM>((b - k) * 16 + k + (4 * q + v) * L, line)] = f2{accr[v], acci[v]};
      }
      lds_barrier();
      if (st + 1 < A) {
#pragma unroll
        for (int g = 0; g < 4; ++g) {
          int line, b;
          mtile<NF, LM>(wave, g, n, line, b);
#pragma unroll
          for (int v = 0; v < 4; ++v)
            x[4 * g + v] = lds[mlidx<NF, LM>(b + TT * mrow<MODE>(q, v), line)];
        }
      } else {
#pragma unroll
        for (int m = 0; m < 16; ++m) x[m] = lds[mlidx<NF, LM>(t + TT * m, c)];
      }
      lds_barrier();
    }

    // ---- last radix-RL stage on the vector ALUs (local to the thread) ----
    // butterflies b = t + T*i (i < 16/RL), inputs m = i + r*(16/RL), twiddle
    // W_NF^(r*b); outputs stay in the same registers
    if constexpr (RL > 1) {
      constexpr int NB = 16 / RL;
#pragma unroll
      for (int i = 0; i < NB; ++i) {
        const int b = t + TT * i;
        f2 a[RL];
#pragma unroll
        for (int r = 0; r < RL; ++r) {
          a[r] = x[i + r * NB];
          if (r > 0) a[r] = f2mul(a[r], twl[r * b]);
        }
        if constexpr (RL == 2) {
          x[i] = f2{a[0].x + a[1].x, a[0].y + a[1].y};
          x[i + NB] = f2{a[0].x - a[1].x, a[0].y - a[1].y};
        } else {   // RL == 4, W4 = SIGN*i
          const f2 e0{a[0].x + a[2].x, a[0].y + a[2].y}, e1{a[0].x - a[2].x, a[0].y - a[2].y};
          const f2 e2{a[1].x + a[3].x, a[1].y + a[3].y};
          const f2 d3{a[1].x - a[3].x, a[1].y - a[3].y};
          const f2 e3 = SIGN < 0 ? f2{d3.y, -d3.x} : f2{-d3.y, d3.x};
          x[i] = f2{e0.x + e2.x, e0.y + e2.y};
          x[i + 2 * NB] = f2{e0.x - e2.x, e0.y - e2.y};
          x[i + NB] = f2{e1.x + e3.x, e1.y + e3.y};
          x[i + 3 * NB] = f2{e1.x - e3.x, e1.y - e3.y};
        }
      }
    }

    // ---- store (trunc + output modulation + undo the per-stage scaling) ----
    if (c < tc.nl) {
      const float unscale = (MODE == 1) ? (float)(1 << (2 * A)) : 1.0f;
      f2 *d = reinterpret_cast<f2 *>(dst) + tc.db + (int64_t)c * P.tsout;
#pragma unroll
      for (int m = 0; m < 16; ++m) {
        const int kk = t + TT * m - zout;
        if (kk >= 0 && kk < nout) {
          const float f = msign_m(P.omod, kk) * unscale;
          d[(int64_t)kk * P.sout] = f2{x[m].x * f, x[m].y * f};
        }
      }
    }
    if (!more) return false;
    item = ni;
    tc = tn;
    return true;   // the next step's first LDS write follows a barrier already
  };
  f2 xa[16], xb[16];
  {
    const int lane = tid & 63;
    mload<NF, MODE, LM>(P, src, tc, true, tid >> 6, lane & 15, lane >> 4, xa);
  }
  lds_barrier();   // twiddle table
  while (step(xa, xb) && step(xb, xa)) {
  }
}

template <int NF, int MODE, int SIGN, bool LM>
static int launch_mfma_reg_s(const pfk_pass *p, int64_t ntiles, const void *src, void *dst) {
  static int64_t pgrid = 0;   // resident workgroups (per instantiation)
  if (!pgrid) {
    int dev = 0, ncu = 0, occ = 0;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k_c2c_mfma_reg<NF, MODE, SIGN, LM>,
                                                     NF, 0) != hipSuccess || occ < 1)
      occ = 1;
    if (ncu < 8) ncu = 8;
    pgrid = (int64_t)occ * (ncu / 8) * 8;
  }
  const float *s = (const float *)src;
  float *d = (float *)dst;
  if (!(p->variant & 4)) {
    const int64_t g = ntiles < pgrid ? ntiles : pgrid;
    hipLaunchKernelGGL((k_c2c_mfma_reg<NF, MODE, SIGN, LM>), dim3((unsigned)g), dim3(NF), 0,
                       g_stream, *p, (int64_t)0, ntiles, ntiles, s, d);
  } else {
    const int64_t chunk = pfk_max_grid(NF);
    for (int64_t base = 0; base < ntiles; base += chunk) {
      const int64_t nblk = (ntiles - base) < chunk ? (ntiles - base) : chunk;
      hipLaunchKernelGGL((k_c2c_mfma_reg<NF, MODE, SIGN, LM>), dim3((unsigned)nblk), dim3(NF), 0,
                         g_stream, *p, base, base + nblk, ntiles, s, d);
    }
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) { set_error("pfk_launch_pass(mfma)", e); return -1; }
  return 0;
}

template <int NF, int MODE>
static int launch_mfma_reg(const pfk_pass *p, const void *src, void *dst) {
  int64_t ntiles = (p->text + 15) / 16;
  for (int i = 0; i < p->nb; ++i) ntiles *= p->bext[i];
  if (ntiles <= 0) return 0;
  if (p->variant & 8)
    return p->sign < 0 ? launch_mfma_reg_s<NF, MODE, -1, true>(p, ntiles, src, dst)
                       : launch_mfma_reg_s<NF, MODE, +1, true>(p, ntiles, src, dst);
  return p->sign < 0 ? launch_mfma_reg_s<NF, MODE, -1, false>(p, ntiles, src, dst)
                     : launch_mfma_reg_s<NF, MODE, +1, false>(p, ntiles, src, dst);
}

// LDS-resident kernel (2048 and 4096 points, where 16 lines of the
// register kernel would not fit the LDS)
template <int NF, int MODE>
static int launch_mfma_nf(const pfk_pass *p, const void *src, void *dst) {
  constexpr int C = (NF <= 1024) ? 16 : (NF == 2048 ? 8 : 4);
  int64_t ntiles = (p->text + C - 1) / C;
  for (int i = 0; i < p->nb; ++i) ntiles *= p->bext[i];
  if (ntiles <= 0) return 0;
  const int64_t chunk = pfk_max_grid(512);
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
  if (p->n1 || p->dec_in || p->dec_out || p->twn) return 0;   // steps of long lines
  if (p->mfma == 0 || p->prec != PFK_F32 || p->kind != PFK_C2C || p->hw) return 0;
  // the MFMA kernels coalesce lines that are contiguous on both sides or
  // adjacent along the tile dimension on both sides; a transposing pass
  // (one of each) runs the vector-ALU kernels, which stage one side
  // through LDS
  if (!((p->sin == 1 && p->sout == 1) || (p->tsin == 1 && p->tsout == 1))) return 0;
  return p->nf >= 256 && p->nf <= 4096 && (p->nf & (p->nf - 1)) == 0;
}

// 0 launched, 1 not covered, -1 error.  Plain passes of 256..1024 points
// run the chained kernels (kernels_mfma_chain.hip); embedded, truncated or
// shifted passes the register kernel above.
int launch_mfma(const pfk_pass *p, const void *src, void *dst) {
  if (!mfma_covers(p)) return 1;
  const int rc = launch_mfma_chain(p, src, dst);
  if (rc != 1) return rc;
  const int mode = p->mfma - 1;
  switch (p->nf) {
    case 256: return mode ? launch_mfma_reg<256, 1>(p, src, dst) : launch_mfma_reg<256, 0>(p, src, dst);
    case 512: return mode ? launch_mfma_reg<512, 1>(p, src, dst) : launch_mfma_reg<512, 0>(p, src, dst);
    case 1024: return mode ? launch_mfma_reg<1024, 1>(p, src, dst) : launch_mfma_reg<1024, 0>(p, src, dst);
    case 2048: return mode ? launch_mfma_nf<2048, 1>(p, src, dst) : launch_mfma_nf<2048, 0>(p, src, dst);
    case 4096: return mode ? launch_mfma_nf<4096, 1>(p, src, dst) : launch_mfma_nf<4096, 0>(p, src, dst);
    default: return 1;
  }
}

}  // namespace pfk
