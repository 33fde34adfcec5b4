// kernels_small.hip -- per-pass launch of the register-resident short-length
// kernel (small_tile.hpp): c2c passes one line per lane (256 lines per
// workgroup) or G lanes per line (256/G lines) for the longer lengths,
// including the mid lengths 144..960 that only this per-pass launch
// instantiates; and even-length r2c / c2r passes whose half length has a
// one-chunk split (small_split_real_tile).
#include "small_tile.hpp"

namespace pfk {

template <typename T, int NF>
__global__ void __launch_bounds__(256)
k_small(const pfk_pass P, const int64_t tile_base, const T *__restrict__ src,
        T *__restrict__ dst) {
  extern __shared__ unsigned char smem_raw[];
  small_tile<T, NF>(P, tile_base + blockIdx.x, src, dst, smem_raw);
}

template <typename T, int NF, int KIND>
__global__ void __launch_bounds__(256)
k_small_real(const pfk_pass P, const int64_t tile_base, const T *__restrict__ src,
             T *__restrict__ dst) {
  extern __shared__ unsigned char smem_raw[];
  small_split_real_tile<T, NF, KIND>(P, tile_base + blockIdx.x, src, dst, smem_raw);
}

// The length rule of small_covers for even r2c / c2r passes (without the
// device's LDS limit), for callers that enumerate the covered lengths.
extern "C" int pfk_small_real_length(int nf, int f64) {
  return split_real_length(nf, f64 != 0) && (small_length(nf, f64 != 0) || mid_length(nf, f64 != 0));
}

int small_covers(const pfk_pass *p) {
  if (p->n1 || p->dec_in || p->dec_out || p->twn) return 0;   // steps of long lines
  if (p->nf > 4096) return 0;
  const int nf = (int)p->nf;
  const bool f64 = p->prec == PFK_F64;
  if (p->kind == PFK_C2C) {
    if (p->hw || p->nf != p->n) return 0;
  } else if (p->kind == PFK_R2C || p->kind == PFK_C2R) {
    // even real lines through the half-length complex line (one-chunk shapes)
    if ((p->n & 1) || 2 * p->nf != p->n || !split_real_length(nf, f64)) return 0;
  } else {
    return 0;
  }
  const bool real = p->kind != PFK_C2C;
  if (small::lds_bytes(nf, f64, real) > (size_t)g_max_lds) return 0;
  return (small_length(nf, f64) || mid_length(nf, f64)) ? 1 : 0;
}

template <typename T, int NF>
static int launch_small_nf(const pfk_pass *p, const void *src, void *dst) {
  if constexpr (!small_length(NF, sizeof(T) == 8) && !mid_length(NF, sizeof(T) == 8)) {
    return 1;
  } else {
    int64_t nlines = p->text;
    for (int i = 0; i < p->nb; ++i) nlines *= p->bext[i];
    constexpr bool F64 = sizeof(T) == 8;
    const bool real = p->kind != PFK_C2C;   // real passes take the one-chunk shape
    const int lpt = small_lines_per_tile(NF, F64, real);
    const int64_t ntiles = (nlines + lpt - 1) / lpt;
    const int64_t chunk = pfk_max_grid(small_threads(NF, F64, real));
    for (int64_t base = 0; base < ntiles; base += chunk) {
      const int64_t nblk = (ntiles - base) < chunk ? (ntiles - base) : chunk;
      const dim3 grid((unsigned)nblk), blk(small_threads(NF, F64, real));
      const size_t lds = small::lds_bytes(NF, F64, real);
      if (p->kind == PFK_C2C) {
        hipLaunchKernelGGL((k_small<T, NF>), grid, blk, lds, g_stream, *p, base, (const T *)src,
                           (T *)dst);
      } else if constexpr (split_real_length(NF, F64)) {
        if (p->kind == PFK_R2C)
          hipLaunchKernelGGL((k_small_real<T, NF, PFK_R2C>), grid, blk, lds, g_stream, *p, base,
                             (const T *)src, (T *)dst);
        else
          hipLaunchKernelGGL((k_small_real<T, NF, PFK_C2R>), grid, blk, lds, g_stream, *p, base,
                             (const T *)src, (T *)dst);
      } else {
        return 1;
      }
      hipError_t e = hipGetLastError();
      if (e != hipSuccess) { set_error("pfk_launch_pass(small)", e); return -1; }
    }
    return 0;
  }
}

template <typename T>
static int launch_small_t(const pfk_pass *p, const void *src, void *dst) {
  switch (p->nf) {
#define PFK_SMALL_LAUNCH(N) case N: return launch_small_nf<T, N>(p, src, dst);
    PFK_SMALL_LAUNCH(2) PFK_SMALL_LAUNCH(3) PFK_SMALL_LAUNCH(4) PFK_SMALL_LAUNCH(5)
    PFK_SMALL_LAUNCH(6) PFK_SMALL_LAUNCH(8) PFK_SMALL_LAUNCH(10) PFK_SMALL_LAUNCH(12)
    PFK_SMALL_LAUNCH(15) PFK_SMALL_LAUNCH(16) PFK_SMALL_LAUNCH(20) PFK_SMALL_LAUNCH(24)
    PFK_SMALL_LAUNCH(30) PFK_SMALL_LAUNCH(32) PFK_SMALL_LAUNCH(40) PFK_SMALL_LAUNCH(48)
    PFK_SMALL_LAUNCH(60) PFK_SMALL_LAUNCH(64) PFK_SMALL_LAUNCH(80) PFK_SMALL_LAUNCH(96)
    PFK_SMALL_LAUNCH(120) PFK_SMALL_LAUNCH(128)
    PFK_SMALL_LAUNCH(9) PFK_SMALL_LAUNCH(18) PFK_SMALL_LAUNCH(36) PFK_SMALL_LAUNCH(50)
    PFK_SMALL_LAUNCH(54) PFK_SMALL_LAUNCH(72) PFK_SMALL_LAUNCH(90) PFK_SMALL_LAUNCH(100)
    PFK_SMALL_LAUNCH(108) PFK_SMALL_LAUNCH(25) PFK_SMALL_LAUNCH(27) PFK_SMALL_LAUNCH(45)
    PFK_SMALL_LAUNCH(144) PFK_SMALL_LAUNCH(160) PFK_SMALL_LAUNCH(192) PFK_SMALL_LAUNCH(200)
    PFK_SMALL_LAUNCH(216) PFK_SMALL_LAUNCH(240) PFK_SMALL_LAUNCH(288) PFK_SMALL_LAUNCH(320)
    PFK_SMALL_LAUNCH(384) PFK_SMALL_LAUNCH(400) PFK_SMALL_LAUNCH(432) PFK_SMALL_LAUNCH(480)
    PFK_SMALL_LAUNCH(576) PFK_SMALL_LAUNCH(640) PFK_SMALL_LAUNCH(768) PFK_SMALL_LAUNCH(800)
    PFK_SMALL_LAUNCH(864) PFK_SMALL_LAUNCH(960)
#undef PFK_SMALL_LAUNCH
    default: return 1;
  }
}

// 0 launched, 1 not covered, -1 error
int launch_small(const pfk_pass *p, const void *src, void *dst) {
  if (!small_covers(p)) return 1;
  return (p->prec == PFK_F64) ? launch_small_t<double>(p, src, dst)
                              : launch_small_t<float>(p, src, dst);
}

}  // namespace pfk
