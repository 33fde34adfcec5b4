// kernels_probe.hip -- attainable-bandwidth probe of the GPU's HBM.
//
// The pass kernels are HBM-bound, so their yardstick is what a plain
// streaming kernel reaches on the same device.  One kernel template covers
// three access mixes (read only, write only, copy), three widths (16, 32 or
// 64 contiguous bytes per lane per access: one, two or four dwordx4
// operations), cached or nontemporal accesses, and any grid: a persistent
// grid walks the buffer with a grid stride, each thread keeping UNR
// independent accesses in flight per loop trip.  The host sweeps the
// combinations and reports the best of each mix.
#include "pfk_internal.hpp"

namespace pfk {

typedef unsigned int u4 __attribute__((ext_vector_type(4)));

template <bool NT>
__device__ __forceinline__ u4 pld(const u4 *p) {
  if constexpr (NT) return __builtin_nontemporal_load(p);
  else return *p;
}
template <bool NT>
__device__ __forceinline__ void pst(u4 *p, u4 v) {
  if constexpr (NT) __builtin_nontemporal_store(v, p);
  else *p = v;
}

// MODE 0 read, 1 write, 2 copy.  VEC 16-byte vectors per lane and access,
// UNR accesses per thread per loop trip.  n = number of 16-byte vectors
// (a multiple of VEC).  The read mode folds what it loads into one word per
// thread, written to `sink` only when it equals a value no input holds, so
// the loads are live but no bytes are stored.
template <int MODE, int VEC, bool NT>
__global__ void __launch_bounds__(256) k_bw_probe(const u4 *__restrict__ a, u4 *__restrict__ b,
                                                  int64_t n, unsigned *sink) {
  constexpr int UNR = 4;
  const int64_t nchunk = n / VEC;   // VEC-vector chunks
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  unsigned acc = 0;
  for (; i + (UNR - 1) * stride < nchunk; i += UNR * stride) {
    u4 v[UNR][VEC];
    if constexpr (MODE != 1) {
#pragma unroll
      for (int u = 0; u < UNR; ++u)
#pragma unroll
        for (int w = 0; w < VEC; ++w) v[u][w] = pld<NT>(a + (i + u * stride) * VEC + w);
    } else {
#pragma unroll
      for (int u = 0; u < UNR; ++u)
#pragma unroll
        for (int w = 0; w < VEC; ++w) {
          const unsigned s = (unsigned)(i + u * stride) * 2654435761u + w;
          v[u][w] = u4{s, s + 1, s + 2, s + 3};
        }
    }
    if constexpr (MODE == 0) {
#pragma unroll
      for (int u = 0; u < UNR; ++u)
#pragma unroll
        for (int w = 0; w < VEC; ++w) acc ^= v[u][w].x ^ v[u][w].y ^ v[u][w].z ^ v[u][w].w;
    } else {
#pragma unroll
      for (int u = 0; u < UNR; ++u)
#pragma unroll
        for (int w = 0; w < VEC; ++w) pst<NT>(b + (i + u * stride) * VEC + w, v[u][w]);
    }
  }
  for (; i < nchunk; i += stride) {
#pragma unroll
    for (int w = 0; w < VEC; ++w) {
      if constexpr (MODE == 0) {
        const u4 x = pld<NT>(a + i * VEC + w);
        acc ^= x.x ^ x.y ^ x.z ^ x.w;
      } else if constexpr (MODE == 1) {
        const unsigned s = (unsigned)i * 2654435761u + w;
        pst<NT>(b + i * VEC + w, u4{s, s + 1, s + 2, s + 3});
      } else {
        pst<NT>(b + i * VEC + w, pld<NT>(a + i * VEC + w));
      }
    }
  }
  if (MODE == 0 && acc == 0x9E3779B9u) sink[0] = acc;
}

template <int MODE, int VEC, bool NT>
static void launch_probe(int64_t grid, const u4 *a, u4 *b, int64_t n, unsigned *sink) {
  hipLaunchKernelGGL((k_bw_probe<MODE, VEC, NT>), dim3((unsigned)grid), dim3(256), 0, g_stream, a,
                     b, n, sink);
}

template <int MODE>
static void dispatch_probe(int vec, int nt, int64_t grid, const u4 *a, u4 *b, int64_t n,
                           unsigned *sink) {
  if (vec == 1) nt ? launch_probe<MODE, 1, true>(grid, a, b, n, sink)
                   : launch_probe<MODE, 1, false>(grid, a, b, n, sink);
  else if (vec == 2) nt ? launch_probe<MODE, 2, true>(grid, a, b, n, sink)
                        : launch_probe<MODE, 2, false>(grid, a, b, n, sink);
  else nt ? launch_probe<MODE, 4, true>(grid, a, b, n, sink)
          : launch_probe<MODE, 4, false>(grid, a, b, n, sink);
}

}  // namespace pfk

using namespace pfk;

// Best bandwidth (GB/s of bytes read plus bytes written) of `iters` timed
// launches of one probe configuration over `bytes` per buffer.  mode 0 read,
// 1 write, 2 copy; width 16, 32 or 64 bytes per lane; nt nonzero for
// nontemporal accesses; grid = number of 256-thread workgroups, or 0 for
// one pass over the buffer (every thread does one loop trip).  Returns a
// negative value on error.  The buffers are allocated per call.
extern "C" double pfk_bw_probe(int mode, int width, int nt, int64_t grid, size_t bytes,
                               int iters) {
  if (pfk_init()) return -1;
  if (mode < 0 || mode > 2 || (width != 16 && width != 32 && width != 64) || iters < 1)
    return -1;
  const int vec = width / 16;
  int64_t n = (int64_t)(bytes / 16);
  n -= n % (4 * vec);
  if (n <= 0) return -1;
  if (grid <= 0) grid = (n / vec + 4 * 256 - 1) / (4 * 256);
  if (grid > 0x7fffffff / 256) return -1;
  u4 *a = nullptr, *b = nullptr;
  unsigned *sink = nullptr;
  if (hipMalloc(&a, n * 16) != hipSuccess) return -1;
  if (hipMalloc(&b, n * 16) != hipSuccess) { (void)hipFree(a); return -1; }
  if (hipMalloc(&sink, 256) != hipSuccess) { (void)hipFree(a); (void)hipFree(b); return -1; }
  (void)hipMemsetAsync(a, 1, n * 16, g_stream);
  (void)hipMemsetAsync(b, 0, n * 16, g_stream);
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  double best = 0;
  const double moved = (mode == 2 ? 2.0 : 1.0) * (double)n * 16;
  for (int it = 0; it < iters + 1; it++) {
    (void)hipEventRecord(e0, g_stream);
    if (mode == 0) dispatch_probe<0>(vec, nt, grid, a, b, n, sink);
    else if (mode == 1) dispatch_probe<1>(vec, nt, grid, a, b, n, sink);
    else dispatch_probe<2>(vec, nt, grid, a, b, n, sink);
    (void)hipEventRecord(e1, g_stream);
    if (hipEventSynchronize(e1) != hipSuccess) { best = -1; break; }
    float ms = 0;
    (void)hipEventElapsedTime(&ms, e0, e1);
    const double gbs = moved / (ms * 1e-3) / 1e9;
    if (it > 0 && gbs > best) best = gbs;   // the first launch is a warm-up
  }
  if (hipGetLastError() != hipSuccess) best = -1;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(a);
  (void)hipFree(b);
  (void)hipFree(sink);
  return best;
}
