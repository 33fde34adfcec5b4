// kernels_gcell.hip -- ghost-cell exchange and reduce of a 3-d block on one
// device (the reference's gcell module, gcell/gcells_plan.c and
// gcell/gcells_sendrecv.c, on a one-process mesh).
//
// On one process every neighbour is the process itself, so the ghost
// cells of the box [0, ngc0) x [0, ngc1) x [0, ngc2) are periodic images of
// the local block n0 x n1 x n2: box index k along dimension t shows local
// index (k - gb[t]) mod n[t].  Elements are tuples of `tuple` reals
// (2*howmany for complex data), innermost.
//
// exchange: the local block at the front of `data` is copied to a scratch
//   buffer, then one gather kernel writes every element of the box
//   (coalesced stores; reads follow the periodic source index).
// reduce: a gather kernel copies every local element's own box value to a
//   scratch buffer (the whole sum for elements without further images), a
//   second kernel recomputes the full sum for the thin shell of elements
//   that have further images, and a third writes the scratch to the front
//   of `data` and zeroes the rest of the box.  The sum follows the
//   reference's association exactly (see gc_chain), so the result is
//   bitwise the reference's.
//
// Both are HBM-bound byte movement: 32-bit index arithmetic, eight elements
// per thread with all loads issued before the stores.
#include "pfk_internal.hpp"

namespace pfk {
namespace {

constexpr int GC_NT = 256;
constexpr int GC_EPT = 8;
constexpr int64_t GC_CH = (int64_t)GC_NT * GC_EPT;

struct GcDims {
  uint32_t n[3];     // periods = local block
  uint32_t gb[3];    // ghost cells below
  uint32_t ngc[3];   // box extents
  uint32_t off[3];   // (k + off) mod n = (k - gb) mod n for box index k
  uint32_t tuple;
  uint32_t wbox;     // ngc2 * tuple: reals per box row
  uint32_t wloc;     // n2 * tuple: reals per local row
};

// exchange: dst[b] for every box element b <- src (local block) at the
// periodic source index
template <typename T>
__global__ void __launch_bounds__(GC_NT)
k_gc_fill(const GcDims d, const int64_t total, const T *__restrict__ src, T *__restrict__ dst) {
  const int64_t b0 = (int64_t)blockIdx.x * GC_CH;
  const int64_t row0 = b0 / d.wbox;
  const uint32_t e0 = (uint32_t)(b0 - row0 * d.wbox);
  int64_t si[GC_EPT];
  T v[GC_EPT];
#pragma unroll
  for (int j = 0; j < GC_EPT; j++) {
    const uint32_t o = (uint32_t)(j * GC_NT) + threadIdx.x;
    uint32_t e = e0 + o;
    const uint32_t rr = e / d.wbox;
    e -= rr * d.wbox;
    const uint32_t row = (uint32_t)row0 + rr;
    const uint32_t k1 = row % d.ngc[1], k0 = row / d.ngc[1];
    const uint32_t k2 = e / d.tuple, h = e - k2 * d.tuple;
    const uint32_t s0 = (k0 + d.off[0]) % d.n[0];
    const uint32_t s1 = (k1 + d.off[1]) % d.n[1];
    const uint32_t s2 = (k2 + d.off[2]) % d.n[2];
    const int64_t s = (((int64_t)s0 * d.n[1] + s1) * d.n[2] + s2) * d.tuple + h;
    si[j] = (b0 + o < total) ? s : 0;
  }
#pragma unroll
  for (int j = 0; j < GC_EPT; j++) v[j] = src[si[j]];
#pragma unroll
  for (int j = 0; j < GC_EPT; j++) {
    const int64_t b = b0 + (int64_t)(j * GC_NT) + threadIdx.x;
    if (b < total) dst[b] = v[j];
  }
}

// Sum of the images of local index i along one dimension, in the
// reference's order (gcells_sendrecv.c:425-536 at one process): the images
// below are added chunk by chunk from the farthest inwards, then the chain
// is added to the local value, then the same for the images above:
//   (f(q) + (f(q-n) + (f(q-2n) + ...))) + (f(q+n) + (f(q+2n) + ...))
// with q = gb + i the box index of the local element.
template <typename T, typename F>
__device__ __forceinline__ T gc_chain(uint32_t q, uint32_t n, uint32_t ngc, F f) {
  T v = f(q);
  const uint32_t nb = q / n;
  if (nb) {
    T acc = f(q - nb * n);
    for (uint32_t j = nb - 1; j >= 1; j--) acc = f(q - j * n) + acc;
    v = v + acc;
  }
  const uint32_t na = (ngc - 1 - q) / n;
  if (na) {
    T acc = f(q + na * n);
    for (uint32_t j = na - 1; j >= 1; j--) acc = f(q + j * n) + acc;
    v = v + acc;
  }
  return v;
}

// reduce, part 1: out[l] <- the box value at local element l's own
// position (the whole sum for the elements without further images).
template <typename T>
__global__ void __launch_bounds__(GC_NT)
k_gc_own(const GcDims d, const int64_t total, const T *__restrict__ src, T *__restrict__ out) {
  const int64_t b0 = (int64_t)blockIdx.x * GC_CH;
  const int64_t row0 = b0 / d.wloc;
  const uint32_t e0 = (uint32_t)(b0 - row0 * d.wloc);
  const int64_t s1 = (int64_t)d.ngc[2] * d.tuple;     // box strides (reals)
  const int64_t s0 = s1 * d.ngc[1];
  int64_t pos[GC_EPT];
  T v[GC_EPT];
#pragma unroll
  for (int j = 0; j < GC_EPT; j++) {
    const uint32_t o = (uint32_t)(j * GC_NT) + threadIdx.x;
    uint32_t e = e0 + o;
    const uint32_t rr = e / d.wloc;
    e -= rr * d.wloc;
    const uint32_t row = (uint32_t)row0 + rr;
    const uint32_t i2 = e / d.tuple, h = e - i2 * d.tuple;
    // rows past the block (dead lanes) are wrapped into range
    const uint32_t p0 = d.gb[0] + (row / d.n[1]) % d.n[0];
    const uint32_t p1 = d.gb[1] + row % d.n[1];
    pos[j] = (int64_t)p0 * s0 + (int64_t)p1 * s1 + (int64_t)(d.gb[2] + i2) * d.tuple + h;
  }
#pragma unroll
  for (int j = 0; j < GC_EPT; j++) v[j] = src[pos[j]];
#pragma unroll
  for (int j = 0; j < GC_EPT; j++) {
    const int64_t l = b0 + (int64_t)(j * GC_NT) + threadIdx.x;
    if (l < total) out[l] = v[j];
  }
}

// Local elements with images besides themselves: along dimension t those
// with i < ga[t] (an image above) or i >= n[t] - gb[t] (an image below),
// i.e. i in E_t = [0, lo[t]) u [hi[t], n[t]).  The shell is enumerated as
// three disjoint parts  E_0 x all x all,  I_0 x E_1 x all,  I_0 x I_1 x E_2
// (I_t = [lo[t], hi[t])), tuple innermost.
struct GcShell {
  uint32_t lo[3], hi[3];
  int64_t c0, c1, c2;   // reals in the three parts
};

// reduce, part 2: out[l] <- the full sum of local element l's images, for
// the shell elements only (gc_chain order)
template <typename T>
__global__ void __launch_bounds__(GC_NT)
k_gc_shell(const GcDims d, const GcShell sh, const T *__restrict__ src, T *__restrict__ out) {
  const int64_t g = (int64_t)blockIdx.x * GC_NT + threadIdx.x;
  if (g >= sh.c0 + sh.c1 + sh.c2) return;
  const int64_t tu = d.tuple, n1 = d.n[1], n2 = d.n[2];
  auto edge = [&](int t, int64_t e) -> uint32_t {
    return e < sh.lo[t] ? (uint32_t)e : sh.hi[t] + (uint32_t)(e - sh.lo[t]);
  };
  const int64_t E1 = sh.lo[1] + (d.n[1] - sh.hi[1]), E2 = sh.lo[2] + (d.n[2] - sh.hi[2]);
  const int64_t I1 = sh.hi[1] - sh.lo[1];
  uint32_t i0, i1, i2, h;
  int64_t r;
  if (g < sh.c0) {
    r = g;
    h = (uint32_t)(r % tu); r /= tu;
    i2 = (uint32_t)(r % n2); r /= n2;
    i1 = (uint32_t)(r % n1); r /= n1;
    i0 = edge(0, r);
  } else if (g < sh.c0 + sh.c1) {
    r = g - sh.c0;
    h = (uint32_t)(r % tu); r /= tu;
    i2 = (uint32_t)(r % n2); r /= n2;
    i1 = edge(1, r % E1); r /= E1;
    i0 = sh.lo[0] + (uint32_t)r;
  } else {
    r = g - sh.c0 - sh.c1;
    h = (uint32_t)(r % tu); r /= tu;
    i2 = edge(2, r % E2); r /= E2;
    i1 = sh.lo[1] + (uint32_t)(r % I1);
    i0 = sh.lo[0] + (uint32_t)(r / I1);
  }
  const int64_t s1 = (int64_t)d.ngc[2] * tu;
  const int64_t s0 = s1 * d.ngc[1];
  const T v = gc_chain<T>(d.gb[2] + i2, d.n[2], d.ngc[2], [&](uint32_t p2) {
    return gc_chain<T>(d.gb[1] + i1, d.n[1], d.ngc[1], [&](uint32_t p1) {
      return gc_chain<T>(d.gb[0] + i0, d.n[0], d.ngc[0], [&](uint32_t p0) {
        return src[(int64_t)p0 * s0 + (int64_t)p1 * s1 + (int64_t)p2 * tu + h];
      });
    });
  });
  out[((int64_t)i0 * n1 + i1) * n2 * tu + (int64_t)i2 * tu + h] = v;
}

// reduce, part 3: data[b] <- b < nloc ? loc[b] : 0 over the box (with
// nloc == total: the block copy of exchange)
template <typename T>
__global__ void __launch_bounds__(GC_NT)
k_gc_compact(const int64_t nloc, const int64_t total, const T *__restrict__ loc,
             T *__restrict__ data) {
  const int64_t b0 = (int64_t)blockIdx.x * GC_CH;
  T v[GC_EPT];
#pragma unroll
  for (int j = 0; j < GC_EPT; j++) {
    const int64_t b = b0 + (int64_t)(j * GC_NT) + threadIdx.x;
    v[j] = loc[b < nloc ? b : 0];
    if (b >= nloc) v[j] = (T)0;
  }
#pragma unroll
  for (int j = 0; j < GC_EPT; j++) {
    const int64_t b = b0 + (int64_t)(j * GC_NT) + threadIdx.x;
    if (b < total) data[b] = v[j];
  }
}

int gc_dims(const pfk_gc *g, GcDims *d, int64_t *nloc, int64_t *nbox) {
  int64_t L = g->tuple, G = g->tuple;
  for (int t = 0; t < 3; t++) {
    if (g->n[t] < 1 || g->gb[t] < 0 || g->ngc[t] < g->gb[t] + g->n[t] || g->ngc[t] > INT32_MAX) {
      set_error_msg("pfk_gc: bad ghost-cell geometry");
      return 1;
    }
    d->n[t] = (uint32_t)g->n[t];
    d->gb[t] = (uint32_t)g->gb[t];
    d->ngc[t] = (uint32_t)g->ngc[t];
    d->off[t] = (uint32_t)((g->n[t] - g->gb[t] % g->n[t]) % g->n[t]);
    L *= g->n[t];
    G *= g->ngc[t];
  }
  // 32-bit index arithmetic inside the kernels: rows, row widths and the
  // in-row offsets (< width + GC_CH) must fit
  if (g->tuple < 1 || g->ngc[0] * g->ngc[1] > INT32_MAX ||
      g->ngc[2] * g->tuple > INT32_MAX - GC_CH) {
    set_error_msg("pfk_gc: box too large for 32-bit row indexing");
    return 1;
  }
  d->tuple = (uint32_t)g->tuple;
  d->wbox = (uint32_t)(g->ngc[2] * g->tuple);
  d->wloc = (uint32_t)(g->n[2] * g->tuple);
  *nloc = L;
  *nbox = G;
  return 0;
}

int gc_event(void *const *ev, int k) {
  if (!ev || !ev[k]) return 0;
  hipError_t e = hipEventRecord((hipEvent_t)ev[k], g_stream);
  if (e != hipSuccess) { set_error("pfk_gc event", e); return 1; }
  return 0;
}

template <typename T>
int gc_exchange(const pfk_gc *g, T *data, T *scratch, void *const *ev) {
  GcDims d;
  int64_t L, G;
  if (gc_dims(g, &d, &L, &G)) return 1;
  if (gc_event(ev, 0)) return 1;
  // block copy: the compact kernel with nothing to zero
  hipLaunchKernelGGL(k_gc_compact<T>, dim3((unsigned)((L + GC_CH - 1) / GC_CH)), dim3(GC_NT), 0,
                     g_stream, L, L, (const T *)data, scratch);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) { set_error("pfk_gc_exchange copy", e); return 1; }
  if (gc_event(ev, 1)) return 1;
  const int64_t nb = (G + GC_CH - 1) / GC_CH;
  hipLaunchKernelGGL(k_gc_fill<T>, dim3((unsigned)nb), dim3(GC_NT), 0, g_stream, d, G,
                     (const T *)scratch, data);
  e = hipGetLastError();
  if (e != hipSuccess) { set_error("pfk_gc_exchange", e); return 1; }
  return gc_event(ev, 2);
}

template <typename T>
int gc_reduce(const pfk_gc *g, T *data, T *scratch, void *const *ev) {
  GcDims d;
  int64_t L, G;
  if (gc_dims(g, &d, &L, &G)) return 1;
  GcShell sh;
  int64_t E[3], I[3];
  for (int t = 0; t < 3; t++) {
    const int64_t ga = g->ngc[t] - g->gb[t] - g->n[t];
    const int64_t lo = ga < g->n[t] ? ga : g->n[t];                    // i < ga: image above
    const int64_t hi = (g->n[t] - g->gb[t] > lo) ? g->n[t] - g->gb[t] : lo;   // i >= n - gb: below
    sh.lo[t] = (uint32_t)lo;
    sh.hi[t] = (uint32_t)hi;
    E[t] = lo + (g->n[t] - hi);
    I[t] = hi - lo;
  }
  sh.c0 = E[0] * g->n[1] * g->n[2] * g->tuple;
  sh.c1 = I[0] * E[1] * g->n[2] * g->tuple;
  sh.c2 = I[0] * I[1] * E[2] * g->tuple;
  if (gc_event(ev, 0)) return 1;
  hipLaunchKernelGGL(k_gc_own<T>, dim3((unsigned)((L + GC_CH - 1) / GC_CH)), dim3(GC_NT), 0,
                     g_stream, d, L, (const T *)data, scratch);
  hipError_t e = hipGetLastError();
  const int64_t nshell = sh.c0 + sh.c1 + sh.c2;
  if (e == hipSuccess && nshell > 0) {
    hipLaunchKernelGGL(k_gc_shell<T>, dim3((unsigned)((nshell + GC_NT - 1) / GC_NT)),
                       dim3(GC_NT), 0, g_stream, d, sh, (const T *)data, scratch);
    e = hipGetLastError();
  }
  if (e != hipSuccess) { set_error("pfk_gc_reduce sum", e); return 1; }
  if (gc_event(ev, 1)) return 1;
  hipLaunchKernelGGL(k_gc_compact<T>, dim3((unsigned)((G + GC_CH - 1) / GC_CH)), dim3(GC_NT), 0,
                     g_stream, L, G, (const T *)scratch, data);
  e = hipGetLastError();
  if (e != hipSuccess) { set_error("pfk_gc_reduce compact", e); return 1; }
  return gc_event(ev, 2);
}

}  // namespace
}  // namespace pfk

using namespace pfk;

extern "C" int pfk_gc_exchange(const pfk_gc *g, void *data, void *scratch, void *const *ev) {
  if (!g || !data || !scratch) { set_error_msg("pfk_gc_exchange: null argument"); return 1; }
  return g->prec == PFK_F64 ? gc_exchange<double>(g, (double *)data, (double *)scratch, ev)
                            : gc_exchange<float>(g, (float *)data, (float *)scratch, ev);
}

extern "C" int pfk_gc_reduce(const pfk_gc *g, void *data, void *scratch, void *const *ev) {
  if (!g || !data || !scratch) { set_error_msg("pfk_gc_reduce: null argument"); return 1; }
  return g->prec == PFK_F64 ? gc_reduce<double>(g, (double *)data, (double *)scratch, ev)
                            : gc_reduce<float>(g, (float *)data, (float *)scratch, ev);
}
