// pfk_internal.hpp -- shared definitions of the HIP translation units.
#pragma once
#include <hip/hip_runtime.h>
#include "pfk.h"

namespace pfk {

template <typename T>
struct alignas(2 * sizeof(T)) cpx {
  T x, y;
};

template <typename T>
__device__ __forceinline__ cpx<T> cmul(cpx<T> a, cpx<T> b) {
  return cpx<T>{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}

// Workgroup barrier ordering LDS accesses only.  The pass kernels share
// nothing through global memory inside a launch, so a barrier need not wait
// for outstanding global stores (or prefetched loads, which complete in
// order with them) the way __syncthreads() does.
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

extern hipStream_t g_stream;
extern int g_generic_only;   // PFFT_AMD_GENERIC_ONLY=1: route every pass to the general kernel
extern int g_max_lds;
void set_error(const char *what, hipError_t e);
void set_error_msg(const char *msg);

// Specialised kernels: return 0 if the pass was launched, nonzero if the
// pass shape is not covered (the caller then uses the general kernel).
int launch_fast(const pfk_pass *p, const void *src, void *dst);
int launch_generic(const pfk_pass *p, const void *src, void *dst);
int launch_mfma(const pfk_pass *p, const void *src, void *dst);
int launch_mfma_chain(const pfk_pass *p, const void *src, void *dst);
int launch_small(const pfk_pass *p, const void *src, void *dst);
int fast_covers(const pfk_pass *p);   // 1 if launch_fast has a kernel for the shape
int mfma_covers(const pfk_pass *p);   // 1 if launch_mfma has a kernel for the shape
int mfma_chain_covers(const pfk_pass *p);   // plain 256..1024-point passes (chained MFMA)
int small_covers(const pfk_pass *p);  // 1 if launch_small has a kernel for the shape

}  // namespace pfk
