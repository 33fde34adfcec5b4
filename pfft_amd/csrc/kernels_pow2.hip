// kernels_pow2.hip -- specialised pass kernels for the hot shapes.
#include "pfk_internal.hpp"

namespace pfk {

int launch_fast(const pfk_pass *p, const void *src, void *dst) {
  (void)p; (void)src; (void)dst;
  return 1;
}

}  // namespace pfk
