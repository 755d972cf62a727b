// Split-precision Legendre GEMMs on MFMA: written in a following change.
#include "dino_common.h"

namespace dino {
int legendre_inverse_mfma(const dino_plan*, const void*, void*, int64_t, int prec, hipStream_t) {
  set_error("split precision %d is not available in this build", prec);
  return DINO_EINVAL;
}
int legendre_forward_mfma(const dino_plan*, const void*, void*, int64_t, int prec, hipStream_t) {
  set_error("split precision %d is not available in this build", prec);
  return DINO_EINVAL;
}
}  // namespace dino
