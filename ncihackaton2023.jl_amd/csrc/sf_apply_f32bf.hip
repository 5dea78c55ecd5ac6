// sf_apply_f32bf.hip -- FP32 contractions with the geometric factor stored in
// bfloat16.  Rounding G is a perturbation of the diffusion coefficient, so
// the operator stays symmetric positive definite and spectrally equivalent
// to the exact one (within (1 +- c) with c ~ 3 * 2^-9 * cond(G_q)); as the
// inner operator of iterative refinement it still contracts the error by
// about c per outer step while reading half the G bytes of FP32.
#include "sf_apply.hpp"
#include "sf_kernels.h"

namespace sf {

hipError_t apply_f32_gbf16(int dim, int n, int q, const float* x, float* y,
                           const int32_t* cells, const bf16s* G, const float* tB,
                           const float* tD, int64_t ncell, const BlockArgs* k,
                           int accumulate, hipStream_t s) {
  ApplyArgs<float, bf16s> a{x, y, cells, G, tB, tD, ncell, accumulate};
  return launch_apply<float, bf16s>(dim, n, q, a, k, s);
}

}  // namespace sf
