// sf_apply_f32.hip -- FP32 instantiations of the apply kernels: the inner
// operator of the mixed-precision solver (FP32 vectors, tables, G and
// vector-ALU contractions).
#include "sf_apply.hpp"
#include "sf_kernels.h"

namespace sf {

hipError_t apply_f32(int dim, int n, int q, const float* x, float* y,
                     const int32_t* cells, const float* G, const float* tB,
                     const float* tD, int64_t ncell, const BlockArgs* k,
                     int accumulate, hipStream_t s, const float* Xv,
                     const float* tP, const float* tW, double* dotp,
                     const float* kap) {
  ApplyArgs<float, float> a{x, y, cells, G, tB, tD, ncell, accumulate, Xv, tP, tW, dotp, kap};
  return launch_apply<float, float>(dim, n, q, a, k, s);
}

}  // namespace sf
