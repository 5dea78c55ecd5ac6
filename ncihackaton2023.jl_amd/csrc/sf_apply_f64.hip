// sf_apply_f64.hip -- FP64 instantiations of the apply kernels (the plain
// form: FP64 vectors, tables, geometric factor and contractions).
#include "sf_apply.hpp"
#include "sf_kernels.h"

namespace sf {

hipError_t apply_f64(int dim, int n, int q, const double* x, double* y,
                     const int32_t* cells, const double* G, const double* tB,
                     const double* tD, int64_t ncell, const BlockArgs* k,
                     int accumulate, hipStream_t s, const double* Xv,
                     const double* tP, const double* tW, double* dotp,
                     const double* kap) {
  ApplyArgs<double, double> a{x, y, cells, G, tB, tD, ncell, accumulate, Xv, tP, tW, dotp, kap};
  return launch_apply<double, double>(dim, n, q, a, k, s);
}

void apply_shape_f64(int dim, int n, int q, int* nt, int* eb) {
  apply_shape(dim, n, q, nt, eb);
}

}  // namespace sf
