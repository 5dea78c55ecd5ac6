// fp64_pipe_probe.hip -- whether FP64 matrix-core work (v_mfma_f64_16x16x4_f64)
// runs beside FP64 vector FMAs on gfx950, or shares their pipe.  Not part of
// the library: a stand-alone measurement behind the choice of the vector ALU
// for the FP64 contractions.
//
// Each thread runs ITER rounds of NV FP64 FMAs (8 independent chains) and NM
// 16x16x4 FP64 MFMAs (4 independent accumulators), on registers only; one
// store per thread at the end.  The grid holds 8 workgroups of 256 threads
// per CU.  If the two pipes overlapped, the mixed rows would take the time
// of the longer part alone; if they share one pipe, the times add.
//   make -C ncihackaton2023.jl_amd/csrc probe && ncihackaton2023.jl_amd/csrc/_build/fp64_pipe_probe
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

#define CHECK(x)                                                       \
  do {                                                                 \
    hipError_t e_ = (x);                                               \
    if (e_ != hipSuccess) {                                            \
      fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_));          \
      exit(1);                                                         \
    }                                                                  \
  } while (0)

typedef double d4 __attribute__((ext_vector_type(4)));

template <int NV, int NM, int ITER>
__global__ void __launch_bounds__(256) probe(double* out, double s) {
  double v[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) v[k] = threadIdx.x * 1e-3 + k;
  d4 acc[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) acc[k] = d4{0.0, 0.0, 0.0, (double)k};
  const double a = s * 0.5, b = s * 0.25;
  for (int it = 0; it < ITER; ++it) {
#pragma unroll
    for (int r = 0; r < NV / 8; ++r)
#pragma unroll
      for (int k = 0; k < 8; ++k) v[k] = fma(v[k], s, a);
#pragma unroll
    for (int r = 0; r < NM; ++r)
      acc[r % 4] = __builtin_amdgcn_mfma_f64_16x16x4f64(a + r, b, acc[r % 4], 0, 0, 0);
  }
  double t = 0.0;
#pragma unroll
  for (int k = 0; k < 8; ++k) t += v[k];
#pragma unroll
  for (int k = 0; k < 4; ++k) t += acc[k].x + acc[k].y + acc[k].z + acc[k].w;
  out[blockIdx.x * 256 + threadIdx.x] = t;
}

template <int NV, int NM>
void run(double* d, int nblk, const char* name) {
  constexpr int ITER = 2000;
  hipEvent_t e0, e1;
  CHECK(hipEventCreate(&e0));
  CHECK(hipEventCreate(&e1));
  hipLaunchKernelGGL((probe<NV, NM, ITER>), dim3(nblk), dim3(256), 0, 0, d, 1.0000001);
  CHECK(hipDeviceSynchronize());
  float best = 1e30f;
  for (int rep = 0; rep < 5; ++rep) {
    CHECK(hipEventRecord(e0, 0));
    hipLaunchKernelGGL((probe<NV, NM, ITER>), dim3(nblk), dim3(256), 0, 0, d, 1.0000001);
    CHECK(hipEventRecord(e1, 0));
    CHECK(hipEventSynchronize(e1));
    float ms = 0.f;
    CHECK(hipEventElapsedTime(&ms, e0, e1));
    if (ms < best) best = ms;
  }
  CHECK(hipEventDestroy(e0));
  CHECK(hipEventDestroy(e1));
  const double waves = (double)nblk * 4;
  const double valu = waves * 64.0 * NV * 2.0 * ITER;  // FLOPs
  const double mfma = waves * NM * 2048.0 * ITER;
  printf("%-12s NV=%3d NM=%2d  %8.3f ms  VALU %6.1f TF  MFMA %6.1f TF  total %6.1f TF\n", name,
         NV, NM, best, valu / best / 1e9, mfma / best / 1e9, (valu + mfma) / best / 1e9);
  fflush(stdout);
}

int main() {
  int dev = 0, ncu = 0;
  CHECK(hipGetDevice(&dev));
  CHECK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  const int nblk = ncu * 8;
  double* d = nullptr;
  CHECK(hipMalloc(&d, (size_t)nblk * 256 * sizeof(double)));
  run<64, 0>(d, nblk, "VALU only");
  run<0, 4>(d, nblk, "MFMA only");
  run<0, 8>(d, nblk, "MFMA only");
  run<64, 4>(d, nblk, "VALU + MFMA");
  run<64, 8>(d, nblk, "VALU + MFMA");
  run<32, 8>(d, nblk, "VALU + MFMA");
  run<128, 4>(d, nblk, "VALU + MFMA");
  CHECK(hipFree(d));
  return 0;
}
