// sf_diag.hip -- the diagonal of the assembled operator, matrix-free, for
// the Jacobi preconditioner of the device-resident CG solvers.
//
// The diagonal of the element matrix K_e = sum_t grad(phi_a)^T G_t
// grad(phi_b) at a = b, with the reference gradients of a tensor-product
// basis function (g_x, g_y, g_z) = (D_x B_y B_z, B_x D_y B_z, B_x B_y D_z)
// (1D tables evaluated at the dof's 1D nodes), is
//     sum_t  G00 g_x^2 + G11 g_y^2 + G22 g_z^2
//          + 2 G01 g_x g_y + 2 G02 g_x g_z + 2 G12 g_y g_z,
// and every product of two 1D factors is a 1D table of its own:
// B^2, D^2 and M = B D (elementwise).  So each of the six terms is a sum
// factorisation with those tables, e.g. 2 G01 g_x g_y -> M_x M_y (B^2)_z.
// Per cell and per z node c the kernel contracts z (six q x q tensors),
// then y (grouped by the x table into three q x n tensors), then x, and
// adds the n x n values into the free dofs of the cell.  A one-time set-up
// kernel (called when a preconditioner is set or the operator changes);
// the reference has no counterpart (its CG is unpreconditioned,
// profiling/drivers.jl:85-95), the arithmetic is the element-matrix
// diagonal of the reference's weak form (src/SumFactorizationKernels.jl:
// 174-216 with a = b).
#include "sf_apply.hpp"
#include "sf_common.hpp"
#include "sf_kernels.h"

namespace sf {

namespace {

constexpr int kDiagThreads = 256;  // >= max(q^2, q n, n^2) for q <= 13, n <= 10

// One workgroup per cell.  G in the block-interleaved layout (g3_index),
// entries G00 G01 G02 G11 G12 G22; tables B, Dm row-major q x n.
__global__ void __launch_bounds__(kDiagThreads)
    k_diag3d(const double* __restrict__ G, const double* __restrict__ B,
             const double* __restrict__ Dm, const int32_t* __restrict__ cells,
             const int32_t* __restrict__ cell_block,
             const int64_t* __restrict__ bcell, int n, int q, int64_t ncell,
             double* __restrict__ diag) {
  __shared__ double sT[6 * kMaxQ * kMaxQ];  // z pass: [term][j][i]
  __shared__ double sP[3 * kMaxN * kMaxQ];  // y pass: [x table][b][i]
  __shared__ double sB2[kMaxQ * kMaxN], sD2[kMaxQ * kMaxN], sM[kMaxQ * kMaxN];
  const int64_t e = blockIdx.x;
  if (e >= ncell) return;  // uniform over the workgroup
  const int tid = threadIdx.x;
  const int qq = q * q;
  for (int t = tid; t < q * n; t += blockDim.x) {
    const double b = B[t], d = Dm[t];
    sB2[t] = b * b;
    sD2[t] = d * d;
    sM[t] = b * d;
  }
  const int cb = cell_block[e];
  const int64_t bs = bcell[cb];
  const int ncb = (int)(bcell[cb + 1] - bs);
  __syncthreads();
  for (int c = 0; c < n; ++c) {
    // z pass: thread (j, i) of the q x q quadrature columns
    if (tid < qq) {
      double t1 = 0, t2 = 0, t3 = 0, t4 = 0, t5 = 0, t6 = 0;
      for (int l = 0; l < q; ++l) {
        const double b2 = sB2[l * n + c], d2 = sD2[l * n + c], m = sM[l * n + c];
        const double g00 = G[g3_index(bs, ncb, e, 0, l, tid, q)];
        const double g01 = G[g3_index(bs, ncb, e, 1, l, tid, q)];
        const double g02 = G[g3_index(bs, ncb, e, 2, l, tid, q)];
        const double g11 = G[g3_index(bs, ncb, e, 3, l, tid, q)];
        const double g12 = G[g3_index(bs, ncb, e, 4, l, tid, q)];
        const double g22 = G[g3_index(bs, ncb, e, 5, l, tid, q)];
        t1 = fma(g00, b2, t1);
        t2 = fma(g11, b2, t2);
        t3 = fma(2.0 * g01, b2, t3);
        t4 = fma(g22, d2, t4);
        t5 = fma(2.0 * g02, m, t5);
        t6 = fma(2.0 * g12, m, t6);
      }
      sT[0 * qq + tid] = t1;
      sT[1 * qq + tid] = t2;
      sT[2 * qq + tid] = t3;
      sT[3 * qq + tid] = t4;
      sT[4 * qq + tid] = t5;
      sT[5 * qq + tid] = t6;
    }
    __syncthreads();
    // y pass: thread (b, i), grouped by the x table (D^2, B^2, M)
    if (tid < q * n) {
      const int b = tid / q, i = tid - b * q;
      double pd = 0, pb = 0, pm = 0;
      for (int j = 0; j < q; ++j) {
        const double b2 = sB2[j * n + b], d2 = sD2[j * n + b], m = sM[j * n + b];
        const int o = j * q + i;
        pd = fma(b2, sT[0 * qq + o], pd);
        pb = fma(d2, sT[1 * qq + o], pb);
        pb = fma(b2, sT[3 * qq + o], pb);
        pb = fma(m, sT[5 * qq + o], pb);
        pm = fma(m, sT[2 * qq + o], pm);
        pm = fma(b2, sT[4 * qq + o], pm);
      }
      sP[0 * n * q + tid] = pd;
      sP[1 * n * q + tid] = pb;
      sP[2 * n * q + tid] = pm;
    }
    __syncthreads();
    // x pass: thread (b, a); one value per cell dof of the z layer c
    if (tid < n * n) {
      const int b = tid / n, a = tid - b * n;
      double v = 0;
      for (int i = 0; i < q; ++i) {
        const int o = b * q + i;
        v = fma(sD2[i * n + a], sP[0 * n * q + o], v);
        v = fma(sB2[i * n + a], sP[1 * n * q + o], v);
        v = fma(sM[i * n + a], sP[2 * n * q + o], v);
      }
      const int32_t id = cells[e * (int64_t)n * n * n + (int64_t)c * n * n + tid];
      if (id >= 0) atomic_add(diag + id, v);
    }
    __syncthreads();  // sT / sP are rewritten for the next layer
  }
}

// 2D: one thread per cell dof (b, a); G in the caller's layout
// [cell][G00 G01 G11][qp], qp = j q + i.
__global__ void __launch_bounds__(kDiagThreads)
    k_diag2d(const double* __restrict__ G, const double* __restrict__ B,
             const double* __restrict__ Dm, const int32_t* __restrict__ cells,
             int n, int q, int64_t ncell, double* __restrict__ diag) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int nn = n * n;
  if (t >= ncell * nn) return;
  const int64_t e = t / nn;
  const int r = (int)(t - e * nn);
  const int b = r / n, a = r - b * n;
  const int qq = q * q;
  const double* Ge = G + e * 3 * qq;
  double v = 0;
  for (int j = 0; j < q; ++j) {
    const double by = B[j * n + b], dy = Dm[j * n + b];
    for (int i = 0; i < q; ++i) {
      const double bx = B[i * n + a], dx = Dm[i * n + a];
      const double gx = dx * by, gy = bx * dy;
      const int o = j * q + i;
      v += Ge[o] * gx * gx + 2.0 * Ge[qq + o] * gx * gy + Ge[2 * qq + o] * gy * gy;
    }
  }
  const int32_t id = cells[t];
  if (id >= 0) atomic_add(diag + id, v);
}

// inv[i] = 1 / d[i] in FP32; flags a non-positive or non-finite entry.
__global__ void __launch_bounds__(256)
    k_inv_diag(const double* __restrict__ d, float* __restrict__ inv, int64_t n,
               int* __restrict__ bad) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const double v = d[i];
    if (!(v > 0.0) || !isfinite(v)) {
      atomicOr(bad, 1);
      inv[i] = 0.f;
    } else {
      inv[i] = (float)(1.0 / v);
    }
  }
}

}  // namespace

hipError_t launch_diagonal(int dim, int n, int q, const double* G, const double* tB,
                           const double* tD, const int32_t* cells,
                           const int32_t* cell_block, const int64_t* bcell,
                           int64_t ncell, double* diag, hipStream_t s) {
  if (ncell <= 0) return hipSuccess;
  if (n > kMaxN || q > kMaxQ || q * q > kDiagThreads || q * n > kDiagThreads)
    return hipErrorNotSupported;
  if (dim == 3) {
    if (ncell > 0x7fffffffLL) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_diag3d, dim3((unsigned)ncell), dim3(kDiagThreads), 0, s, G, tB, tD,
                       cells, cell_block, bcell, n, q, ncell, diag);
  } else {
    const int64_t nt = ncell * n * n;
    const int64_t g = (nt + kDiagThreads - 1) / kDiagThreads;
    if (g > 0x7fffffffLL) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_diag2d, dim3((unsigned)g), dim3(kDiagThreads), 0, s, G, tB, tD,
                       cells, n, q, ncell, diag);
  }
  return hipGetLastError();
}

hipError_t launch_inv_diag(const double* d, float* inv, int64_t n, int* bad, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  int64_t g = (n + 255) / 256;
  if (g > kStreamBlocks) g = kStreamBlocks;
  hipLaunchKernelGGL(k_inv_diag, dim3((unsigned)g), dim3(256), 0, s, d, inv, n, bad);
  return hipGetLastError();
}

}  // namespace sf
