// sf_geom.hip -- one-time geometry set-up on the device (FP64).
//
// From the 2^D vertices of each cell (multilinear map, vertex order
// tensor-lexicographic) compute at every tensor quadrature point
//     G = w_q |det J| (J^T J)^{-1} = (w_q / |det J|) C^T C,
// C the cofactor matrix of J = dx/dxi, so no division beyond 1/|det J|.
// The reference forms the same product per call from wq, jq = J^{-T} and
// djq = |det J| (src/LazyMatrices.jl:37-41, src/SumFactorizationKernels.jl:
// 191-210); here it is formed once and stored as D(D+1)/2 entries.
#include "sf_apply.hpp"
#include "sf_common.hpp"
#include "sf_kernels.h"

namespace sf {

// G is written in the block-interleaved layout of the 3D apply (g3_index).
__global__ void k_geom3d(const double* __restrict__ X,
                         const double* __restrict__ pts,
                         const double* __restrict__ w1d, int q, int64_t ncell,
                         const int32_t* __restrict__ cell_block,
                         const int64_t* __restrict__ bcell,
                         double* __restrict__ G, int* __restrict__ bad) {
  const int64_t qqq = (int64_t)q * q * q;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ncell * qqq) return;
  const int64_t e = t / qqq;
  const int r = (int)(t - e * qqq);
  const int i = r % q, k = (r / q) % q, l = r / (q * q);
  const double xi0 = pts[i], xi1 = pts[k], xi2 = pts[l];
  const double w = w1d[i] * w1d[k] * w1d[l];
  const double* Xe = X + e * 24;
  double J[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
#pragma unroll
  for (int v = 0; v < 8; ++v) {
    const int b0 = v & 1, b1 = (v >> 1) & 1, b2 = (v >> 2) & 1;
    const double s0 = b0 ? xi0 : 1.0 - xi0, d0 = b0 ? 1.0 : -1.0;
    const double s1 = b1 ? xi1 : 1.0 - xi1, d1 = b1 ? 1.0 : -1.0;
    const double s2 = b2 ? xi2 : 1.0 - xi2, d2 = b2 ? 1.0 : -1.0;
    const double dn0 = d0 * s1 * s2, dn1 = s0 * d1 * s2, dn2 = s0 * s1 * d2;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const double xa = Xe[v * 3 + a];
      J[a][0] += xa * dn0;
      J[a][1] += xa * dn1;
      J[a][2] += xa * dn2;
    }
  }
  // cofactors C[a][b] = (-1)^(a+b) minor(a, b)
  double C[3][3];
  C[0][0] = J[1][1] * J[2][2] - J[1][2] * J[2][1];
  C[0][1] = J[1][2] * J[2][0] - J[1][0] * J[2][2];
  C[0][2] = J[1][0] * J[2][1] - J[1][1] * J[2][0];
  C[1][0] = J[0][2] * J[2][1] - J[0][1] * J[2][2];
  C[1][1] = J[0][0] * J[2][2] - J[0][2] * J[2][0];
  C[1][2] = J[0][1] * J[2][0] - J[0][0] * J[2][1];
  C[2][0] = J[0][1] * J[1][2] - J[0][2] * J[1][1];
  C[2][1] = J[0][2] * J[1][0] - J[0][0] * J[1][2];
  C[2][2] = J[0][0] * J[1][1] - J[0][1] * J[1][0];
  const double det = J[0][0] * C[0][0] + J[0][1] * C[0][1] + J[0][2] * C[0][2];
  if (!(det > 0.0) || !isfinite(det)) atomicOr(bad, 1);
  const double s = w / fabs(det);
  double M[6];
  const int pa[6] = {0, 0, 0, 1, 1, 2}, pb[6] = {0, 1, 2, 1, 2, 2};
#pragma unroll
  for (int c = 0; c < 6; ++c) {
    const int a = pa[c], b = pb[c];
    M[c] = s * (C[0][a] * C[0][b] + C[1][a] * C[1][b] + C[2][a] * C[2][b]);
  }
  const int cb = cell_block[e];
  const int64_t bs = bcell[cb];
  const int ncb = (int)(bcell[cb + 1] - bs);
  const int rr = r % (q * q);
#pragma unroll
  for (int c = 0; c < 6; ++c) G[g3_index(bs, ncb, e, c, l, rr, q)] = M[c];
}

__global__ void k_geom2d(const double* __restrict__ X,
                         const double* __restrict__ pts,
                         const double* __restrict__ w1d, int q, int64_t ncell,
                         double* __restrict__ G, int* __restrict__ bad) {
  const int64_t qq = (int64_t)q * q;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ncell * qq) return;
  const int64_t e = t / qq;
  const int r = (int)(t - e * qq);
  const int i = r % q, k = r / q;
  const double xi0 = pts[i], xi1 = pts[k];
  const double w = w1d[i] * w1d[k];
  const double* Xe = X + e * 8;
  double J[2][2] = {{0, 0}, {0, 0}};
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const int b0 = v & 1, b1 = (v >> 1) & 1;
    const double s0 = b0 ? xi0 : 1.0 - xi0, d0 = b0 ? 1.0 : -1.0;
    const double s1 = b1 ? xi1 : 1.0 - xi1, d1 = b1 ? 1.0 : -1.0;
    const double dn0 = d0 * s1, dn1 = s0 * d1;
#pragma unroll
    for (int a = 0; a < 2; ++a) {
      const double xa = Xe[v * 2 + a];
      J[a][0] += xa * dn0;
      J[a][1] += xa * dn1;
    }
  }
  // C = cofactor matrix: [[J11, -J10], [-J01, J00]]
  const double c00 = J[1][1], c01 = -J[1][0], c10 = -J[0][1], c11 = J[0][0];
  const double det = J[0][0] * J[1][1] - J[0][1] * J[1][0];
  if (!(det > 0.0) || !isfinite(det)) atomicOr(bad, 1);
  const double s = w / fabs(det);
  double* Ge = G + e * 3 * qq + r;
  Ge[0] = s * (c00 * c00 + c10 * c10);
  Ge[qq] = s * (c00 * c01 + c10 * c11);
  Ge[2 * qq] = s * (c01 * c01 + c11 * c11);
}

// 3D: between the caller's layout [cell][entry][qp] and the block-
// interleaved one (to_int = 1: abi -> int, 0: int -> abi).
__global__ void k_g3_relayout(const double* __restrict__ src,
                              double* __restrict__ dst, int q, int64_t ncell,
                              const int32_t* __restrict__ cell_block,
                              const int64_t* __restrict__ bcell, int to_int) {
  const int64_t qqq = (int64_t)q * q * q;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ncell * 6 * qqq) return;
  const int64_t e = t / (6 * qqq);
  const int c = (int)((t / qqq) % 6);
  const int qp = (int)(t % qqq);
  const int cb = cell_block[e];
  const int64_t bs = bcell[cb];
  const int ncb = (int)(bcell[cb + 1] - bs);
  const int64_t gi = g3_index(bs, ncb, e, c, qp / (q * q), qp % (q * q), q);
  if (to_int)
    dst[gi] = src[t];
  else
    dst[t] = src[gi];
}

// 3D: FP64 block-interleaved G -> reduced-precision pair-interleaved G.
template <typename TG>
__device__ inline TG to_reduced(double v);
template <>
__device__ inline float to_reduced<float>(double v) { return (float)v; }
template <>
__device__ inline bf16s to_reduced<bf16s>(double v) {
  bf16s h;
  h.bits = f32_to_bf16_bits((float)v);
  return h;
}

template <typename TG>
__global__ void k_g3_to_pairs(const double* __restrict__ G, TG* __restrict__ out,
                              int q, int64_t ncell,
                              const int32_t* __restrict__ cell_block,
                              const int64_t* __restrict__ bcell) {
  const int64_t qqq = (int64_t)q * q * q;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ncell * 6 * qqq) return;
  const int64_t e = t / (6 * qqq);
  const int c = (int)((t / qqq) % 6);
  const int qp = (int)(t % qqq);
  const int cb = cell_block[e];
  const int64_t bs = bcell[cb];
  const int ncb = (int)(bcell[cb + 1] - bs);
  const int l = qp / (q * q), r = qp % (q * q);
  out[g3_pair_index(bs, ncb, e, c, l, r, q)] =
      to_reduced<TG>(G[g3_index(bs, ncb, e, c, l, r, q)]);
}

// 3D: Gk = kappa * G in the block-interleaved layout, kappa [cell][qp].
__global__ void k_scale_kappa3(const double* __restrict__ G,
                               const double* __restrict__ kappa, int q,
                               int64_t ncell,
                               const int32_t* __restrict__ cell_block,
                               const int64_t* __restrict__ bcell,
                               double* __restrict__ Gk) {
  const int64_t qqq = (int64_t)q * q * q;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ncell * qqq) return;
  const int64_t e = t / qqq;
  const int qp = (int)(t % qqq);
  const int cb = cell_block[e];
  const int64_t bs = bcell[cb];
  const int ncb = (int)(bcell[cb + 1] - bs);
  const double k = kappa[t];
#pragma unroll
  for (int c = 0; c < 6; ++c) {
    const int64_t gi = g3_index(bs, ncb, e, c, qp / (q * q), qp % (q * q), q);
    Gk[gi] = k * G[gi];
  }
}

// 2D: Gk[cell][c][qp] = kappa[cell][qp] * G[cell][c][qp]
__global__ void k_scale_kappa(const double* __restrict__ G,
                              const double* __restrict__ kappa, int ng,
                              int64_t nqc, int64_t ncell,
                              double* __restrict__ Gk) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ncell * ng * nqc) return;
  const int64_t e = t / (ng * nqc);
  const int64_t r = t % nqc;
  Gk[t] = kappa[e * nqc + r] * G[t];
}

// Finite, positive kappa check.
__global__ void k_check_positive(const double* __restrict__ v, int64_t n,
                                 int* __restrict__ bad) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const double a = v[t];
  if (!(a > 0.0) || !isfinite(a)) atomicOr(bad, 1);
}

__global__ void k_to_f32(const double* __restrict__ a, int64_t n,
                         float* __restrict__ b) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) b[t] = (float)a[t];
}

__global__ void k_to_bf16(const double* __restrict__ a, int64_t n,
                          bf16s* __restrict__ b) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) b[t].bits = f32_to_bf16_bits((float)a[t]);
}

static inline unsigned grid_for(int64_t n, int bs) {
  return (unsigned)((n + bs - 1) / bs);
}

hipError_t launch_geometry(int dim, const double* X, const double* pts,
                           const double* w1d, int q, int64_t ncell,
                           const int32_t* cell_block, const int64_t* bcell,
                           double* G, int* bad, hipStream_t s) {
  const int bs = 256;
  if (dim == 3) {
    const int64_t n = ncell * q * q * q;
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_geom3d, dim3(grid_for(n, bs)), dim3(bs), 0, s, X, pts,
                       w1d, q, ncell, cell_block, bcell, G, bad);
  } else {
    const int64_t n = ncell * q * q;
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_geom2d, dim3(grid_for(n, bs)), dim3(bs), 0, s, X, pts,
                       w1d, q, ncell, G, bad);
  }
  return hipGetLastError();
}

hipError_t launch_g3_relayout(const double* src, double* dst, int q,
                              int64_t ncell, const int32_t* cell_block,
                              const int64_t* bcell, int to_int,
                              hipStream_t s) {
  const int64_t n = ncell * 6 * q * q * q;
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_g3_relayout, dim3(grid_for(n, 256)), dim3(256), 0, s,
                     src, dst, q, ncell, cell_block, bcell, to_int);
  return hipGetLastError();
}

hipError_t launch_scale_kappa3(const double* G, const double* kappa, int q,
                               int64_t ncell, const int32_t* cell_block,
                               const int64_t* bcell, double* Gk,
                               hipStream_t s) {
  const int64_t n = ncell * q * q * q;
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_scale_kappa3, dim3(grid_for(n, 256)), dim3(256), 0, s,
                     G, kappa, q, ncell, cell_block, bcell, Gk);
  return hipGetLastError();
}

hipError_t launch_scale_kappa(const double* G, const double* kappa, int ng,
                              int64_t nqc, int64_t ncell, double* Gk,
                              hipStream_t s) {
  const int64_t n = ncell * ng * nqc;
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_scale_kappa, dim3(grid_for(n, 256)), dim3(256), 0, s, G,
                     kappa, ng, nqc, ncell, Gk);
  return hipGetLastError();
}

hipError_t launch_check_positive(const double* v, int64_t n, int* bad,
                                 hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_check_positive, dim3(grid_for(n, 256)), dim3(256), 0, s,
                     v, n, bad);
  return hipGetLastError();
}

hipError_t launch_g3_to_f32_pairs(const double* G, float* out, int q,
                                  int64_t ncell, const int32_t* cell_block,
                                  const int64_t* bcell, hipStream_t s) {
  const int64_t n = ncell * 6 * q * q * q;
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_g3_to_pairs<float>, dim3(grid_for(n, 256)), dim3(256), 0,
                     s, G, out, q, ncell, cell_block, bcell);
  return hipGetLastError();
}

hipError_t launch_g3_to_bf16_pairs(const double* G, bf16s* out, int q,
                                   int64_t ncell, const int32_t* cell_block,
                                   const int64_t* bcell, hipStream_t s) {
  const int64_t n = ncell * 6 * q * q * q;
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_g3_to_pairs<bf16s>, dim3(grid_for(n, 256)), dim3(256), 0,
                     s, G, out, q, ncell, cell_block, bcell);
  return hipGetLastError();
}

hipError_t launch_to_f32(const double* a, int64_t n, float* b, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_to_f32, dim3(grid_for(n, 256)), dim3(256), 0, s, a, n,
                     b);
  return hipGetLastError();
}

hipError_t launch_to_bf16(const double* a, int64_t n, bf16s* b,
                          hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_to_bf16, dim3(grid_for(n, 256)), dim3(256), 0, s, a, n,
                     b);
  return hipGetLastError();
}

}  // namespace sf
