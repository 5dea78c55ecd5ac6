// sf_precond.hip -- the Gauss-Lobatto-basis Jacobi preconditioner of the
// device-resident CG solvers (3D, block structure of the apply).
//
// The reference's Lagrange basis sits on equispaced nodes [0, 1, 1/p, ...]
// (src/SumFactorizationMaps.jl:47-65).  At high order that basis is badly
// conditioned, and neither plain CG (the reference's profiling/drivers.jl:
// 85-95) nor Jacobi on its diagonal converges in useful time at Q9.  The
// same space has a well-conditioned nodal basis on the Gauss-Lobatto (GLL)
// points.  T maps GLL nodal values to equispaced nodal values: per cell the
// tensor product T1 (x) T1 (x) T1 with T1[a][b] = l_b(t_a), l_b the GLL
// Lagrange polynomials and t_a the equispaced nodes; it is conforming (the
// rows of T1 at the vertex nodes are unit rows, so a dof on a face, edge or
// vertex only sees the GLL dofs of that entity, which every cell holding
// the dof also holds).  The stiffness in the GLL basis is T^T A T, and its
// diagonal is the diagonal of the same operator with the 1D tables B T1,
// D T1 (computed by the diagonal kernel).  The preconditioner is
//     z = M^{-1} r = T diag(T^T A T)^{-1} T^T r,
// symmetric positive definite.  Its two tensor passes are:
//   * k_pc_gll<.., TRANS = true>:  w = T^T (minv r), minv = 1 / (number of
//     cells holding the dof): every cell holding dof i and j holds the
//     entity of i, so the per-cell products T_e^T (r_e / mult) add up to
//     exactly T^T r;
//   * k_pc_gll<.., TRANS = false>: z = T (dginv w): every cell holding a
//     dof computes the same value for it (the unit vertex rows make the
//     contraction along a face normal an exact copy), so each block stores
//     its dofs' values with plain stores.  (scale may be null: 1.)
// The solvers do not apply M^{-1} per iteration: PCG with this M on A is CG
// with the Jacobi preconditioner on T^T A T, so they run in the GLL basis
// (the apply with the GLL tables) and use these passes only to form T^T b,
// T^{-1} x0 (the forward pass with T1^{-1}, whose vertex rows are unit rows
// too) and x = T y; sf_precond_apply runs both passes.
// Each pass: one workgroup per block of the apply, the block's distinct
// dofs gathered once (sorted ids, coalesced), three 1D contractions per
// cell through LDS (one N-long pencil per thread), and one global access
// per distinct dof on the way out (the transposed pass adds its shared
// dofs with atomics into a zeroed w).  The LDS per block (its dof values
// and one tensor per cell) is what bounds the occupancy of these light
// passes.
#include "sf_apply.hpp"
#include "sf_common.hpp"
#include "sf_kernels.h"

namespace sf {

namespace {

// Cells of the block run side by side: thread t owns pencil t % N^2 of
// cell t / N^2 in each pass.  LDS (dynamic): the block's nu distinct-dof
// values (the gathered inputs, later the outputs), then one N x N x NP
// tensor per cell (x stride padded to an odd NP: the x pass walks a
// pencil per thread).  The y pass rewrites its own pencil in place (a
// thread reads and writes only its pencil's N positions), so one tensor
// per cell suffices.  The vertex rows of T1 are unit rows, so part of each
// contraction is a copy (forward) or skips the two vertex inputs (TRANS).
template <int N, typename TR, typename TZ, bool TRANS>
__global__ void __launch_bounds__(1024)
    k_pc_gll(const TR* __restrict__ in, const float* __restrict__ scale,
             TZ* __restrict__ out, const double* __restrict__ M,
             const uint16_t* __restrict__ pidx, const int32_t* __restrict__ bdofs,
             const int64_t* __restrict__ boff, const int32_t* __restrict__ bexcl,
             const int64_t* __restrict__ bcell, int nblk, int eb,
             const CGScalars* __restrict__ sc) {
  constexpr int NN = N * N, NNN = N * N * N, NP = N | 1, CS = NN * NP;
  if (sc && sc->done) return;  // uniform: written only by other kernels
  extern __shared__ double pc_smem[];
  TZ* const V = (TZ*)pc_smem;
  const int b = xcd_remap(blockIdx.x, nblk);
  const int64_t base = bcell[b];
  const int ncb = (int)(bcell[b + 1] - base);
  const int64_t d0 = boff[b];
  const int nu = (int)(boff[b + 1] - d0);
  const int nexcl = bexcl[b];
  const int32_t* __restrict__ dofs = bdofs + d0;
  TZ* const S0 = V + ((nu + 1) & ~1);  // 8-byte aligned for FP32 too
  const int tid = threadIdx.x, nt = blockDim.x;
  const int e = tid / NN, r = tid - e * NN;
  const bool act = e < ncb && e < eb;
  TZ* const A = S0 + (size_t)(act ? e : 0) * CS;
  const uint16_t* lp = pidx + (base + (act ? e : 0)) * NNN;
  const int c1 = r / N, c0 = r - c1 * N;  // pencil coordinates (slow, fast)
  // the local indices of this thread's x pencil (z, y) = (c1, c0) and z
  // pencil (y, x) = (c1, c0), loaded with the gather so that their latency
  // overlaps it
  uint16_t lx[N], lz[N];
#pragma unroll
  for (int i = 0; i < N; ++i) {
    lx[i] = act ? __builtin_nontemporal_load(lp + r * N + i) : kNoDof;
    lz[i] = act ? __builtin_nontemporal_load(lp + (i * N + c1) * N + c0) : kNoDof;
  }
  // the block's nu <= eb N^3 distinct dofs, at most N per thread (the
  // launch has nt >= eb N^2 threads): all loads issued at once
  int32_t dv[N];
#pragma unroll
  for (int k = 0; k < N; ++k) {
    const int i = tid + k * nt;
    dv[k] = i < nu ? __builtin_nontemporal_load(dofs + i) : -1;
  }
  TZ xv[N];
#pragma unroll
  for (int k = 0; k < N; ++k)
    xv[k] = dv[k] >= 0 ? (TZ)((double)in[dv[k]] * (scale ? (double)scale[dv[k]] : 1.0))
                       : TZ(0);
#pragma unroll
  for (int k = 0; k < N; ++k)
    if (dv[k] >= 0) V[tid + k * nt] = xv[k];
  __syncthreads();
  // out[o] = sum_i M[o][i] u[i] (TRANS: M = T1^T, else M = T1)
  auto contract = [&](const TZ (&u)[N], TZ (&w)[N]) {
#pragma unroll
    for (int o = 0; o < N; ++o) {
      if (!TRANS && o < 2) {
        w[o] = u[o];  // unit rows of T1
        continue;
      }
      TZ v = TRANS && o < 2 ? u[o] : TZ(0);
#pragma unroll
      for (int i = TRANS ? 2 : 0; i < N; ++i) v = fma((TZ)M[o * N + i], u[i], v);
      w[o] = v;
    }
  };
  // x pass: pencil (z, y) = (c1, c0), gathered straight from V
  if (act) {
    TZ u[N], w[N];
#pragma unroll
    for (int i = 0; i < N; ++i) u[i] = lx[i] != kNoDof ? V[lx[i]] : TZ(0);
    contract(u, w);
#pragma unroll
    for (int o = 0; o < N; ++o) A[r * NP + o] = w[o];
  }
  __syncthreads();
  if (TRANS) {  // V is read: it becomes the accumulator
#pragma unroll
    for (int k = 0; k < N; ++k)
      if (dv[k] >= 0) V[tid + k * nt] = TZ(0);
  }
  // y pass, in place: pencil (z, x) = (c1, c0)
  if (act) {
    TZ u[N], w[N];
#pragma unroll
    for (int i = 0; i < N; ++i) u[i] = A[(c1 * N + i) * NP + c0];
    contract(u, w);
#pragma unroll
    for (int o = 0; o < N; ++o) A[(c1 * N + o) * NP + c0] = w[o];
  }
  __syncthreads();
  // z pass: pencil (y, x) = (c1, c0), into the block's values
  if (act) {
    TZ u[N], w[N];
#pragma unroll
    for (int i = 0; i < N; ++i) u[i] = A[(i * N + c1) * NP + c0];
    contract(u, w);
#pragma unroll
    for (int o = 0; o < N; ++o) {
      if (lz[o] != kNoDof) {
        if (TRANS)
          lds_atomic_add(V + lz[o], w[o]);
        else
          V[lz[o]] = w[o];  // every cell holding the dof computes the same value
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < N; ++k) {
    const int i = tid + k * nt;
    if (dv[k] < 0) continue;
    if (!TRANS || i < nexcl)
      out[dv[k]] = V[i];
    else
      atomic_add(out + dv[k], V[i]);
  }
}

// scale[d] = 1 / (number of cells holding free dof d): cnt counts them.
__global__ void __launch_bounds__(256)
    k_count_cells(const int32_t* __restrict__ cells, int64_t n, float* __restrict__ cnt) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int32_t d = cells[i];
    if (d >= 0) atomicAdd(cnt + d, 1.0f);
  }
}

__global__ void __launch_bounds__(256) k_recip(float* __restrict__ a, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    a[i] = a[i] > 0.f ? 1.0f / a[i] : 0.f;
}

__global__ void __launch_bounds__(256)
    k_scale_vec(const double* __restrict__ r, const float* __restrict__ d,
                double* __restrict__ z, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    z[i] = (double)d[i] * r[i];
}

inline unsigned grid_for(int64_t n) {
  int64_t g = (n + 255) / 256;
  if (g > kStreamBlocks) g = kStreamBlocks;
  if (g < 1) g = 1;
  return (unsigned)g;
}

template <typename TR, typename TZ, bool TRANS>
hipError_t launch_pass(int n, const TR* in, const float* scale, TZ* out, const double* M,
                       const BlockArgs& k, int eb, int max_nu, const CGScalars* sc,
                       hipStream_t s) {
  if (k.nblk <= 0) return hipSuccess;
  const int nt = round_up(eb * n * n, 64);
  const size_t lds = (size_t)(((max_nu + 1) & ~1) + eb * n * n * (n | 1)) * sizeof(TZ);
  // every thread gathers at most n of the block's distinct dofs
  if (nt > 1024 || lds > 65536 || max_nu > n * nt) return hipErrorNotSupported;
  switch (n) {
#define SF_PC_CASE(N_)                                                                     \
  case N_:                                                                                 \
    hipLaunchKernelGGL((k_pc_gll<N_, TR, TZ, TRANS>), dim3(k.nblk), dim3(nt), lds, s, in, \
                       scale, out, M, k.pidx, k.bdofs, k.boff, k.bexcl, k.bcell, k.nblk,   \
                       eb, sc);                                                            \
    break;
    SF_PC_CASE(2) SF_PC_CASE(3) SF_PC_CASE(4) SF_PC_CASE(5) SF_PC_CASE(6)
    SF_PC_CASE(7) SF_PC_CASE(8) SF_PC_CASE(9) SF_PC_CASE(10)
#undef SF_PC_CASE
    default:
      return hipErrorNotSupported;
  }
  return hipGetLastError();
}

}  // namespace

hipError_t launch_mult_inverse(const int32_t* cells, int64_t ncell_dofs, float* minv,
                               int64_t nfree, hipStream_t s) {
  if (nfree <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_count_cells, dim3(grid_for(ncell_dofs)), dim3(256), 0, s, cells,
                     ncell_dofs, minv);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_recip, dim3(grid_for(nfree)), dim3(256), 0, s, minv, nfree);
  return hipGetLastError();
}

hipError_t launch_scale_vec(const double* r, const float* dinv, double* z, int64_t n,
                            hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_scale_vec, dim3(grid_for(n)), dim3(256), 0, s, r, dinv, z, n);
  return hipGetLastError();
}

template <typename TR, typename TZ>
hipError_t launch_pc_gll(int n, const TR* r, TZ* z, TZ* w, int64_t nfree, const float* minv,
                         const float* dginv, const double* T1, const double* T1t,
                         const BlockArgs& k, int eb, int max_nu, const CGScalars* sc,
                         hipStream_t s) {
  if (nfree <= 0) return hipSuccess;
  hipError_t e = launch_fill<TZ>(w, nfree, TZ(0), s);
  if (e != hipSuccess) return e;
  e = launch_pass<TR, TZ, true>(n, r, minv, w, T1t, k, eb, max_nu, sc, s);
  if (e != hipSuccess) return e;
  return launch_pass<TZ, TZ, false>(n, w, dginv, z, T1, k, eb, max_nu, sc, s);
}

template hipError_t launch_pc_gll<double, double>(int, const double*, double*, double*, int64_t,
                                                  const float*, const float*, const double*,
                                                  const double*, const BlockArgs&, int, int,
                                                  const CGScalars*, hipStream_t);

hipError_t launch_gll_pass(int n, bool trans, const double* in, const float* scale, double* out,
                           const double* M, int64_t nfree, const BlockArgs& k, int eb,
                           int max_nu, hipStream_t s) {
  if (nfree <= 0) return hipSuccess;
  if (trans) {
    hipError_t e = launch_fill<double>(out, nfree, 0.0, s);
    if (e != hipSuccess) return e;
    return launch_pass<double, double, true>(n, in, scale, out, M, k, eb, max_nu, nullptr, s);
  }
  return launch_pass<double, double, false>(n, in, scale, out, M, k, eb, max_nu, nullptr, s);
}

}  // namespace sf
