// sf_blas.hip -- BLAS-1 pieces of the device-resident CG / iterative
// refinement.  Every reduction accumulates in FP64 (also for FP32 vectors)
// and the CG scalars alpha, beta, r.r stay in device memory.  The
// reductions of a CG step run in two launches: the streaming grid stores
// one share per workgroup, then one workgroup sums the shares and updates
// the scalars.  The stand-alone dot product and the fused step's p.Ap
// (a few thousand values) finish in the last workgroup to arrive instead.
// With the Jacobi preconditioner every piece also takes dinv = 1 / diag(A)
// (FP32) and carries r.z next to r.r.
#include "sf_common.hpp"
#include "sf_kernels.h"

namespace sf {

namespace {

// Block sum over NT threads (kRedThreads unless given); result valid on
// thread 0.
template <int NT = kRedThreads>
__device__ double block_sum(double v) {
  __shared__ double ws[NT / 64];
  v = wave_sum_f64(v);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();  // ws may still be read by a previous call
  if (lane == 0) ws[wid] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int w = 0; w < NT / 64; ++w) s += ws[w];
  }
  return s;
}

// Threads of the one-workgroup CG finish: the shares of a streaming grid
// (up to 4096 per sum) are read four per thread in one round of loads,
// where 256 threads needed 16 dependent rounds.
constexpr int kFinThreads = 1024;

// Publish this block's partial; true in every thread of the last block.
__device__ bool arrive(Reducer red, double partial) {
  __shared__ int am_last;
  if (threadIdx.x == 0) {
    red.part[blockIdx.x] = partial;
    __threadfence();
    const unsigned prev = atomicAdd(red.counter, 1u);
    am_last = (prev == gridDim.x - 1);
    if (am_last) __threadfence();
  }
  __syncthreads();
  return am_last != 0;
}

// Sum of all partials, called by every thread of the last block; valid on
// thread 0, which also resets the arrival counter.
__device__ double gather_partials(Reducer red) {
  double s = 0.0;
  for (unsigned i = threadIdx.x; i < gridDim.x; i += blockDim.x)
    s += __hip_atomic_load(red.part + i, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
  s = block_sum(s);
  if (threadIdx.x == 0) *red.counter = 0u;
  return s;
}

template <typename V>
__global__ void __launch_bounds__(256) k_fill(V* __restrict__ a, int64_t n,
                                              V v) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    a[i] = v;
}

template <typename V>
__global__ void __launch_bounds__(kRedThreads)
    k_dot(const V* __restrict__ a, const V* __restrict__ b, int64_t n,
          Reducer red, double* __restrict__ out) {
  double s = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    s += (double)a[i] * (double)b[i];
  s = block_sum(s);
  if (arrive(red, s)) {
    const double tot = gather_partials(red);
    if (threadIdx.x == 0) *out = tot;
  }
}

// ||b - t||^2 (the true residual's norm without storing it).
__global__ void __launch_bounds__(kRedThreads)
    k_diff_norm2(const double* __restrict__ b, const double* __restrict__ t, int64_t n,
                 Reducer red, double* __restrict__ out) {
  double s = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const double d = b[i] - t[i];
    s += d * d;
  }
  s = block_sum(s);
  if (arrive(red, s)) {
    const double tot = gather_partials(red);
    if (threadIdx.x == 0) *out = tot;
  }
}

// a . (w a) with an FP32 weight (r.z of the Jacobi preconditioner).
template <typename V>
__global__ void __launch_bounds__(kRedThreads)
    k_wdot(const V* __restrict__ a, const float* __restrict__ w, int64_t n, Reducer red,
           double* __restrict__ out) {
  double s = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const double ai = (double)a[i];
    s += ai * ((double)w[i] * ai);
  }
  s = block_sum(s);
  if (arrive(red, s)) {
    const double tot = gather_partials(red);
    if (threadIdx.x == 0) *out = tot;
  }
}

// The Jacobi preconditioner z = dinv r (dinv = 1 / diag(A) in FP32, one
// vector for every solver form); PC = false: z = r.
template <bool PC, typename V>
__device__ __forceinline__ V precond(const float* __restrict__ dinv, int64_t i, V r) {
  if constexpr (PC) return (V)((double)dinv[i] * (double)r);
  return r;
}

// r = b - Ax, p = z, q = 0 on the streaming grid; workgroup shares of r.r
// and (PC) r.z into part[blk] and part[g + blk] (finished by
// k_cg_start_finish).
template <typename V, bool PC>
__global__ void __launch_bounds__(256)
    k_cg_start(const V* __restrict__ b, const V* __restrict__ Ax, V* __restrict__ r,
               V* __restrict__ p, V* __restrict__ q, const float* __restrict__ dinv,
               int64_t n, double* __restrict__ part) {
  double s = 0.0, sz = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const V ri = b[i] - Ax[i];
    const V zi = precond<PC>(dinv, i, ri);
    r[i] = ri;
    p[i] = zi;
    q[i] = V(0);
    s += (double)ri * (double)ri;
    if constexpr (PC) sz += (double)ri * (double)zi;
  }
  s = block_sum(s);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
  if constexpr (PC) {
    sz = block_sum(sz);
    if (threadIdx.x == 0) part[gridDim.x + blockIdx.x] = sz;
  }
}

// One workgroup: res2 = r.r, rr = r.z (= r.r without a preconditioner).
__global__ void __launch_bounds__(kRedThreads)
    k_cg_start_finish(const double* __restrict__ part, int np, int pc,
                      CGScalars* __restrict__ sc, double stop_rr) {
  double s = 0.0, sz = 0.0;
  for (int i = threadIdx.x; i < np; i += blockDim.x) {
    s += part[i];
    if (pc) sz += part[np + i];
  }
  s = block_sum(s);
  if (pc) sz = block_sum(sz);
  if (threadIdx.x == 0) {
    const double rz = pc ? sz : s;
    sc->res2 = s;
    sc->rr = rz;
    sc->rr_new = rz;
    sc->pq = 0.0;
    sc->alpha = 0.0;
    sc->beta = 0.0;
    sc->stop_rr = stop_rr;
    sc->iters = 0;
    sc->xpend = 0;
    const bool fin = isfinite(s) && isfinite(rz);
    sc->nonfinite = fin ? 0 : 1;
    sc->done = (s <= stop_rr || !fin) ? 1 : 0;
  }
}

// p.q on the streaming grid, per-workgroup shares (finished by
// k_cg_alpha_finish; see k_cg_resid for why there is no arrival counter).
template <typename V, typename W = V>
__global__ void __launch_bounds__(256)
    k_cg_alpha(const V* __restrict__ p, const W* __restrict__ q, int64_t n,
               double* __restrict__ part, const CGScalars* __restrict__ sc) {
  if (sc->done) return;  // uniform over the grid: set only by other kernels
  double s = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    s += (double)p[i] * (double)q[i];
  s = block_sum(s);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// One workgroup: pq = sum of the np shares, alpha = rr / pq.
__global__ void __launch_bounds__(kRedThreads)
    k_cg_alpha_finish(const double* __restrict__ part, int np,
                      CGScalars* __restrict__ sc) {
  if (sc->done) return;
  double s = 0.0;
  for (int i = threadIdx.x; i < np; i += blockDim.x) s += part[i];
  s = block_sum(s);
  if (threadIdx.x == 0) {
    sc->pq = s;
    const double a = sc->rr / s;
    sc->alpha = a;
    if (!isfinite(a)) {
      sc->nonfinite = 1;
      sc->done = 1;
    }
  }
}

// x += alpha p, r -= alpha q on the streaming grid, per-workgroup shares of
// r.r and (PC) r.z (finished by k_cg_resid_finish).
template <typename V, typename W, bool PC>
__global__ void __launch_bounds__(256)
    k_cg_update(V* __restrict__ x, V* __restrict__ r,
                const V* __restrict__ p, const W* __restrict__ q,
                const float* __restrict__ dinv, int64_t n,
                double* __restrict__ part, const CGScalars* __restrict__ sc) {
  if (sc->done) return;
  const V a = (V)sc->alpha;
  double s = 0.0, sz = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    __builtin_nontemporal_store(fma(a, p[i], x[i]), x + i);
    const V ri = fma(-a, (V)q[i], r[i]);
    __builtin_nontemporal_store(ri, r + i);
    s += (double)ri * (double)ri;
    if constexpr (PC) sz += (double)ri * (double)precond<PC>(dinv, i, ri);
  }
  s = block_sum(s);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
  if constexpr (PC) {
    sz = block_sum(sz);
    if (threadIdx.x == 0) part[gridDim.x + blockIdx.x] = sz;
  }
}

// ---- fused CG step (3D): the apply returns per-wave shares of p.Ap,
// so one step is  apply -> alpha_parts -> resid -> xdir  and p is read
// once outside the apply:
//   alpha_parts: alpha = rr / sum(parts)     (one value per wave of the apply)
//   resid      : r -= alpha q, r.r (and r.z), beta (reads r, q; two
//                launches: per-workgroup shares, then a one-workgroup sum)
//   xdir       : x += alpha p, p = z + beta p, q = 0    (reads x, p, r)
// xpend carries "this step's x update is still to come" from resid to
// xdir, so the step that converges still moves x, and nothing moves it
// twice.
__global__ void __launch_bounds__(kRedThreads)
    k_cg_alpha_parts(const double* __restrict__ part, int64_t np, Reducer red,
                     CGScalars* __restrict__ sc) {
  if (sc->done) {  // uniform: written only by other kernels
    if (blockIdx.x == 0 && threadIdx.x == 0) sc->xpend = 0;
    return;
  }
  double s = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < np;
       i += (int64_t)gridDim.x * blockDim.x)
    s += part[i];
  s = block_sum(s);
  if (arrive(red, s)) {
    const double pq = gather_partials(red);
    if (threadIdx.x == 0) {
      sc->xpend = 0;
      sc->pq = pq;
      const double a = sc->rr / pq;
      sc->alpha = a;
      if (!isfinite(a)) {
        sc->nonfinite = 1;
        sc->done = 1;
      }
    }
  }
}

// r -= alpha q on the streaming grid; each workgroup stores its share of
// r.r (and r.z) (plain store: the finish kernel runs after the grid has
// ended, so no fence or arrival counter -- on this chip each such fence
// waits for the L2 to write back the r lines just stored).
template <typename V, typename W, bool PC>
__global__ void __launch_bounds__(256)
    k_cg_resid(V* __restrict__ r, const W* __restrict__ q, const float* __restrict__ dinv,
               int64_t n, double* __restrict__ part, const CGScalars* __restrict__ sc) {
  if (sc->done) return;
  const V a = (V)sc->alpha;
  double s = 0.0, sz = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const V ri = fma(-a, (V)q[i], r[i]);
    __builtin_nontemporal_store(ri, r + i);
    s += (double)ri * (double)ri;
    if constexpr (PC) sz += (double)ri * (double)precond<PC>(dinv, i, ri);
  }
  s = block_sum(s);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
  if constexpr (PC) {
    sz = block_sum(sz);
    if (threadIdx.x == 0) part[gridDim.x + blockIdx.x] = sz;
  }
}

// One workgroup: res2 = r.r and rr_new = r.z (r.r without a
// preconditioner) from the np shares, beta = rr_new / rr, the stop test on
// res2.
__global__ void __launch_bounds__(kFinThreads)
    k_cg_resid_finish(const double* __restrict__ part, int np, int pc,
                      CGScalars* __restrict__ sc) {
  if (sc->done) return;
  double s = 0.0, sz = 0.0;
#pragma unroll 4
  for (int i = threadIdx.x; i < np; i += kFinThreads) {
    s += part[i];
    if (pc) sz += part[np + i];
  }
  s = block_sum<kFinThreads>(s);
  if (pc) sz = block_sum<kFinThreads>(sz);
  if (threadIdx.x == 0) {
    const double rr_new = pc ? sz : s;
    sc->res2 = s;
    sc->rr_new = rr_new;
    sc->beta = rr_new / sc->rr;
    sc->rr = rr_new;
    sc->iters += 1;
    sc->xpend = 1;
    if (!isfinite(s) || !isfinite(rr_new)) {
      sc->nonfinite = 1;
      sc->done = 1;
    } else if (s <= sc->stop_rr) {
      sc->done = 1;
    }
  }
}

template <typename V, bool PC>
__global__ void __launch_bounds__(256)
    k_cg_xdir(V* __restrict__ x, V* __restrict__ p, const V* __restrict__ r,
              V* __restrict__ q, const float* __restrict__ dinv, int64_t n,
              const CGScalars* __restrict__ sc) {
  if (!sc->xpend) return;
  const V a = (V)sc->alpha;
  if (sc->done) {  // the converging step: x only
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
      __builtin_nontemporal_store(fma(a, p[i], x[i]), x + i);
    return;
  }
  const V b = (V)sc->beta;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const V pi = p[i];
    __builtin_nontemporal_store(fma(a, pi, x[i]), x + i);
    __builtin_nontemporal_store(fma(b, pi, precond<PC>(dinv, i, r[i])), p + i);
    __builtin_nontemporal_store(V(0), q + i);
  }
}

// The same with an FP32 operator: also the FP32 copy of p for the next
// apply, and q32 = 0.
template <bool PC>
__global__ void __launch_bounds__(256)
    k_cg_xdir_mixed(double* __restrict__ x, double* __restrict__ p,
                    float* __restrict__ p32, const double* __restrict__ r,
                    float* __restrict__ q32, const float* __restrict__ dinv, int64_t n,
                    const CGScalars* __restrict__ sc) {
  if (!sc->xpend) return;
  const double a = sc->alpha;
  if (sc->done) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
      __builtin_nontemporal_store(fma(a, p[i], x[i]), x + i);
    return;
  }
  const double b = sc->beta;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const double pi = p[i];
    __builtin_nontemporal_store(fma(a, pi, x[i]), x + i);
    const double pn = fma(b, pi, precond<PC>(dinv, i, r[i]));
    __builtin_nontemporal_store(pn, p + i);
    __builtin_nontemporal_store((float)pn, p32 + i);
    __builtin_nontemporal_store(0.f, q32 + i);
  }
}

// The same with the search direction held only in FP32 (x and r FP64):
// x += alpha p32, p32 = z + beta p32, q32 = 0.  36 bytes per dof instead of
// 48 (40 with the preconditioner).
template <bool PC>
__global__ void __launch_bounds__(256)
    k_cg_xdir_p32(double* __restrict__ x, float* __restrict__ p32,
                  const double* __restrict__ r, float* __restrict__ q32,
                  const float* __restrict__ dinv, int64_t n,
                  const CGScalars* __restrict__ sc) {
  if (!sc->xpend) return;
  const double a = sc->alpha;
  if (sc->done) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
      __builtin_nontemporal_store(fma(a, (double)p32[i], x[i]), x + i);
    return;
  }
  const double b = sc->beta;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const double pi = (double)p32[i];
    __builtin_nontemporal_store(fma(a, pi, x[i]), x + i);
    __builtin_nontemporal_store((float)fma(b, pi, precond<PC>(dinv, i, r[i])), p32 + i);
    __builtin_nontemporal_store(0.f, q32 + i);
  }
}

// p32 = z + c p32, q32 = 0: the FP32-direction form of the direction update
// after a residual replacement.
template <bool PC>
__global__ void __launch_bounds__(256)
    k_dir_coef_p32(float* __restrict__ p32, const double* __restrict__ r,
                   float* __restrict__ q32, const float* __restrict__ dinv, double c,
                   int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    p32[i] = (float)fma(c, (double)p32[i], precond<PC>(dinv, i, r[i]));
    q32[i] = 0.f;
  }
}

template <typename V, bool PC>
__global__ void __launch_bounds__(256)
    k_cg_dir(V* __restrict__ p, const V* __restrict__ r, V* __restrict__ q,
             const float* __restrict__ dinv, int64_t n, const CGScalars* __restrict__ sc) {
  if (sc->done) return;
  const V b = (V)sc->beta;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    p[i] = fma(b, p[i], precond<PC>(dinv, i, r[i]));
    q[i] = V(0);
  }
}

// FP64 recurrence with an FP32 operator: p = z + beta p (FP64), its FP32
// copy for the next apply, and q32 = 0 (the apply adds into it).
template <bool PC>
__global__ void __launch_bounds__(256)
    k_cg_dir_mixed(double* __restrict__ p, float* __restrict__ p32,
                   const double* __restrict__ r, float* __restrict__ q32,
                   const float* __restrict__ dinv, int64_t n,
                   const CGScalars* __restrict__ sc) {
  if (sc->done) return;
  const double b = sc->beta;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const double pi = fma(b, p[i], precond<PC>(dinv, i, r[i]));
    p[i] = pi;
    p32[i] = (float)pi;
    q32[i] = 0.f;
  }
}

// The same with the coefficient given by the host (after a residual
// replacement).
template <bool PC>
__global__ void __launch_bounds__(256)
    k_dir_coef_mixed(double* __restrict__ p, float* __restrict__ p32,
                     const double* __restrict__ r, float* __restrict__ q32,
                     const float* __restrict__ dinv, double c, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const double pi = fma(c, p[i], precond<PC>(dinv, i, r[i]));
    p[i] = pi;
    p32[i] = (float)pi;
    q32[i] = 0.f;
  }
}

// p = z + c p, q = 0: the direction update after a residual replacement,
// with the coefficient computed on the host.
template <bool PC>
__global__ void __launch_bounds__(256)
    k_dir_coef(float* __restrict__ p, const float* __restrict__ r,
               float* __restrict__ q, const float* __restrict__ dinv, float c, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    p[i] = fma(c, p[i], precond<PC>(dinv, i, r[i]));
    q[i] = 0.f;
  }
}

template <typename V>
__global__ void __launch_bounds__(256)
    k_zero_list(V* __restrict__ y, const int32_t* __restrict__ idx,
                int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    y[idx[i]] = V(0);
}

__global__ void __launch_bounds__(256)
    k_residual(const double* __restrict__ b, const double* __restrict__ Ax,
               double* __restrict__ r, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    r[i] = b[i] - Ax[i];
}

__global__ void __launch_bounds__(256)
    k_scale_to_f32(const double* __restrict__ r, double s,
                   float* __restrict__ r32, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    r32[i] = (float)(r[i] * s);
}

__global__ void __launch_bounds__(256)
    k_axpy_f32_to_f64(double* __restrict__ x, const float* __restrict__ d,
                      double s, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    x[i] = fma(s, (double)d[i], x[i]);
}

inline unsigned stream_grid(int64_t n) {
  // Enough workgroups to fill 256 CUs several times over; grid-stride loops
  // cover the rest.
  int64_t g = (n + 255) / 256;
  if (g > kStreamBlocks) g = kStreamBlocks;
  if (g < 1) g = 1;
  return (unsigned)g;
}

inline unsigned red_grid(int64_t n) {
  int64_t g = (n + kRedThreads - 1) / kRedThreads;
  if (g > kRedBlocks) g = kRedBlocks;
  if (g < 1) g = 1;
  return (unsigned)g;
}

}  // namespace

template <typename V>
hipError_t launch_fill(V* a, int64_t n, V v, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill<V>, dim3(stream_grid(n)), dim3(256), 0, s, a, n,
                     v);
  return hipGetLastError();
}

template <typename V>
hipError_t launch_zero_list(V* y, const int32_t* idx, int64_t n,
                            hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_zero_list<V>, dim3(stream_grid(n)), dim3(256), 0, s, y,
                     idx, n);
  return hipGetLastError();
}

template <typename V>
hipError_t launch_dot(const V* a, const V* b, int64_t n, Reducer red,
                      double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_dot<V>, dim3(red_grid(n)), dim3(kRedThreads), 0, s, a,
                     b, n, red, out);
  return hipGetLastError();
}

hipError_t launch_diff_norm2(const double* b, const double* t, int64_t n, Reducer red,
                            double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_diff_norm2, dim3(red_grid(n)), dim3(kRedThreads), 0, s, b, t, n, red,
                     out);
  return hipGetLastError();
}

template <typename V>
hipError_t launch_wdot(const V* a, const float* w, int64_t n, Reducer red, double* out,
                       hipStream_t s) {
  hipLaunchKernelGGL(k_wdot<V>, dim3(red_grid(n)), dim3(kRedThreads), 0, s, a, w, n, red,
                     out);
  return hipGetLastError();
}

template <typename V>
hipError_t launch_cg_start(const V* b, const V* Ax, V* r, V* p, V* q,
                           int64_t n, Reducer red, CGScalars* sc,
                           double stop_rr, hipStream_t s, const float* dinv) {
  const unsigned g = stream_grid(n);
  if (dinv)
    hipLaunchKernelGGL((k_cg_start<V, true>), dim3(g), dim3(256), 0, s, b, Ax, r, p, q, dinv,
                       n, red.part);
  else
    hipLaunchKernelGGL((k_cg_start<V, false>), dim3(g), dim3(256), 0, s, b, Ax, r, p, q, dinv,
                       n, red.part);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_cg_start_finish, dim3(1), dim3(kRedThreads), 0, s, red.part, (int)g,
                     dinv ? 1 : 0, sc, stop_rr);
  return hipGetLastError();
}

template <typename V, typename W>
hipError_t alpha_two_phase(const V* p, const W* q, int64_t n, Reducer red, CGScalars* sc,
                           hipStream_t s) {
  const unsigned g = stream_grid(n);
  hipLaunchKernelGGL((k_cg_alpha<V, W>), dim3(g), dim3(256), 0, s, p, q, n, red.part, sc);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_cg_alpha_finish, dim3(1), dim3(kRedThreads), 0, s, red.part,
                     (int)g, sc);
  return hipGetLastError();
}

template <typename V, typename W>
hipError_t update_two_phase(V* x, V* r, const V* p, const W* q, const float* dinv,
                            int64_t n, Reducer red, CGScalars* sc, hipStream_t s) {
  const unsigned g = stream_grid(n);
  if (dinv)
    hipLaunchKernelGGL((k_cg_update<V, W, true>), dim3(g), dim3(256), 0, s, x, r, p, q, dinv,
                       n, red.part, sc);
  else
    hipLaunchKernelGGL((k_cg_update<V, W, false>), dim3(g), dim3(256), 0, s, x, r, p, q, dinv,
                       n, red.part, sc);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_cg_resid_finish, dim3(1), dim3(kFinThreads), 0, s, red.part,
                     (int)g, dinv ? 1 : 0, sc);
  return hipGetLastError();
}

template <typename V>
hipError_t launch_cg_alpha(const V* p, const V* q, int64_t n, Reducer red,
                           CGScalars* sc, hipStream_t s) {
  return alpha_two_phase<V, V>(p, q, n, red, sc, s);
}

template <typename V>
hipError_t launch_cg_update(V* x, V* r, const V* p, const V* q, int64_t n,
                            Reducer red, CGScalars* sc, hipStream_t s, const float* dinv) {
  return update_two_phase<V, V>(x, r, p, q, dinv, n, red, sc, s);
}

template <typename V>
hipError_t launch_cg_dir(V* p, const V* r, V* q, int64_t n, CGScalars* sc,
                         hipStream_t s, const float* dinv) {
  if (n <= 0) return hipSuccess;
  if (dinv)
    hipLaunchKernelGGL((k_cg_dir<V, true>), dim3(stream_grid(n)), dim3(256), 0, s, p, r, q,
                       dinv, n, sc);
  else
    hipLaunchKernelGGL((k_cg_dir<V, false>), dim3(stream_grid(n)), dim3(256), 0, s, p, r, q,
                       dinv, n, sc);
  return hipGetLastError();
}

hipError_t launch_cg_alpha_parts(const double* part, int64_t np, Reducer red,
                                 CGScalars* sc, hipStream_t s) {
  hipLaunchKernelGGL(k_cg_alpha_parts, dim3(red_grid(np)), dim3(kRedThreads), 0, s,
                     part, np, red, sc);
  return hipGetLastError();
}

template <typename V, typename W>
hipError_t resid_two_phase(V* r, const W* q, const float* dinv, int64_t n, Reducer red,
                           CGScalars* sc, hipStream_t s) {
  const unsigned g = stream_grid(n);
  if (dinv)
    hipLaunchKernelGGL((k_cg_resid<V, W, true>), dim3(g), dim3(256), 0, s, r, q, dinv, n,
                       red.part, sc);
  else
    hipLaunchKernelGGL((k_cg_resid<V, W, false>), dim3(g), dim3(256), 0, s, r, q, dinv, n,
                       red.part, sc);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_cg_resid_finish, dim3(1), dim3(kFinThreads), 0, s, red.part,
                     (int)g, dinv ? 1 : 0, sc);
  return hipGetLastError();
}

template <typename V>
hipError_t launch_cg_resid(V* r, const V* q, int64_t n, Reducer red, CGScalars* sc,
                           hipStream_t s, const float* dinv) {
  return resid_two_phase<V, V>(r, q, dinv, n, red, sc, s);
}

template <typename V>
hipError_t launch_cg_xdir(V* x, V* p, const V* r, V* q, int64_t n, CGScalars* sc,
                          hipStream_t s, const float* dinv) {
  if (n <= 0) return hipSuccess;
  if (dinv)
    hipLaunchKernelGGL((k_cg_xdir<V, true>), dim3(stream_grid(n)), dim3(256), 0, s, x, p, r,
                       q, dinv, n, sc);
  else
    hipLaunchKernelGGL((k_cg_xdir<V, false>), dim3(stream_grid(n)), dim3(256), 0, s, x, p, r,
                       q, dinv, n, sc);
  return hipGetLastError();
}

hipError_t launch_cg_resid_mixed(double* r, const float* q, int64_t n, Reducer red,
                                 CGScalars* sc, hipStream_t s, const float* dinv) {
  return resid_two_phase<double, float>(r, q, dinv, n, red, sc, s);
}

hipError_t launch_cg_xdir_mixed(double* x, double* p, float* p32, const double* r,
                                float* q32, int64_t n, CGScalars* sc, hipStream_t s,
                                const float* dinv) {
  if (n <= 0) return hipSuccess;
  if (dinv)
    hipLaunchKernelGGL(k_cg_xdir_mixed<true>, dim3(stream_grid(n)), dim3(256), 0, s, x, p,
                       p32, r, q32, dinv, n, sc);
  else
    hipLaunchKernelGGL(k_cg_xdir_mixed<false>, dim3(stream_grid(n)), dim3(256), 0, s, x, p,
                       p32, r, q32, dinv, n, sc);
  return hipGetLastError();
}

hipError_t launch_cg_alpha_mixed(const double* p, const float* q, int64_t n,
                                 Reducer red, CGScalars* sc, hipStream_t s) {
  return alpha_two_phase<double, float>(p, q, n, red, sc, s);
}

hipError_t launch_cg_update_mixed(double* x, double* r, const double* p, const float* q,
                                  int64_t n, Reducer red, CGScalars* sc, hipStream_t s,
                                  const float* dinv) {
  return update_two_phase<double, float>(x, r, p, q, dinv, n, red, sc, s);
}

hipError_t launch_cg_dir_mixed(double* p, float* p32, const double* r, float* q32,
                               int64_t n, CGScalars* sc, hipStream_t s, const float* dinv) {
  if (n <= 0) return hipSuccess;
  if (dinv)
    hipLaunchKernelGGL(k_cg_dir_mixed<true>, dim3(stream_grid(n)), dim3(256), 0, s, p, p32,
                       r, q32, dinv, n, sc);
  else
    hipLaunchKernelGGL(k_cg_dir_mixed<false>, dim3(stream_grid(n)), dim3(256), 0, s, p, p32,
                       r, q32, dinv, n, sc);
  return hipGetLastError();
}

hipError_t launch_cg_xdir_p32(double* x, float* p32, const double* r, float* q32,
                              int64_t n, const CGScalars* sc, hipStream_t s,
                              const float* dinv) {
  if (n <= 0) return hipSuccess;
  if (dinv)
    hipLaunchKernelGGL(k_cg_xdir_p32<true>, dim3(stream_grid(n)), dim3(256), 0, s, x, p32, r,
                       q32, dinv, n, sc);
  else
    hipLaunchKernelGGL(k_cg_xdir_p32<false>, dim3(stream_grid(n)), dim3(256), 0, s, x, p32, r,
                       q32, dinv, n, sc);
  return hipGetLastError();
}

hipError_t launch_dir_coef_p32(float* p32, const double* r, float* q32, double c,
                               int64_t n, hipStream_t s, const float* dinv) {
  if (n <= 0) return hipSuccess;
  if (dinv)
    hipLaunchKernelGGL(k_dir_coef_p32<true>, dim3(stream_grid(n)), dim3(256), 0, s, p32, r,
                       q32, dinv, c, n);
  else
    hipLaunchKernelGGL(k_dir_coef_p32<false>, dim3(stream_grid(n)), dim3(256), 0, s, p32, r,
                       q32, dinv, c, n);
  return hipGetLastError();
}

hipError_t launch_dir_coef_mixed(double* p, float* p32, const double* r, float* q32,
                                 double c, int64_t n, hipStream_t s, const float* dinv) {
  if (n <= 0) return hipSuccess;
  if (dinv)
    hipLaunchKernelGGL(k_dir_coef_mixed<true>, dim3(stream_grid(n)), dim3(256), 0, s, p, p32,
                       r, q32, dinv, c, n);
  else
    hipLaunchKernelGGL(k_dir_coef_mixed<false>, dim3(stream_grid(n)), dim3(256), 0, s, p, p32,
                       r, q32, dinv, c, n);
  return hipGetLastError();
}

hipError_t launch_dir_coef(float* p, const float* r, float* q, float c,
                           int64_t n, hipStream_t s, const float* dinv) {
  if (n <= 0) return hipSuccess;
  if (dinv)
    hipLaunchKernelGGL(k_dir_coef<true>, dim3(stream_grid(n)), dim3(256), 0, s, p, r, q, dinv,
                       c, n);
  else
    hipLaunchKernelGGL(k_dir_coef<false>, dim3(stream_grid(n)), dim3(256), 0, s, p, r, q, dinv,
                       c, n);
  return hipGetLastError();
}

hipError_t launch_residual(const double* b, const double* Ax, double* r,
                           int64_t n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_residual, dim3(stream_grid(n)), dim3(256), 0, s, b, Ax,
                     r, n);
  return hipGetLastError();
}

hipError_t launch_scale_to_f32(const double* r, double inv_scale, float* r32,
                               int64_t n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_scale_to_f32, dim3(stream_grid(n)), dim3(256), 0, s, r,
                     inv_scale, r32, n);
  return hipGetLastError();
}

hipError_t launch_axpy_f32_to_f64(double* x, const float* d32, double scale,
                                  int64_t n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_axpy_f32_to_f64, dim3(stream_grid(n)), dim3(256), 0, s,
                     x, d32, scale, n);
  return hipGetLastError();
}

#define SF_INST(V)                                                            \
  template hipError_t launch_fill<V>(V*, int64_t, V, hipStream_t);            \
  template hipError_t launch_zero_list<V>(V*, const int32_t*, int64_t,        \
                                          hipStream_t);                       \
  template hipError_t launch_dot<V>(const V*, const V*, int64_t, Reducer,     \
                                    double*, hipStream_t);                    \
  template hipError_t launch_wdot<V>(const V*, const float*, int64_t, Reducer, \
                                     double*, hipStream_t);                   \
  template hipError_t launch_cg_start<V>(const V*, const V*, V*, V*, V*,      \
                                         int64_t, Reducer, CGScalars*, double, \
                                         hipStream_t, const float*);          \
  template hipError_t launch_cg_alpha<V>(const V*, const V*, int64_t,         \
                                         Reducer, CGScalars*, hipStream_t);   \
  template hipError_t launch_cg_update<V>(V*, V*, const V*, const V*,         \
                                          int64_t, Reducer, CGScalars*,       \
                                          hipStream_t, const float*);         \
  template hipError_t launch_cg_dir<V>(V*, const V*, V*, int64_t,             \
                                       CGScalars*, hipStream_t, const float*); \
  template hipError_t launch_cg_resid<V>(V*, const V*, int64_t, Reducer,      \
                                         CGScalars*, hipStream_t, const float*); \
  template hipError_t launch_cg_xdir<V>(V*, V*, const V*, V*, int64_t,        \
                                        CGScalars*, hipStream_t, const float*);
SF_INST(double)
SF_INST(float)
#undef SF_INST
template hipError_t launch_fill<unsigned>(unsigned*, int64_t, unsigned, hipStream_t);

}  // namespace sf
