// sf_adjoint.hip -- pullback of y = A(kappa) u for a cotangent lam, fused:
//
//   dJdk[cell][qp] = grad_ref(lam)_qp . G_qp grad_ref(u)_qp   (G without kappa)
//   ybar          += A(kappa) lam  (= A(kappa)^T lam, A is symmetric)
//
// A(kappa) is linear in kappa, so d<lam, A(kappa) u>/d kappa_qp is exactly the
// first line.  One kernel gathers u and lam, forms both reference gradients
// at the quadrature points with the same sum-factorised passes as the apply
// (sf_apply.hpp), uses them at once and never stores them in HBM; the only
// HBM traffic besides the vectors is G, kappa and the dJdk stream.
#include "sf_apply.hpp"
#include "sf_kernels.h"

namespace sf {

template <int N, int Q>
struct Adj3Shape {
  static constexpr int NN = N * N, NQ = N * Q, QQ = Q * Q, QQQ = Q * Q * Q;
  static constexpr int QP = (Q % 2 == 0) ? Q + 1 : Q;
  static constexpr int SA1 = NN * QP;
  static constexpr int SB1 = N * QQ;
  static constexpr int SE = 4 * SA1 + 6 * SB1;  // u and lam tensors
  static constexpr int MAXT = cmax(NN, cmax(NQ, QQ));
  static constexpr int EB_T = MAXT > 100 ? 1 : cmax(1, 256 / MAXT);
  static constexpr int EB_L = cmax(1, 65536 / (SE * 8));
  static constexpr int EB = EB_T < EB_L ? EB_T : EB_L;
  static constexpr int NT = round_up(EB * MAXT, 64);
  // gfx950 allows up to 160 KiB of LDS per workgroup (high orders, EB = 1)
  static_assert(EB * SE * 8 <= 163840, "adjoint3d: LDS above 160 KiB");
};

template <int N, int Q>
__global__ void __launch_bounds__((Adj3Shape<N, Q>::NT))
    k_adjoint3d(const double* __restrict__ u, const double* __restrict__ lam,
                double* __restrict__ dJdk, double* __restrict__ ybar,
                const int32_t* __restrict__ cells,
                const double* __restrict__ G,
                const double* __restrict__ kappa,
                const double* __restrict__ tB, const double* __restrict__ tD,
                int64_t ncell, int nblk,
                const int32_t* __restrict__ cell_block,
                const int64_t* __restrict__ bcell) {
  using S = Adj3Shape<N, Q>;
  __shared__ double smem[S::EB * S::SE];
  const int blk = xcd_remap(blockIdx.x, nblk);
  const int64_t base = (int64_t)blk * S::EB;
  const int tid = threadIdx.x;
  // per cell: [A_u (2 SA1)][A_l (2 SA1)][B_u (3 SB1)][B_l (3 SB1)]

  // stage 1: gather u, lam; x-pencils.  NT >= EB * NN, so the thread of
  // x-pencil (e, r) also scatters it in stage 5 and keeps its ids.
  static_assert(S::NT >= S::EB * S::NN, "one x-pencil per thread");
  int ids[N];
  const bool own1 = tid < S::EB * S::NN;
  const int e1 = tid / S::NN, r1 = tid - e1 * S::NN;
  const int64_t ge1 = base + e1;
  const bool act1 = own1 && ge1 < ncell;
  if (own1) {
    double uu[N], ll[N];
    if (act1) {
      const int32_t* cid = cells + (ge1 * S::NN + r1) * N;
#pragma unroll
      for (int j = 0; j < N; ++j) ids[j] = cid[j];
#pragma unroll
      for (int j = 0; j < N; ++j) {
        const int k = ids[j] >= 0 ? ids[j] : 0;
        uu[j] = u[k];
        ll[j] = lam[k];
      }
#pragma unroll
      for (int j = 0; j < N; ++j) {
        uu[j] = ids[j] >= 0 ? uu[j] : 0.0;
        ll[j] = ids[j] >= 0 ? ll[j] : 0.0;
      }
    } else {
#pragma unroll
      for (int j = 0; j < N; ++j) uu[j] = ll[j] = 0.0;
    }
    double* sA = smem + e1 * S::SE;
#pragma unroll
    for (int i = 0; i < Q; ++i) {
      const auto b_ = table_row(tB);
      const auto d_ = table_row(tD);
      double bu = 0, du = 0, bl = 0, dl = 0;
#pragma unroll
      for (int j = 0; j < N; ++j) {
        bu = fma(b_[i * N + j], uu[j], bu);
        du = fma(d_[i * N + j], uu[j], du);
        bl = fma(b_[i * N + j], ll[j], bl);
        dl = fma(d_[i * N + j], ll[j], dl);
      }
      sA[r1 * S::QP + i] = bu;
      sA[S::SA1 + r1 * S::QP + i] = du;
      sA[2 * S::SA1 + r1 * S::QP + i] = bl;
      sA[3 * S::SA1 + r1 * S::QP + i] = dl;
    }
  }
  __syncthreads();

  // stage 2: y-pencils for u and lam
  for (int t = tid; t < S::EB * 2 * S::NQ; t += S::NT) {
    const int e = t / (2 * S::NQ);
    int r = t - e * 2 * S::NQ;
    const int which = r / S::NQ;  // 0: u, 1: lam
    r -= which * S::NQ;
    const int z = r / Q, i = r - z * Q;
    const double* sA = smem + e * S::SE + which * 2 * S::SA1;
    double* sB = smem + e * S::SE + 4 * S::SA1 + which * 3 * S::SB1;
    double a[N], d[N];
#pragma unroll
    for (int j = 0; j < N; ++j) {
      a[j] = sA[(z * N + j) * S::QP + i];
      d[j] = sA[S::SA1 + (z * N + j) * S::QP + i];
    }
#pragma unroll
    for (int k = 0; k < Q; ++k) {
      const auto b_ = table_row(tB);
      const auto d_ = table_row(tD);
      double bb = 0, bd = 0, db = 0;
#pragma unroll
      for (int j = 0; j < N; ++j) {
        bb = fma(b_[k * N + j], a[j], bb);
        bd = fma(d_[k * N + j], a[j], bd);
        db = fma(b_[k * N + j], d[j], db);
      }
      const int o = z * S::QQ + k * Q + i;
      sB[o] = db;
      sB[S::SB1 + o] = bd;
      sB[2 * S::SB1 + o] = bb;
    }
  }
  __syncthreads();

  // stage 3: z-columns: both gradients, dJ/dkappa, kappa G grad(lam), z^T
  const bool want_y = ybar != nullptr;
  for (int t = tid; t < S::EB * S::QQ; t += S::NT) {
    const int e = t / S::QQ, r = t - e * S::QQ;
    const int64_t ge = base + e;
    const bool valid = ge < ncell;
    double* sBu = smem + e * S::SE + 4 * S::SA1;
    double* sBl = sBu + 3 * S::SB1;
    double cux[N], cuy[N], cuz[N], clx[N], cly[N], clz[N];
#pragma unroll
    for (int z = 0; z < N; ++z) {
      cux[z] = sBu[z * S::QQ + r];
      cuy[z] = sBu[S::SB1 + z * S::QQ + r];
      cuz[z] = sBu[2 * S::SB1 + z * S::QQ + r];
      clx[z] = sBl[z * S::QQ + r];
      cly[z] = sBl[S::SB1 + z * S::QQ + r];
      clz[z] = sBl[2 * S::SB1 + z * S::QQ + r];
    }
    double tx[Q], ty[Q], tz[Q];
    // G in the apply's block-interleaved layout (g3_index)
    int64_t bs = 0;
    int ncb = 1;
    if (valid) {
      const int cb = cell_block[ge];
      bs = bcell[cb];
      ncb = (int)(bcell[cb + 1] - bs);
    }
    const double* g = G + g3_index(bs, ncb, valid ? ge : bs, 0, 0, r, Q);
    const int64_t kst = (int64_t)ncb * S::QQ;
#pragma unroll
    for (int l = 0; l < Q; ++l) {
      const auto b_ = table_row(tB);
      const auto d_ = table_row(tD);
      double gux = 0, guy = 0, guz = 0, glx = 0, gly = 0, glz = 0;
#pragma unroll
      for (int z = 0; z < N; ++z) {
        gux = fma(b_[l * N + z], cux[z], gux);
        guy = fma(b_[l * N + z], cuy[z], guy);
        guz = fma(d_[l * N + z], cuz[z], guz);
        glx = fma(b_[l * N + z], clx[z], glx);
        gly = fma(b_[l * N + z], cly[z], gly);
        glz = fma(d_[l * N + z], clz[z], glz);
      }
      double g0 = 0, g1 = 0, g2 = 0, g3 = 0, g4 = 0, g5 = 0, kq = 1.0;
      if (valid) {
        g0 = load_g<double, double>(g + (0 * Q + l) * kst);
        g1 = load_g<double, double>(g + (1 * Q + l) * kst);
        g2 = load_g<double, double>(g + (2 * Q + l) * kst);
        g3 = load_g<double, double>(g + (3 * Q + l) * kst);
        g4 = load_g<double, double>(g + (4 * Q + l) * kst);
        g5 = load_g<double, double>(g + (5 * Q + l) * kst);
        if (kappa) kq = kappa[ge * S::QQQ + l * S::QQ + r];
      }
      if (dJdk && valid) {
        const double hx = g0 * gux + g1 * guy + g2 * guz;
        const double hy = g1 * gux + g3 * guy + g4 * guz;
        const double hz = g2 * gux + g4 * guy + g5 * guz;
        dJdk[ge * S::QQQ + l * S::QQ + r] = glx * hx + gly * hy + glz * hz;
      }
      tx[l] = kq * (g0 * glx + g1 * gly + g2 * glz);
      ty[l] = kq * (g1 * glx + g3 * gly + g4 * glz);
      tz[l] = kq * (g2 * glx + g4 * gly + g5 * glz);
    }
    if (want_y) {
#pragma unroll
      for (int z = 0; z < N; ++z) {
        const auto b_ = table_row(tB);
        const auto d_ = table_row(tD);
        double rx = 0, ry = 0, rz = 0;
#pragma unroll
        for (int l = 0; l < Q; ++l) {
          rx = fma(b_[l * N + z], tx[l], rx);
          ry = fma(b_[l * N + z], ty[l], ry);
          rz = fma(d_[l * N + z], tz[l], rz);
        }
        sBl[z * S::QQ + r] = rx;
        sBl[S::SB1 + z * S::QQ + r] = ry;
        sBl[2 * S::SB1 + z * S::QQ + r] = rz;
      }
    }
  }
  if (!want_y) return;  // uniform over the grid
  __syncthreads();

  // stage 4: transposed y-pencils (lam path)
  for (int t = tid; t < S::EB * S::NQ; t += S::NT) {
    const int e = t / S::NQ, r = t - e * S::NQ;
    const int z = r / Q, i = r - z * Q;
    double* sA = smem + e * S::SE;  // reuse the u x-pass tensors
    const double* sB = smem + e * S::SE + 4 * S::SA1 + 3 * S::SB1;
    double rx[Q], ry[Q], rz[Q];
#pragma unroll
    for (int k = 0; k < Q; ++k) {
      const int o = z * S::QQ + k * Q + i;
      rx[k] = sB[o];
      ry[k] = sB[S::SB1 + o];
      rz[k] = sB[2 * S::SB1 + o];
    }
#pragma unroll
    for (int j = 0; j < N; ++j) {
      const auto b_ = table_row(tB);
      const auto d_ = table_row(tD);
      double sx = 0, so = 0;
#pragma unroll
      for (int k = 0; k < Q; ++k) {
        sx = fma(b_[k * N + j], rx[k], sx);
        so = fma(d_[k * N + j], ry[k], so);
        so = fma(b_[k * N + j], rz[k], so);
      }
      sA[(z * N + j) * S::QP + i] = sx;
      sA[S::SA1 + (z * N + j) * S::QP + i] = so;
    }
  }
  __syncthreads();

  // stage 5: transposed x-pencils + scatter-add
  if (act1) {
    const double* sA = smem + e1 * S::SE;
    double sx[Q], so[Q];
#pragma unroll
    for (int i = 0; i < Q; ++i) {
      sx[i] = sA[r1 * S::QP + i];
      so[i] = sA[S::SA1 + r1 * S::QP + i];
    }
#pragma unroll
    for (int m = 0; m < N; ++m) {
      const auto b_ = table_row(tB);
      const auto d_ = table_row(tD);
      double v = 0;
#pragma unroll
      for (int i = 0; i < Q; ++i) {
        v = fma(d_[i * N + m], sx[i], v);
        v = fma(b_[i * N + m], so[i], v);
      }
      if (ids[m] >= 0) atomic_add(ybar + ids[m], v);
    }
  }
}

// 2D version
template <int N, int Q>
struct Adj2Shape {
  static constexpr int QQ = Q * Q;
  static constexpr int QP = (Q % 2 == 0) ? Q + 1 : Q;
  static constexpr int SA1 = N * QP;
  static constexpr int SE = 4 * SA1;
  static constexpr int MAXT = cmax(N, Q);
  static constexpr int EB = cmax(1, 256 / MAXT);
  static constexpr int NT = round_up(EB * MAXT, 64);
};

template <int N, int Q>
__global__ void __launch_bounds__((Adj2Shape<N, Q>::NT))
    k_adjoint2d(const double* __restrict__ u, const double* __restrict__ lam,
                double* __restrict__ dJdk, double* __restrict__ ybar,
                const int32_t* __restrict__ cells,
                const double* __restrict__ G,
                const double* __restrict__ kappa,
                const double* __restrict__ tB, const double* __restrict__ tD,
                int64_t ncell, int nblk) {
  using S = Adj2Shape<N, Q>;
  __shared__ double smem[S::EB * S::SE];
  const int blk = xcd_remap(blockIdx.x, nblk);
  const int64_t base = (int64_t)blk * S::EB;
  const int tid = threadIdx.x;

  static_assert(S::NT >= S::EB * N, "one x-pencil per thread");
  int ids[N];
  const bool own1 = tid < S::EB * N;
  const int e1 = tid / N, r1 = tid - e1 * N;
  const int64_t ge1 = base + e1;
  const bool act1 = own1 && ge1 < ncell;
  if (own1) {
    double uu[N], ll[N];
    if (act1) {
      const int32_t* cid = cells + (ge1 * N + r1) * N;
#pragma unroll
      for (int j = 0; j < N; ++j) ids[j] = cid[j];
#pragma unroll
      for (int j = 0; j < N; ++j) {
        const int k = ids[j] >= 0 ? ids[j] : 0;
        uu[j] = u[k];
        ll[j] = lam[k];
      }
#pragma unroll
      for (int j = 0; j < N; ++j) {
        uu[j] = ids[j] >= 0 ? uu[j] : 0.0;
        ll[j] = ids[j] >= 0 ? ll[j] : 0.0;
      }
    } else {
#pragma unroll
      for (int j = 0; j < N; ++j) uu[j] = ll[j] = 0.0;
    }
    double* sA = smem + e1 * S::SE;
#pragma unroll
    for (int i = 0; i < Q; ++i) {
      const auto b_ = table_row(tB);
      const auto d_ = table_row(tD);
      double bu = 0, du = 0, bl = 0, dl = 0;
#pragma unroll
      for (int j = 0; j < N; ++j) {
        bu = fma(b_[i * N + j], uu[j], bu);
        du = fma(d_[i * N + j], uu[j], du);
        bl = fma(b_[i * N + j], ll[j], bl);
        dl = fma(d_[i * N + j], ll[j], dl);
      }
      sA[r1 * S::QP + i] = bu;
      sA[S::SA1 + r1 * S::QP + i] = du;
      sA[2 * S::SA1 + r1 * S::QP + i] = bl;
      sA[3 * S::SA1 + r1 * S::QP + i] = dl;
    }
  }
  __syncthreads();

  const bool want_y = ybar != nullptr;
  for (int t = tid; t < S::EB * Q; t += S::NT) {
    const int e = t / Q, i = t - e * Q;
    const int64_t ge = base + e;
    const bool valid = ge < ncell;
    double* sA = smem + e * S::SE;
    double au[N], du[N], al[N], dl[N];
#pragma unroll
    for (int j = 0; j < N; ++j) {
      au[j] = sA[j * S::QP + i];
      du[j] = sA[S::SA1 + j * S::QP + i];
      al[j] = sA[2 * S::SA1 + j * S::QP + i];
      dl[j] = sA[3 * S::SA1 + j * S::QP + i];
    }
    const double* g = G + ge * 3 * S::QQ + i;
    double tx[Q], ty[Q];
#pragma unroll
    for (int k = 0; k < Q; ++k) {
      const auto b_ = table_row(tB);
      const auto d_ = table_row(tD);
      double gux = 0, guy = 0, glx = 0, gly = 0;
#pragma unroll
      for (int j = 0; j < N; ++j) {
        gux = fma(b_[k * N + j], du[j], gux);
        guy = fma(d_[k * N + j], au[j], guy);
        glx = fma(b_[k * N + j], dl[j], glx);
        gly = fma(d_[k * N + j], al[j], gly);
      }
      double g0 = 0, g1 = 0, g2 = 0, kq = 1.0;
      if (valid) {
        g0 = g[0 * S::QQ + k * Q];
        g1 = g[1 * S::QQ + k * Q];
        g2 = g[2 * S::QQ + k * Q];
        if (kappa) kq = kappa[ge * S::QQ + k * Q + i];
      }
      if (dJdk && valid)
        dJdk[ge * S::QQ + k * Q + i] =
            glx * (g0 * gux + g1 * guy) + gly * (g1 * gux + g2 * guy);
      tx[k] = kq * (g0 * glx + g1 * gly);
      ty[k] = kq * (g1 * glx + g2 * gly);
    }
    if (want_y) {
#pragma unroll
      for (int j = 0; j < N; ++j) {
        const auto b_ = table_row(tB);
        const auto d_ = table_row(tD);
        double rx = 0, ry = 0;
#pragma unroll
        for (int k = 0; k < Q; ++k) {
          rx = fma(b_[k * N + j], tx[k], rx);
          ry = fma(d_[k * N + j], ty[k], ry);
        }
        sA[j * S::QP + i] = rx;
        sA[S::SA1 + j * S::QP + i] = ry;
      }
    }
  }
  if (!want_y) return;
  __syncthreads();

  if (act1) {
    const double* sA = smem + e1 * S::SE;
    double sx[Q], so[Q];
#pragma unroll
    for (int i = 0; i < Q; ++i) {
      sx[i] = sA[r1 * S::QP + i];
      so[i] = sA[S::SA1 + r1 * S::QP + i];
    }
#pragma unroll
    for (int m = 0; m < N; ++m) {
      const auto b_ = table_row(tB);
      const auto d_ = table_row(tD);
      double v = 0;
#pragma unroll
      for (int i = 0; i < Q; ++i) {
        v = fma(d_[i * N + m], sx[i], v);
        v = fma(b_[i * N + m], so[i], v);
      }
      if (ids[m] >= 0) atomic_add(ybar + ids[m], v);
    }
  }
}

template <int N, int Q>
hipError_t launch_adj3(const double* u, const double* lam, double* dJdk,
                       double* ybar, const int32_t* cells, const double* G,
                       const double* kappa, const double* tB,
                       const double* tD, int64_t ncell,
                       const int32_t* cell_block, const int64_t* bcell,
                       hipStream_t s) {
  using S = Adj3Shape<N, Q>;
  const int64_t nb = (ncell + S::EB - 1) / S::EB;
  if (nb <= 0) return hipSuccess;
  if (nb > 0x7fffffff) return hipErrorInvalidValue;
  hipLaunchKernelGGL((k_adjoint3d<N, Q>), dim3((int)nb), dim3(S::NT), 0, s, u,
                     lam, dJdk, ybar, cells, G, kappa, tB, tD, ncell, (int)nb,
                     cell_block, bcell);
  return hipGetLastError();
}

template <int N, int Q>
hipError_t launch_adj2(const double* u, const double* lam, double* dJdk,
                       double* ybar, const int32_t* cells, const double* G,
                       const double* kappa, const double* tB,
                       const double* tD, int64_t ncell, hipStream_t s) {
  using S = Adj2Shape<N, Q>;
  const int64_t nb = (ncell + S::EB - 1) / S::EB;
  if (nb <= 0) return hipSuccess;
  if (nb > 0x7fffffff) return hipErrorInvalidValue;
  hipLaunchKernelGGL((k_adjoint2d<N, Q>), dim3((int)nb), dim3(S::NT), 0, s, u,
                     lam, dJdk, ybar, cells, G, kappa, tB, tD, ncell, (int)nb);
  return hipGetLastError();
}

hipError_t adjoint_f64(int dim, int n, int q, const double* u,
                       const double* lam, double* dJdk, double* ybar,
                       const int32_t* cells, const double* G,
                       const double* kappa, const double* tB,
                       const double* tD, int64_t ncell,
                       const int32_t* cell_block, const int64_t* bcell,
                       hipStream_t s) {
  if (dim == 3) {
    switch (n * 16 + q) {
#define SF_CASE(N_, Q_)                                                   \
  case N_ * 16 + Q_:                                                      \
    return launch_adj3<N_, Q_>(u, lam, dJdk, ybar, cells, G, kappa, tB, tD, \
                               ncell, cell_block, bcell, s);
      SF_FOR_EACH_NQ(SF_CASE)
#undef SF_CASE
      default:
        return hipErrorNotSupported;
    }
  } else if (dim == 2) {
    switch (n * 16 + q) {
#define SF_CASE(N_, Q_)                                                   \
  case N_ * 16 + Q_:                                                      \
    return launch_adj2<N_, Q_>(u, lam, dJdk, ybar, cells, G, kappa, tB, tD, \
                               ncell, s);
      SF_FOR_EACH_NQ(SF_CASE)
#undef SF_CASE
      default:
        return hipErrorNotSupported;
    }
  }
  return hipErrorNotSupported;
}

}  // namespace sf
