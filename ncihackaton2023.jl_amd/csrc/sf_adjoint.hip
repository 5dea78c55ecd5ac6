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
                const int64_t* __restrict__ bcell,
                const int32_t* __restrict__ corig) {
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
        dJdk[(corig ? (int64_t)corig[ge] : ge) * S::QQQ + l * S::QQ + r] =
            glx * hx + gly * hy + glz * hz;
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

// ---------------------------------------------------------------------------
// Block-assembled 3D adjoint with the geometry recomputed from the cell
// vertices (the fused-geometry apply's design, sf_apply.hpp: same cell
// blocks, distinct-dof gather, LDS assembly and exclusive-store / atomic
// write-back).  At every quadrature point, with C the cofactor matrix of J
// and s = w / |det J|:
//   G grad = s C^T (C grad),  so  dJdk = grad(lam) . G grad(u)
//                                      = s (C grad(lam)) . (C grad(u))
//   and the lam-path flux is kappa s C^T (C grad(lam)).
// u and lam run through the x and y passes side by side (twice the apply's
// LDS per cell, so half the workgroups per CU); the z-column pass forms both
// gradients, writes dJdk (non-temporal, one value per quadrature point) and
// contracts only the lam flux back.  WANT_Y = false stops after the z pass.
template <int N, int Q>
struct AdjB3Shape {
  using AS = Apply3Shape<N, Q, double>;
  static constexpr int EB = AS::EB, NT = AS::NT, KU = AS::KU, TK = AS::TK;
  static constexpr int NN = N * N, NQ = N * Q, QQ = Q * Q, QQQ = Q * Q * Q;
  static constexpr int NNN = N * N * N;
  static constexpr int QP = (Q % 2 == 0) ? Q + 1 : Q;
  // x-pass outputs [t][z][y][i], t = (B u, D u, B lam, D lam)
  static constexpr int YA = QP, ZA = N * QP, TA = N * ZA, CA = 4 * TA;
  // y-pass outputs [t][z][k][i], t = (u: x, y, z paths; lam: x, y, z paths)
  static constexpr int ZB = QQ, TB = N * QQ, CB = 6 * TB;
  static constexpr int RA = EB * CA;
  static constexpr int RX = EB * NNN;  // distinct dofs of a block, at most
  // one region: x-pass outputs with the u and lam values behind them; then
  // the y-pass outputs over both; then the y^T-pass output (two tensors per
  // cell slot) with the assembled values at the end
  static constexpr int RC = cmax(EB * CB, RA + 2 * RX);
  static_assert((RC + EB * kGeoCoef3) * 8 <= 163840, "adjoint3d: LDS above 160 KiB");
  static_assert(NT >= EB * cmax(NN, cmax(NQ, QQ)), "one task per thread per stage");
};

template <int N, int Q, bool KAP, bool WANT_Y>
__global__ void __launch_bounds__((AdjB3Shape<N, Q>::NT))
__attribute__((amdgpu_waves_per_eu(2)))
    k_adjoint3d_blk(const double* __restrict__ u, const double* __restrict__ lam,
                    double* __restrict__ dJdk, double* __restrict__ ybar,
                    const uint16_t* __restrict__ pidx,
                    const int32_t* __restrict__ bdofs,
                    const int64_t* __restrict__ boff,
                    const int32_t* __restrict__ bexcl,
                    const int64_t* __restrict__ bcell, int nblk,
                    const double* __restrict__ tB, const double* __restrict__ tD,
                    const double* __restrict__ Xv, const double* __restrict__ tP,
                    const double* __restrict__ tW, const double* __restrict__ kap,
                    const int32_t* __restrict__ corig) {
  using S = AdjB3Shape<N, Q>;
  __shared__ double smem[S::RC + S::EB * kGeoCoef3];
  double* const regV = smem + S::RC;          // the block's map coefficients
  double* const regA = smem;                  // x / y^T pass outputs
  double* const regB = smem;                  // y pass outputs, z^T in place
  double* const regXu = smem + S::RA;         // the block's u values
  double* const regXl = smem + S::RA + S::RX; // ... and lam values
  double* const regY = smem + (S::RC - S::RX);  // assembled A(kappa) lam
  const int b = xcd_remap(blockIdx.x, nblk);
  const int64_t base = bcell[b];
  const int ncb = (int)(bcell[b + 1] - base);
  const int tid = threadIdx.x;
  const int64_t d0 = boff[b];
  const int nu = (int)(boff[b + 1] - d0);
  const int nexcl = bexcl[b];
  const int32_t* __restrict__ dofs = bdofs + d0;

  // ---- prologue: dof ids, u and lam values, local indices, vertices, kappa
  int32_t dv[S::KU];
  double uv[S::KU], lv[S::KU];
#pragma unroll
  for (int k = 0; k < S::KU; ++k) {
    const int i = tid + k * S::NT;
    dv[k] = i < nu ? __builtin_nontemporal_load(dofs + i) : -1;
  }
#pragma unroll
  for (int k = 0; k < S::KU; ++k) {
    const int j = dv[k] >= 0 ? dv[k] : 0;
    uv[k] = u[j];
    lv[k] = lam[j];
  }
  uint16_t li[N];
  const bool own1 = tid < S::EB * S::NN;
  const int e1 = tid / S::NN, r1 = tid - e1 * S::NN;  // r1 = z*N + y
  const int o1 = (r1 / N) * S::ZA + (r1 % N) * S::YA;
  const bool act1 = own1 && e1 < ncb;
  if (act1) {
    const uint16_t* lp = pidx + ((base + e1) * S::NN + r1) * N;
#pragma unroll
    for (int j = 0; j < N; ++j) li[j] = __builtin_nontemporal_load(lp + j);
  } else {
#pragma unroll
    for (int j = 0; j < N; ++j) li[j] = kNoDof;
  }
  const bool own3 = tid < S::EB * S::QQ;
  const int e3 = tid / S::QQ, r3 = tid - e3 * S::QQ;  // r3 = k*Q + i
  const bool act3 = own3 && e3 < ncb;
  // ncb * 21 map coefficients: more than NT when a block holds more than
  // NT / 21 cells (up to 64 at the lowest orders)
  for (int t = tid; t < ncb * kGeoCoef3; t += S::NT)
    regV[t] = __builtin_nontemporal_load(Xv + base * kGeoCoef3 + t);
  double kq[KAP ? Q : 1];
  if constexpr (KAP) {
    const double* kp = kap + (base + (act3 ? e3 : 0)) * S::QQQ + (act3 ? r3 : 0);
#pragma unroll
    for (int l = 0; l < Q; ++l)
      kq[l] = act3 ? __builtin_nontemporal_load(kp + l * S::QQ) : 0.0;
  }

  // ---- stage 0: the block's distinct u and lam values into LDS ----------
#pragma unroll
  for (int k = 0; k < S::KU; ++k) {
    const int i = tid + k * S::NT;
    if (i < nu) {
      regXu[i] = uv[k];
      regXl[i] = lv[k];
    }
  }
  __syncthreads();

  // ---- stage 1: x-pencils of u and lam ----------------------------------
  if (own1) {
    double uu[N], ll[N];
#pragma unroll
    for (int j = 0; j < N; ++j) {
      uu[j] = li[j] != kNoDof ? regXu[li[j]] : 0.0;
      ll[j] = li[j] != kNoDof ? regXl[li[j]] : 0.0;
    }
    double* sA = regA + e1 * S::CA + o1;
    TableRows<double, S::TK> t_(tB, tD);
#pragma unroll
    for (int i = 0; i < Q; ++i) {
      t_.at(i);
      double bu = 0, du = 0, bl = 0, dl = 0;
#pragma unroll
      for (int j = 0; j < N; ++j) {
        bu = fma(t_.b[i * N + j], uu[j], bu);
        du = fma(t_.d[i * N + j], uu[j], du);
        bl = fma(t_.b[i * N + j], ll[j], bl);
        dl = fma(t_.d[i * N + j], ll[j], dl);
      }
      sA[i] = bu;
      sA[S::TA + i] = du;
      sA[2 * S::TA + i] = bl;
      sA[3 * S::TA + i] = dl;
    }
  }
  __syncthreads();

  // ---- stage 2: y-pencils of both (into registers, barrier, over A) -------
  {
    const bool on2 = tid < S::EB * S::NQ;
    const int e = on2 ? tid / S::NQ : 0, r = on2 ? tid - e * S::NQ : 0;
    const int z = r / Q, i = r - z * Q;
    const double* sA = regA + e * S::CA + z * S::ZA + i;
    double* sB = regB + e * S::CB + z * S::ZB + i;
    double au[N], du[N], al[N], dl[N];
#pragma unroll
    for (int j = 0; j < N; ++j) {
      au[j] = sA[j * S::YA];
      du[j] = sA[S::TA + j * S::YA];
      al[j] = sA[2 * S::TA + j * S::YA];
      dl[j] = sA[3 * S::TA + j * S::YA];
    }
    __syncthreads();
    TableRows<double, S::TK> t_(tB, tD);
    if (on2)
#pragma unroll
      for (int k = 0; k < Q; ++k) {
        t_.at(k);
        double ubb = 0, ubd = 0, udb = 0, lbb = 0, lbd = 0, ldb = 0;
#pragma unroll
        for (int j = 0; j < N; ++j) {
          ubb = fma(t_.b[k * N + j], au[j], ubb);
          ubd = fma(t_.d[k * N + j], au[j], ubd);
          udb = fma(t_.b[k * N + j], du[j], udb);
          lbb = fma(t_.b[k * N + j], al[j], lbb);
          lbd = fma(t_.d[k * N + j], al[j], lbd);
          ldb = fma(t_.b[k * N + j], dl[j], ldb);
        }
        sB[k * Q] = udb;                 // u: x-gradient path
        sB[S::TB + k * Q] = ubd;         // u: y-gradient path
        sB[2 * S::TB + k * Q] = ubb;     // u: z-gradient path
        sB[3 * S::TB + k * Q] = ldb;     // lam: x
        sB[4 * S::TB + k * Q] = lbd;     // lam: y
        sB[5 * S::TB + k * Q] = lbb;     // lam: z
      }
  }
  __syncthreads();

  // ---- stage 3: z-columns: both gradients, dJ/dkappa, the lam flux --------
  // grad u at the column's Q points first (its z pencils then die), then the
  // lam pencils; the lam flux overwrites grad u point by point, so the
  // column holds 3Q + 3N values at a time, as in the apply
  if (own3) {
    double* sB = regB + e3 * S::CB + r3;
    double g0[Q], g1[Q], g2[Q];
    {
      double cx[N], cy[N], cz[N];
#pragma unroll
      for (int z = 0; z < N; ++z) {
        cx[z] = sB[z * S::ZB];
        cy[z] = sB[S::TB + z * S::ZB];
        cz[z] = sB[2 * S::TB + z * S::ZB];
      }
      TableRows<double, S::TK> t_(tB, tD);
#pragma unroll
      for (int l = 0; l < Q; ++l) {
        t_.at(l);
        double a0 = 0, a1 = 0, a2 = 0;
#pragma unroll
        for (int z = 0; z < N; ++z) {
          a0 = fma(t_.b[l * N + z], cx[z], a0);
          a1 = fma(t_.b[l * N + z], cy[z], a1);
          a2 = fma(t_.d[l * N + z], cz[z], a2);
        }
        g0[l] = a0;
        g1[l] = a1;
        g2[l] = a2;
      }
    }
    // pin grad u here: otherwise its FMAs are sunk into the point loop
    // below, which then keeps the u pencils live next to the lam pencils
    // and the flux (about twice the registers)
#pragma unroll
    for (int l = 0; l < Q; ++l) asm volatile("" : "+v"(g0[l]), "+v"(g1[l]), "+v"(g2[l]));
    double lx[N], ly[N], lz[N];
#pragma unroll
    for (int z = 0; z < N; ++z) {
      lx[z] = sB[3 * S::TB + z * S::ZB];
      ly[z] = sB[4 * S::TB + z * S::ZB];
      lz[z] = sB[5 * S::TB + z * S::ZB];
    }
    // edge vectors of the trilinear map along quadrature column (i, k), as
    // in the fused-geometry apply
    double E0[3], dE[3], F0[3], dF[3], H[3];
    const int qi = r3 % Q, qk = r3 / Q;
    const double wik = tW[qi] * tW[qk];
    column_geom3<double>(regV + e3 * kGeoCoef3, tP[qi], tP[qk], E0, dE, F0, dF, H);
    const int64_t crow = base + (act3 ? e3 : 0);  // the plan's cell, the caller's row
    double* kout = dJdk + (corig ? (int64_t)corig[crow] : crow) * S::QQQ + r3;
    TableRows<double, S::TK> t_(tB, tD);
#pragma unroll
    for (int l = 0; l < Q; ++l) {
      t_.at(l);
      double gl0 = 0, gl1 = 0, gl2 = 0;
#pragma unroll
      for (int z = 0; z < N; ++z) {
        gl0 = fma(t_.b[l * N + z], lx[z], gl0);
        gl1 = fma(t_.b[l * N + z], ly[z], gl1);
        gl2 = fma(t_.d[l * N + z], lz[z], gl2);
      }
      const double ze = tP[l];
      double j0[3], j1[3];
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        j0[a] = fma(ze, dE[a], E0[a]);
        j1[a] = fma(ze, dF[a], F0[a]);
      }
      const double* j2 = H;
      double C[3][3];
      C[0][0] = j1[1] * j2[2] - j2[1] * j1[2];
      C[0][1] = j2[1] * j0[2] - j0[1] * j2[2];
      C[0][2] = j0[1] * j1[2] - j1[1] * j0[2];
      C[1][0] = j2[0] * j1[2] - j1[0] * j2[2];
      C[1][1] = j0[0] * j2[2] - j2[0] * j0[2];
      C[1][2] = j1[0] * j0[2] - j0[0] * j1[2];
      C[2][0] = j1[0] * j2[1] - j2[0] * j1[1];
      C[2][1] = j2[0] * j0[1] - j0[0] * j2[1];
      C[2][2] = j0[0] * j1[1] - j1[0] * j0[1];
      const double det = j0[0] * C[0][0] + j1[0] * C[0][1] + j2[0] * C[0][2];
      const double sc = wik * tW[l] * recip_pos(fabs(det));
      const double vu0 = C[0][0] * g0[l] + C[0][1] * g1[l] + C[0][2] * g2[l];
      const double vu1 = C[1][0] * g0[l] + C[1][1] * g1[l] + C[1][2] * g2[l];
      const double vu2 = C[2][0] * g0[l] + C[2][1] * g1[l] + C[2][2] * g2[l];
      const double vl0 = C[0][0] * gl0 + C[0][1] * gl1 + C[0][2] * gl2;
      const double vl1 = C[1][0] * gl0 + C[1][1] * gl1 + C[1][2] * gl2;
      const double vl2 = C[2][0] * gl0 + C[2][1] * gl1 + C[2][2] * gl2;
      if (act3)
        __builtin_nontemporal_store(sc * (vl0 * vu0 + vl1 * vu1 + vl2 * vu2),
                                    kout + l * S::QQ);
      if constexpr (WANT_Y) {
        const double sk = KAP ? sc * kq[l] : sc;
        g0[l] = sk * (C[0][0] * vl0 + C[1][0] * vl1 + C[2][0] * vl2);
        g1[l] = sk * (C[0][1] * vl0 + C[1][1] * vl1 + C[2][1] * vl2);
        g2[l] = sk * (C[0][2] * vl0 + C[1][2] * vl1 + C[2][2] * vl2);
      }
      // keep the scheduler from overlapping point iterations (their
      // geometry temporaries otherwise overflow the register file)
      __builtin_amdgcn_sched_barrier(0);
    }
    if constexpr (WANT_Y) {
      TableRows<double, S::TK> tt_(tB + Q * N, tD + Q * N);  // transposed tables
#pragma unroll
      for (int z = 0; z < N; ++z) {
        tt_.at(z);
        double rx = 0, ry = 0, rz = 0;
#pragma unroll
        for (int l = 0; l < Q; ++l) {
          rx = fma(tt_.b[z * Q + l], g0[l], rx);
          ry = fma(tt_.b[z * Q + l], g1[l], ry);
          rz = fma(tt_.d[z * Q + l], g2[l], rz);
        }
        sB[z * S::ZB] = rx;
        sB[S::TB + z * S::ZB] = ry;
        sB[2 * S::TB + z * S::ZB] = rz;
      }
    }
  }
  if constexpr (!WANT_Y) return;
  __syncthreads();

  // ---- stage 4: transposed y-pencils (into registers, barrier, over A) ----
  {
    const bool on4 = tid < S::EB * S::NQ;
    const int e = on4 ? tid / S::NQ : 0, r = on4 ? tid - e * S::NQ : 0;
    const int z = r / Q, i = r - z * Q;
    double* sA = regA + e * S::CA + z * S::ZA + i;
    const double* sB = regB + e * S::CB + z * S::ZB + i;
    double rx[Q], ry[Q], rz[Q];
#pragma unroll
    for (int k = 0; k < Q; ++k) {
      rx[k] = sB[k * Q];
      ry[k] = sB[S::TB + k * Q];
      rz[k] = sB[2 * S::TB + k * Q];
    }
    __syncthreads();
    TableRows<double, S::TK> t_(tB + Q * N, tD + Q * N);
    if (on4)
#pragma unroll
      for (int j = 0; j < N; ++j) {
        t_.at(j);
        double sx = 0, so = 0;
#pragma unroll
        for (int k = 0; k < Q; ++k) {
          sx = fma(t_.b[j * Q + k], rx[k], sx);
          so = fma(t_.d[j * Q + k], ry[k], so);
          so = fma(t_.b[j * Q + k], rz[k], so);
        }
        sA[j * S::YA] = sx;           // then D_x^T
        sA[S::TA + j * S::YA] = so;   // then B_x^T
      }
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < S::KU; ++k) {
    const int i = tid + k * S::NT;
    if (i < nu) regY[i] = 0.0;
  }
  __syncthreads();

  // ---- stage 5: transposed x-pencils, assembled in LDS --------------------
  if (act1) {
    const double* sA = regA + e1 * S::CA + o1;
    double sx[Q], so[Q];
#pragma unroll
    for (int i = 0; i < Q; ++i) {
      sx[i] = sA[i];
      so[i] = sA[S::TA + i];
    }
    TableRows<double, S::TK> t_(tB + Q * N, tD + Q * N);
#pragma unroll
    for (int m = 0; m < N; ++m) {
      t_.at(m);
      double v = 0;
#pragma unroll
      for (int i = 0; i < Q; ++i) {
        v = fma(t_.d[m * Q + i], sx[i], v);
        v = fma(t_.b[m * Q + i], so[i], v);
      }
      if (li[m] != kNoDof) lds_atomic_add(regY + li[m], v);
    }
  }
  __syncthreads();

  // ---- write-back: exclusive dofs stored, shared dofs added ---------------
#pragma unroll
  for (int k = 0; k < S::KU; ++k) {
    const int i = tid + k * S::NT;
    if (dv[k] < 0) continue;
    const double v = regY[i];
    if (i < nexcl)
      ybar[dv[k]] = v;
    else
      atomic_add(ybar + dv[k], v);
  }
}

template <int N, int Q>
hipError_t launch_adj3_blk(const double* u, const double* lam, double* dJdk,
                           double* ybar, const BlockArgs& k, const double* tB,
                           const double* tD, const double* Xv, const double* tP,
                           const double* tW, const double* kap, const int32_t* corig,
                           hipStream_t s) {
  using S = AdjB3Shape<N, Q>;
  if (k.nblk <= 0) return hipSuccess;
#define SF_ADJ_LAUNCH(KAP_, WY_)                                                     \
  hipLaunchKernelGGL((k_adjoint3d_blk<N, Q, KAP_, WY_>), dim3(k.nblk), dim3(S::NT), 0, \
                     s, u, lam, dJdk, ybar, k.pidx, k.bdofs, k.boff, k.bexcl,        \
                     k.bcell, k.nblk, tB, tD, Xv, tP, tW, kap, corig)
  if (kap && ybar)
    SF_ADJ_LAUNCH(true, true);
  else if (ybar)
    SF_ADJ_LAUNCH(false, true);
  else
    SF_ADJ_LAUNCH(false, false);  // kappa scales only the lam flux
#undef SF_ADJ_LAUNCH
  return hipGetLastError();
}

hipError_t adjoint3d_fused_f64(int n, int q, const double* u, const double* lam,
                               double* dJdk, double* ybar, const BlockArgs* k,
                               const double* tB, const double* tD, const double* Xv,
                               const double* tP, const double* tW, const double* kappa,
                               const int32_t* corig, hipStream_t s) {
  if (!k || !dJdk || !Xv) return hipErrorInvalidValue;
  switch (n * 16 + q) {
#define SF_CASE(N_, Q_)                                                            \
  case N_ * 16 + Q_:                                                               \
    return launch_adj3_blk<N_, Q_>(u, lam, dJdk, ybar, *k, tB, tD, Xv, tP, tW, kappa, corig, \
                                   s);
    SF_FOR_EACH_NQ(SF_CASE)
#undef SF_CASE
    default:
      return hipErrorNotSupported;
  }
}

template <int N, int Q>
hipError_t launch_adj3(const double* u, const double* lam, double* dJdk,
                       double* ybar, const int32_t* cells, const double* G,
                       const double* kappa, const double* tB,
                       const double* tD, int64_t ncell,
                       const int32_t* cell_block, const int64_t* bcell,
                       const int32_t* corig, hipStream_t s) {
  using S = Adj3Shape<N, Q>;
  const int64_t nb = (ncell + S::EB - 1) / S::EB;
  if (nb <= 0) return hipSuccess;
  if (nb > 0x7fffffff) return hipErrorInvalidValue;
  hipLaunchKernelGGL((k_adjoint3d<N, Q>), dim3((int)nb), dim3(S::NT), 0, s, u,
                     lam, dJdk, ybar, cells, G, kappa, tB, tD, ncell, (int)nb,
                     cell_block, bcell, corig);
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
                       const int32_t* corig, hipStream_t s) {
  if (dim == 3) {
    switch (n * 16 + q) {
#define SF_CASE(N_, Q_)                                                   \
  case N_ * 16 + Q_:                                                      \
    return launch_adj3<N_, Q_>(u, lam, dJdk, ybar, cells, G, kappa, tB, tD, \
                               ncell, cell_block, bcell, corig, s);
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
