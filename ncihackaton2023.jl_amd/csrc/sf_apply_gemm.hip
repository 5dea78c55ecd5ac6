// sf_apply_gemm.hip -- 3D apply at n = 8 nodes per direction (Q7) with every
// 1D contraction a GEMM on the matrix cores over 32 pencils at a time,
// batched across the cells of the block: v_mfma_f32_32x32x8f16 (FP16
// operands, FP32 accumulation), K = 8 = the nodes of a pencil, so the
// node-index contractions need no K padding.
//
// Same operator, block structure, stored bf16 G (the layout of the other
// MFMA apply) and assembly as k_apply3d.  The cost of a matrix-core apply is
// not the MFMAs but moving the operands into their lane layouts; this
// kernel keeps every per-element step out of the vector ALU:
//   * every table operand is a per-lane constant (built once per launch);
//   * each stage reads its data operand with one 8-byte LDS load per lane
//     (four consecutive FP16 values of the contracted index) and writes its
//     output with one 8-byte store per group of four rows, because each
//     stage enumerates its pencils so that the four consecutive rows a lane
//     holds in the MFMA result are four consecutive values of the index the
//     next stage contracts;
//   * rows whose length is not a multiple of 4 are padded (QP = Q rounded
//     up to 4); the padding, and reads that run past a row into the next,
//     only meet table entries that are zero, and the LDS is zeroed at the
//     start so that everything such a read can meet is finite.
//
// MFMA 32x32x8 lane layout (wave64, lane l, h = l / 32, r = l % 32):
//   A[m][k]: m = r, k = 4h + j        B[k][n]: n = r, k = 4h + j   (j < 4)
//   C[m][n]: n = r, m = 8i + 4h + jj  in register 4i + jj         (i, jj < 4)
//
// Stages (data as the A operand and the tables as B, except Z / Z^T):
//   X   : u[e][c][b][a]  . [B | D]^T      -> Yin[t][e][c][i1][b]  (t: Bx, Dx)
//   Y   : Yin            . [B | D | B]^T  -> Zin[p][e][i2][i1][c] (p: DxBy, BxDy, BxBy)
//   Z   : B or D (A) . Zin (B)            -> grad at (i3, pencil) in registers,
//         h = G grad (vector ALU, FP32, only at the lane's valid points)
//   Z^T : B^T or D^T (A) . h (B: the Z output tile is that operand as it is)
//                                         -> W[f][e][c][i1][i2]   (f: x, y, z)
//   Y^T : W . B (x), W . D (y) + W . B (z) -> S[t][e][c][b][i1]   (t: D-path, B-path)
//   X^T : S . D + S . B                   -> y tensor (FP32), assembled in LDS
//
// FP16 range.  Every conversion to FP16 first scales by a power of two that
// brings the bound of the value below 2^14: the block's x by its max |x|, each
// contraction's output by the bound of its own table (row sums forward,
// column sums transposed: the D table's are about 100 times the B
// table's, so one scale for both would push the B paths into FP16's
// subnormals within two stages), the flux by max |G| and the gradient
// bounds.  Sums of terms at different scales (the gradient components in
// the flux, the two terms of the transposed y pass and of the x pass) are
// formed in FP32 after rescaling to a common one.  A value far below its
// bound keeps FP16's 11 significant bits down to 2^-28 of the bound (the
// flux, for instance, spans the contrast of kappa).
#include "sf_apply.hpp"
#include "sf_common.hpp"
#include "sf_kernels.h"

namespace sf {
namespace {

typedef _Float16 h4 __attribute__((ext_vector_type(4)));
typedef float f16v __attribute__((ext_vector_type(16)));
typedef float f4v __attribute__((ext_vector_type(4)));

constexpr int GT = 256;  // threads per workgroup (4 waves)
constexpr int GW = GT / 64;

__device__ __forceinline__ f16v mma(h4 a, h4 b, f16v c) {
  return __builtin_amdgcn_mfma_f32_32x32x8f16(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ h4 to_h4(float a, float b, float c, float d) {
  h4 v = {(_Float16)a, (_Float16)b, (_Float16)c, (_Float16)d};
  return v;
}

template <int Q>
struct GemmShape {
  static constexpr int EB = Apply3Shape<8, Q, double>::EB;  // the plan's block size
  static constexpr int KS = (Q + 7) / 8;                    // K-steps over a q index
  static constexpr int QP = (Q + 3) & ~3;                   // padded q-indexed row
  static constexpr int U_SZ = EB * 512;
  static constexpr int YIN_SZ = 2 * EB * 8 * Q * 8;
  static constexpr int ZIN_SZ = 3 * EB * Q * Q * 8;
  // W rows ((f EB + e) 8 + c) Q + i1; the Y^T pencils run i1 up to QP - 1 and
  // every read spans 8 KS values from the row start
  static constexpr int W_ROWS = 3 * EB * 8 * Q + (QP - Q);
  static constexpr int W_SZ = (W_ROWS - 1) * QP + 8 * KS;
  static constexpr int S_SZ = (2 * EB * 64 - 1) * QP + 8 * KS;
  static constexpr int YO_SZ = EB * 512 * 2;  // FP32 tensor, in FP16 units
  static constexpr int mx(int a, int b) { return a > b ? a : b; }
  static constexpr int RA = (mx(mx(U_SZ, ZIN_SZ), S_SZ) + 7) & ~7;
  static constexpr int RB = (mx(mx(YIN_SZ, W_SZ), YO_SZ) + 7) & ~7;
  static constexpr int BX = EB * 512;  // the block's distinct dofs (FP32)
  static constexpr int KE = EB * 512 / GT;
  static_assert(EB * 512 % GT == 0, "cell dofs per thread");
  static_assert(Q >= 8 && 3 * Q <= 32, "Q in [8, 10]");
  static_assert((RA + RB) * 2 + BX * 4 <= 65536, "apply3d_gemm: LDS above 64 KiB");
};

template <int Q>
__global__ void __launch_bounds__(GT)
    k_apply3d_gemm(const float* __restrict__ x, float* __restrict__ y,
                   const uint16_t* __restrict__ pidx, const int32_t* __restrict__ bdofs,
                   const int64_t* __restrict__ boff, const int32_t* __restrict__ bexcl,
                   const int64_t* __restrict__ bcell, int nblk,
                   const uint32_t* __restrict__ Gm, const float* __restrict__ tB,
                   const float* __restrict__ tD, GemmScales sc) {
  using S = GemmShape<Q>;
  constexpr int EB = S::EB, KS = S::KS, QP = S::QP, QQ = Q * Q;
  __shared__ __align__(16) _Float16 rA[S::RA];
  __shared__ __align__(16) _Float16 rB[S::RB];
  __shared__ __align__(16) float bx[S::BX];
  __shared__ float wmax[GW];

  const int blk = xcd_remap(blockIdx.x, nblk);
  const int64_t base = bcell[blk];
  const int ncb = (int)(bcell[blk + 1] - base);
  const int tid = threadIdx.x;
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int lane = tid & 63, h = lane >> 5, r = lane & 31;
  const int64_t d0 = boff[blk];
  const int nu = (int)(boff[blk + 1] - d0);
  const int nexcl = bexcl[blk];
  const int32_t* __restrict__ dofs = bdofs + d0;

  // ---- prologue: dof ids, x, local indices ---------------------------------
  int32_t dv[S::KE];
  float xv[S::KE];
#pragma unroll
  for (int k = 0; k < S::KE; ++k) {
    const int i = tid + k * GT;
    dv[k] = i < nu ? dofs[i] : -1;
  }
#pragma unroll
  for (int k = 0; k < S::KE; ++k) xv[k] = dv[k] >= 0 ? x[dv[k]] : 0.f;
  // this thread's cell dofs, tensor order e * 512 + a + 8 b + 64 c
  uint16_t pv[S::KE];
#pragma unroll
  for (int k = 0; k < S::KE; ++k) {
    const int i = tid + k * GT;
    pv[k] = (i >> 9) < ncb ? pidx[base * 512 + i] : kNoDof;
  }
  // zeroed LDS: whatever a padded or spilling read meets is finite
  {
    const uint4 z = {0u, 0u, 0u, 0u};
    for (int i = tid; i < S::RA / 8; i += GT) ((uint4*)rA)[i] = z;
    for (int i = tid; i < S::RB / 8; i += GT) ((uint4*)rB)[i] = z;
  }
  // table operands (per-lane constants), each times the power-of-two
  // scale its output takes (exact in FP16): forward by the output column's
  // table (B, D; B, D, B); Z^T and the D path of Y^T by tB / tD; the two
  // terms of the B path of Y^T by tD / 2 and tB / 2; the B path of X^T by
  // 2 tB / tD, which brings it to the scale of the D path
  h4 opX, opY, aB, aD, TBz[KS], TDz[KS], TBh[KS], TDh[KS], TD[KS], TBx[KS];
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int k = 4 * h + j;  // a node index (< 8)
    float vx = 0.f, vy = 0.f, vb = 0.f, vd = 0.f;
    if (r < Q) {
      vx = tB[r * 8 + k] * sc.pB;
      vy = vx;
      vb = tB[r * 8 + k];
      vd = tD[r * 8 + k];
    } else if (r < 2 * Q) {
      vx = tD[(r - Q) * 8 + k] * sc.pD;
      vy = vx;
    } else if (r < 3 * Q) {
      vy = tB[(r - 2 * Q) * 8 + k] * sc.pB;
    }
    opX[j] = (_Float16)vx;
    opY[j] = (_Float16)vy;
    aB[j] = (_Float16)vb;
    aD[j] = (_Float16)vd;
#pragma unroll
    for (int s = 0; s < KS; ++s) {
      const int i = 8 * s + k;  // a quadrature index
      const bool ok = r < 8 && i < Q;
      const float b = ok ? tB[i * 8 + r] : 0.f, d = ok ? tD[i * 8 + r] : 0.f;
      TBz[s][j] = (_Float16)(b * sc.tB);
      TDz[s][j] = (_Float16)(d * sc.tD);
      TBh[s][j] = (_Float16)(b * sc.hB);
      TDh[s][j] = (_Float16)(d * sc.hD);
      TD[s][j] = (_Float16)d;
      TBx[s][j] = (_Float16)(b * sc.bx);
    }
  }
  // the block's max |x| -> its power-of-two scale
  {
    float m = 0.f;
#pragma unroll
    for (int k = 0; k < S::KE; ++k) m = fmaxf(m, fabsf(xv[k]));
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o));
    if (lane == 0) wmax[w] = m;
  }
#pragma unroll
  for (int k = 0; k < S::KE; ++k) {
    const int i = tid + k * GT;
    if (i < nu) bx[i] = xv[k];
  }
  __syncthreads();
  float mx = wmax[0];
#pragma unroll
  for (int i = 1; i < GW; ++i) mx = fmaxf(mx, wmax[i]);
  int ex = 0;
  (void)frexpf(mx, &ex);
  // mx sx in [2^13, 2^14): every later conversion keeps its bound at 2^14
  // (FP16's largest finite value is about 2^16), so values down to 2^-28 of
  // the bound stay normal
  const float sx = mx > 0.f ? ldexpf(1.f, 14 - ex) : 1.f;
  const float isx = 1.f / sx;  // a power of two: exact
  // u tensor in FP16, a fastest
#pragma unroll
  for (int k = 0; k < S::KE; ++k) {
    const int i = tid + k * GT;
    rA[i] = (_Float16)(pv[k] != kNoDof ? bx[pv[k]] * sx : 0.f);
  }
  __syncthreads();
  // the x values are consumed: bx becomes the block's y accumulator (read
  // again only after several barriers)
#pragma unroll
  for (int k = 0; k < S::KE; ++k) {
    const int i = tid + k * GT;
    if (i < nu) bx[i] = 0.f;
  }
  const f16v zero = {};
  // G of the wave's first Z tile, loaded now so that its latency passes
  // under the X and Y stages (the next tile's is loaded one tile ahead)
  const int npz = QQ * ncb, ntz = (npz + 31) / 32;
  uint32_t gl[KS][4][3], gn[KS][4][3];
  auto load_g = [&](int t, uint32_t (&g)[KS][4][3]) {
    const int p = 32 * t + r;
    const bool ok = p < npz;
    const int pp = ok ? p : 0;
    const int e = pp / QQ, pz = pp - e * QQ;
    const uint32_t* gp = Gm + (base + e) * 3 * QQ * Q + pz;
#pragma unroll
    for (int i = 0; i < KS; ++i)
#pragma unroll
      for (int jj = 0; jj < 4; ++jj) {
        const int i3 = 8 * i + 4 * h + jj;
        const bool v = ok && i3 < Q;
#pragma unroll
        for (int k = 0; k < 3; ++k)
          g[i][jj][k] = v ? __builtin_nontemporal_load(gp + (i3 * 3 + k) * QQ) : 0u;
      }
  };
  if (w < ntz) load_g(w, gl);

  // ---- X: pencils p = b + 8 c + 64 e, K = a -> Yin[t][e][c][i1][b] ----------
  for (int t = w; t < 2 * ncb; t += GW) {
    const int p = 32 * t + r;
    const f16v cx = mma(*(const h4*)(rA + p * 8 + 4 * h), opX, zero);
    if (r < 2 * Q) {
      const int tt = r < Q ? 0 : 1, i1 = r - tt * Q;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int ce = 4 * t + i, c = ce & 7, e = ce >> 3;
        *(h4*)(rB + ((((tt * EB + e) * 8 + c) * Q + i1) * 8 + 4 * h)) =
            to_h4(cx[4 * i], cx[4 * i + 1], cx[4 * i + 2], cx[4 * i + 3]);
      }
    }
  }
  __syncthreads();

  // ---- Y: per t, pencils p = c + 8 (i1 + Q e), K = b -> Zin -----------------
  {
    const int np = 8 * Q * ncb, nt = (np + 31) / 32;
    for (int u = w; u < 2 * nt; u += GW) {
      const int tt = u < nt ? 0 : 1, t = u - tt * nt;
      const int p = 32 * t + r;
      const int pp = p < np ? p : 0;
      const int c = pp & 7, rest = pp >> 3, i1 = rest % Q, e = rest / Q;
      const f16v cy =
          mma(*(const h4*)(rB + ((((tt * EB + e) * 8 + c) * Q + i1) * 8 + 4 * h)), opY, zero);
      // path 0 = DxBy (from Dx, By), 1 = BxDy, 2 = BxBy (from Bx)
      int path = -1, i2 = 0;
      if (tt == 0) {
        if (r < Q) {
          path = 2;
          i2 = r;
        } else if (r < 2 * Q) {
          path = 1;
          i2 = r - Q;
        }
      } else if (r >= 2 * Q && r < 3 * Q) {
        path = 0;
        i2 = r - 2 * Q;
      }
      if (path >= 0) {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int ro = 4 * t + i;  // i1 + Q e of the group's four rows
          if (ro * 8 < np) {
            const int i1o = ro % Q, eo = ro / Q;
            *(h4*)(rA + (((path * EB + eo) * QQ + i1o + Q * i2) * 8 + 4 * h)) =
                to_h4(cy[4 * i], cy[4 * i + 1], cy[4 * i + 2], cy[4 * i + 3]);
          }
        }
      }
    }
  }
  __syncthreads();

  // ---- Z, G, Z^T: pencils p = (i1 + Q i2) + Q^2 e; K = c, then K = i3 -------
  {
    for (int t = w; t < ntz; t += GW) {
      const int p = 32 * t + r;
      const bool ok = p < npz;
      const int pp = ok ? p : 0;
      const int e = pp / QQ, pz = pp - e * QQ;
      // gl: G of the lane's points i3 = 8 i + 4 h + jj (i < KS), pairs
      // xx xy | xz yy | yz zz; the next tile's into gn meanwhile
      if (t + GW < ntz) load_g(t + GW, gn);
      const _Float16* zp = rA + (e * QQ + pz) * 8 + 4 * h;
      const f16v gx = mma(aB, *(const h4*)(zp), zero);
      const f16v gy = mma(aB, *(const h4*)(zp + EB * QQ * 8), zero);
      const f16v gz = mma(aD, *(const h4*)(zp + 2 * EB * QQ * 8), zero);
      f16v wx = zero, wy = zero, wz = zero;
#pragma unroll
      for (int s = 0; s < KS; ++s) {
        // rows 8 s + 4 h + jj: all below Q for s = 0; for the last K-step
        // only jj < Q - 8 s can be (lanes h = 1 read zero G there)
        constexpr int JV0 = 4;
        const int jv = s == 0 ? JV0 : (Q - 8 * s < 4 ? Q - 8 * s : 4);
        float fx[4] = {0.f, 0.f, 0.f, 0.f}, fy[4] = {0.f, 0.f, 0.f, 0.f},
              fz[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int jj = 0; jj < 4; ++jj) {
          if (jj >= jv) continue;
          const float g0 = __uint_as_float(gl[s][jj][0] << 16);
          const float g1 = __uint_as_float(gl[s][jj][0] & 0xffff0000u);
          const float g2 = __uint_as_float(gl[s][jj][1] << 16);
          const float g3 = __uint_as_float(gl[s][jj][1] & 0xffff0000u);
          const float g4 = __uint_as_float(gl[s][jj][2] << 16);
          const float g5 = __uint_as_float(gl[s][jj][2] & 0xffff0000u);
          const int k = 4 * s + jj;
          // the components at one scale (cx, cy, cz include the flux's);
          // G is zero at the lane's invalid points, so the flux is too
          const float ux = gx[k] * sc.cx, uy = gy[k] * sc.cy, uz = gz[k] * sc.cz;
          fx[jj] = g0 * ux + g1 * uy + g2 * uz;
          fy[jj] = g1 * ux + g3 * uy + g4 * uz;
          fz[jj] = g2 * ux + g4 * uy + g5 * uz;
        }
        wx = mma(TBz[s], to_h4(fx[0], fx[1], fx[2], fx[3]), wx);
        wy = mma(TBz[s], to_h4(fy[0], fy[1], fy[2], fy[3]), wy);
        wz = mma(TDz[s], to_h4(fz[0], fz[1], fz[2], fz[3]), wz);
      }
      // rows c = 4 h + jj (register group 0), column = the lane's pencil
      if (ok) {
        const int i2 = pz / Q, i1 = pz - i2 * Q;
        _Float16* o = rB + (((e * 8 + 4 * h) * Q + i1) * QP + i2);
        constexpr int FS = EB * 8 * Q * QP, CS = Q * QP;
#pragma unroll
        for (int jj = 0; jj < 4; ++jj) {
          o[jj * CS] = (_Float16)wx[jj];
          o[FS + jj * CS] = (_Float16)wy[jj];
          o[2 * FS + jj * CS] = (_Float16)wz[jj];
        }
      }
#pragma unroll
      for (int i = 0; i < KS; ++i)
#pragma unroll
        for (int jj = 0; jj < 4; ++jj)
#pragma unroll
          for (int k = 0; k < 3; ++k) gl[i][jj][k] = gn[i][jj][k];
    }
  }
  __syncthreads();

  // ---- Y^T: pencils p = i1 + QP (c + 8 e) (i1 >= Q: padding), K = i2 -> S ---
  {
    const int nt = QP * 8 * ncb / 32;
    constexpr int FS = EB * 8 * Q * QP;
    for (int t = w; t < nt; t += GW) {
      const int p = 32 * t + r;
      const int i1 = p % QP, ce = p / QP, c = ce & 7, e = ce >> 3;
      const _Float16* wp = rB + ((e * 8 + c) * Q + i1) * QP + 4 * h;
      // S_D at scale s3 tB^2, S_B = (y term tD + z term tB) / 2 at s3 tB tD / 2
      f16v sd = zero, sb = zero;
#pragma unroll
      for (int s = 0; s < KS; ++s) {
        sd = mma(*(const h4*)(wp + 8 * s), TBz[s], sd);
        sb = mma(*(const h4*)(wp + FS + 8 * s), TDh[s], sb);
        sb = mma(*(const h4*)(wp + 2 * FS + 8 * s), TBh[s], sb);
      }
      if (r < 8) {
        constexpr int TS = EB * 64 * QP;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int pb = 32 * t + 8 * i + 4 * h;
          const int i1b = pb % QP, cb = pb / QP, cc = cb & 7, eb = cb >> 3;
          _Float16* o = rA + ((eb * 8 + cc) * 8 + r) * QP + i1b;
          *(h4*)o = to_h4(sd[4 * i], sd[4 * i + 1], sd[4 * i + 2], sd[4 * i + 3]);
          *(h4*)(o + TS) = to_h4(sb[4 * i], sb[4 * i + 1], sb[4 * i + 2], sb[4 * i + 3]);
        }
      }
    }
  }
  __syncthreads();

  // ---- X^T: pencils p = b + 8 c + 64 e, K = i1 -> y tensor [e][c][a][b] ------
  float* const yo = (float*)rB;
  {
    constexpr int TS = EB * 64 * QP;
    for (int t = w; t < 2 * ncb; t += GW) {
      const int p = 32 * t + r;
      const _Float16* sp = rA + p * QP + 4 * h;  // row ((e 8 + c) 8 + b)
      f16v yv = zero;
#pragma unroll
      for (int s = 0; s < KS; ++s) {
        yv = mma(*(const h4*)(sp + 8 * s), TD[s], yv);
        yv = mma(*(const h4*)(sp + TS + 8 * s), TBx[s], yv);
      }
      const float yscale = sc.vD * isx;
      if (r < 8) {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int ce = 4 * t + i;  // c + 8 e
          const f4v v = {yv[4 * i] * yscale, yv[4 * i + 1] * yscale, yv[4 * i + 2] * yscale,
                         yv[4 * i + 3] * yscale};
          *(f4v*)(yo + (ce * 8 + r) * 8 + 4 * h) = v;
        }
      }
    }
  }
  __syncthreads();

  // ---- assembly over the block's distinct dofs, then write-back ------------
#pragma unroll
  for (int k = 0; k < S::KE; ++k) {
    if (pv[k] == kNoDof) continue;
    const int i = tid + k * GT;
    const int e = i >> 9, kk = i & 511;
    const int a = kk & 7, b = (kk >> 3) & 7, c = kk >> 6;
    lds_atomic_add(bx + pv[k], yo[((e * 8 + c) * 8 + a) * 8 + b]);
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < S::KE; ++k) {
    const int i = tid + k * GT;
    if (dv[k] < 0) continue;
    const float v = bx[i];
    if (i < nexcl)
      y[dv[k]] = v;
    else
      atomic_add(y + dv[k], v);
  }
}

template <int Q>
hipError_t launch_gemm_q(const float* x, float* y, const uint32_t* Gm, const float* tB,
                         const float* tD, const BlockArgs& k, const GemmScales& sc,
                         hipStream_t s) {
  if (k.nblk <= 0) return hipSuccess;
  hipLaunchKernelGGL((k_apply3d_gemm<Q>), dim3(k.nblk), dim3(GT), 0, s, x, y, k.pidx, k.bdofs,
                     k.boff, k.bexcl, k.bcell, k.nblk, Gm, tB, tD, sc);
  return hipGetLastError();
}

// max |v| over bf16 pairs (positive floats order as their bit patterns)
__global__ void __launch_bounds__(256)
    k_absmax_bf16pairs(const uint32_t* __restrict__ a, int64_t n, unsigned* __restrict__ out) {
  unsigned m = 0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const unsigned v = a[i];
    const unsigned lo = (v << 16) & 0x7fffffffu, hi = v & 0x7fff0000u;
    m = lo > m ? lo : m;
    m = hi > m ? hi : m;
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const unsigned t = __shfl_xor(m, o);
    m = t > m ? t : m;
  }
  if ((threadIdx.x & 63) == 0) atomicMax(out, m);
}

}  // namespace

bool gemm_supported(int n, int q) { return n == 8 && q >= 8 && q <= 10; }

hipError_t apply_gemm(int n, int q, const float* x, float* y, const uint32_t* Gm,
                      const float* tB, const float* tD, const BlockArgs* k,
                      const GemmScales& sc, hipStream_t s) {
  if (!k) return hipErrorInvalidValue;
  if (n != 8) return hipErrorNotSupported;
  switch (q) {
    case 8:
      return launch_gemm_q<8>(x, y, Gm, tB, tD, *k, sc, s);
    case 9:
      return launch_gemm_q<9>(x, y, Gm, tB, tD, *k, sc, s);
    case 10:
      return launch_gemm_q<10>(x, y, Gm, tB, tD, *k, sc, s);
    default:
      return hipErrorNotSupported;
  }
}

hipError_t launch_absmax_bf16pairs(const uint32_t* a, int64_t n, unsigned* out, hipStream_t s) {
  hipError_t e = hipMemsetAsync(out, 0, sizeof(unsigned), s);
  if (e != hipSuccess || n <= 0) return e;
  int64_t g = (n + 255) / 256;
  if (g > 4096) g = 4096;
  hipLaunchKernelGGL(k_absmax_bf16pairs, dim3((unsigned)g), dim3(256), 0, s, a, n, out);
  return hipGetLastError();
}

}  // namespace sf
