// sf_apply.hpp -- the matrix-free, sum-factorized Laplacian apply kernels.
//
// One workgroup processes EB cells.  Per cell the work is the reference's
// Arrays.evaluate! pipeline (src/SumFactorizationMaps.jl:78-89): gather,
// dof->quadrature contractions (src/SumFactorizationKernels.jl:40-65),
// geometric factor (:174-216), quadrature->dof contractions (:105-130),
// scatter-add (the assembly of src/LazyMatrices.jl:57-58).
//
// The contraction schedule is not the reference's D x D one (each gradient
// component interpolated separately in every direction).  Directions share
// partial results: the x-pencil pass produces both B u and D u, the y pass
// forms three tensors from those two, the z pass three gradient components.
// In 3D that is 2N^3Q + 3N^2Q^2 + 3NQ^3 FMAs per direction of travel against
// the reference's 3(N^3Q + N^2Q^2 + NQ^3), and each thread keeps a whole 1D
// pencil in registers, so every LDS value it reads feeds Q (or N) FMAs.
//
// The 1D tables are read through uniform (scalar) loads: every lane uses the
// same B[i][j] at the same time, so they live in SGPRs / the scalar cache and
// cost no LDS bandwidth.  LDS holds only the per-cell intermediate tensors.
//
// LDS per cell (3D):  A = 2 tensors N x N x Q   (x-pass output, and the
//                                                y^T-pass output later)
//                     B = 3 tensors N x Q x Q   (y-pass output, rewritten in
//                                                place by the z column pass)
// The strides of both groups (LdsLayout) are chosen per (N, Q) so that the
// lanes of each stage hit distinct LDS banks: FP64 from sf_lds_layout.h
// (gen_lds_layout.py scores every stage's lane pattern against the 64-bit
// bank model of gfx950: 64 banks of 4 bytes, 16 lanes per pass); otherwise
// dense tensors with the x-pencil rows padded to an odd length QP.
// From N = 6 on (N = 5 in the fused-geometry kernel), A and B share one
// region: the y pass and the y^T pass read their input pencils into
// registers, cross a barrier, then overwrite them, so a workgroup holds
// max(B, A + x values) instead of A + B (more cells per workgroup and more
// workgroups per CU -- the occupancy is what hides the gather latency at
// high order; at N = 5 measured 2-4% faster with the fused geometry, 10%
// slower with the stored G).  Below that the LDS is not what limits the
// occupancy (the registers are) and the two extra barriers would only cost.
#pragma once

#include "sf_common.hpp"
#include "sf_lds_layout.h"

namespace sf {

// LDS strides, in elements, of the per-cell tensors of the 3D apply:
//   A (2 tensors): e*CA + t*N*ZA + z*ZA + y*YA + i
//   B (3 tensors): e*CB + t*N*ZB + z*ZB + k*Q + i
template <int N, int Q, int BYTES>
struct LdsLayout {
  static constexpr int QP = (Q % 2 == 0) ? Q + 1 : Q;
  static constexpr int YA = QP, ZA = N * QP, CA = 2 * N * N * QP;
  static constexpr int ZB = Q * Q, CB = 3 * N * Q * Q;
};
#define SF_LDS_ROW(N_, Q_, YA_, ZA_, CA_, ZB_, CB_)                        \
  template <>                                                             \
  struct LdsLayout<N_, Q_, 8> {                                           \
    static constexpr int YA = YA_, ZA = ZA_, CA = CA_, ZB = ZB_, CB = CB_; \
  };
SF_LDS_LAYOUT_F64(SF_LDS_ROW)
#undef SF_LDS_ROW

template <int N, int Q, typename T, bool GEO = true>
struct Apply3Shape {
  static constexpr int NN = N * N, NQ = N * Q, QQ = Q * Q, QQQ = Q * Q * Q;
  static constexpr int NNN = N * N * N;
  static constexpr int QP = (Q % 2 == 0) ? Q + 1 : Q;
  static constexpr int SA1 = NN * QP;  // one x-pass tensor [z][y][i]
  static constexpr int SB1 = N * QQ;   // one y-pass tensor [z][k][i]
  static constexpr int SE = 2 * SA1 + 3 * SB1;
  static constexpr int MAXT = cmax(NN, cmax(NQ, QQ));
  // shared A / B region (N >= 6, and N = 5 with the fused geometry; the
  // stored-G kernel at N = 5 is faster without the extra barriers): LDS per
  // cell is the y-pass output, or the x-pass output next to the cell's x
  // values.  FP64 fused geometry at N = 3 and 4 too: 35 -> 24 KB and
  // 40 -> 28 KB per workgroup, five workgroups per CU instead of four (Q2
  // -9%; Q3 -4% with the dof ids in LDS and a 5-wave cap, below; in FP32
  // it measured no better)
  static constexpr bool LC =
      N >= 6 || (N == 5 && GEO) || ((N == 3 || N == 4) && GEO && sizeof(T) == 8);
  static constexpr int SC = LC ? cmax(3 * SB1, 2 * SA1 + NNN) : SE;
  // cells per workgroup: fill ~256 lanes in the widest stage, within 64 KiB
  // of FP64 LDS (the same EB for every T, so FP64 and FP32 applies share the
  // host-side block structure)
  static constexpr int EB_T = MAXT > 100 ? 1 : cmax(1, 256 / MAXT);
  static constexpr int EB_L = cmax(1, 65536 / (SC * 8));
  // table rows per scalar-load group (measured: 3 at N = 5, 2 at N = 6..8;
  // one row per load is faster below N = 5, and larger groups at large N, Q
  // spill SGPRs into VGPR lanes)
  static constexpr int TK = N <= 4 ? 1 : (N == 5 ? 3 : (N <= 8 && Q <= N + 2 ? 2 : 1));
  static constexpr int EB = EB_T < EB_L ? EB_T : EB_L;
  static constexpr int NT = round_up(EB * MAXT, 64);
  // minimum waves per SIMD the compiler must fit the FP64 kernel into.  At
  // N = 7 and, fused geometry only, (N = 8, Q <= 9) and (N = 9, Q <= 10) it
  // used 170-206 VGPRs (2 waves per SIMD) while the LDS allows 3
  // workgroups per CU; held to 168 the third wave won (measured: Q6 -13%,
  // Q7 -6%, Q8 -12%).  Since the split z stage (SPLIT below) the fused
  // kernels at N = 8, 9 fit 168 by themselves (144-167 VGPRs, no spills);
  // the cap stays for N = 7, which keeps the interleaved order.  The
  // stored-G kernel spills more: with the cap it is 2% faster at Q6 but 14%
  // slower at Q7 and 26% slower at Q8.  At N = 10 the spills cost more than
  // the wave gains (Q9 with q = 11: +14%).  The FP32 kernel at N = 6, Q <= 7
  // fits five waves without spills (Q5 -7%; at N = 7, 8 the same cap was
  // +1% / +14%).  A 4-wave cap on the FP64 kappa apply at N = 6 (140 VGPRs,
  // 10 spilled under the cap) was 2% slower.  FP64 fused geometry at N = 4
  // (Q <= 5): five waves (96 VGPRs, 4 spilled) now that the shared region
  // and the ids in LDS let five workgroups fit; the kappa apply there (12
  // spills under that cap) keeps the compiler's choice.  Elsewhere the
  // compiler's own choice was as fast or faster.
  static constexpr int WPE =
      (sizeof(T) == 4 && N == 6 && Q <= 7) ? 5 :
      (sizeof(T) == 8 && GEO && N == 5 && Q <= 6) ? 4 :
      (sizeof(T) == 8 && GEO && N == 4 && Q <= 5) ? 5 :
      (sizeof(T) == 8 && (N == 7 || (GEO && ((N == 8 && Q <= 9) || (N == 9 && Q <= 10))))) ? 3
                                                                                            : 1;
  // fused geometry, z stage: SPLIT forms all q gradients of the column
  // first and then applies G point by point in place, so the geometry's
  // registers are never live beside the 3N contraction inputs (FP64 at N =
  // 6: 126 instead of 152 VGPRs, a fourth wave per SIMD; at N = 8 no
  // spills instead of 7).  Otherwise G is applied as each gradient is
  // formed, which gives the scheduler both to interleave: where the
  // registers do not change the waves (FP64 at N = 5 and 7, FP32 below
  // N = 6) that was 2-3% faster (measured A/B/A/B).
  static constexpr bool SPLIT = GEO && (sizeof(T) == 8 ? (N != 5 && N != 7) : N >= 6);
  // the apply kernel's LDS strides (SA1 / SB1 / QP above size EB and serve
  // the adjoint kernel's dense layout).  The FP64 table assumes the shared
  // region from N = 5 on; the stored-G kernel at N = 5 (split regions)
  // keeps the dense layout, which fits three workgroups per CU.
  using L = LdsLayout<N, Q, (LC || N < 5) ? (int)sizeof(T) : 0>;
  static constexpr int YA = L::YA, ZA = L::ZA, TA = N * ZA, CA = L::CA;
  static constexpr int ZB = L::ZB, TB = N * ZB, CB = L::CB;
  static_assert(YA >= Q && ZA >= N * YA && CA >= 2 * TA && ZB >= QQ && CB >= 3 * TB,
                "apply3d: overlapping LDS tensors");
  // LC: one region of RC values -- the x-pass output (A, EB x CA) at its
  // start with the block's gathered x values (at most EB * N^3 distinct
  // dofs) behind it; later the y-pass output (B, EB x CB) over both;
  // later the y^T-pass output over A and the assembled y values at the end.
  // Otherwise A, then B; B also holds the x values and later the y values.
  static constexpr int RA = EB * CA;
  static constexpr int RX = EB * NNN;
  static constexpr int RC = LC ? cmax(EB * CB, RA + RX) : RA + cmax(EB * CB, RX);
  // FP64 fused geometry at N = 5 (Q <= 6) and N = 8 (Q <= 9): the block's
  // dof ids wait for the write-back in LDS (4 bytes per distinct dof), not
  // in KU registers.  At N = 5 that and a 4-wave cap (128 VGPRs, 2-3
  // spilled) give a fourth workgroup per CU (Q4 -4%); at N = 8 it halves
  // the spills (Q7 -3.5%).  Elsewhere the extra LDS costs a workgroup per
  // CU (Q3 +8%, Q6 +20%; measured A/B/A/B).
  // FP64 at N = 4 (Q <= 5), with the shared region and a 5-wave cap: 96
  // VGPRs with 4 spilled (7 without the ids in LDS), Q3 -4%.
  // FP32 at N = 8 (Q <= 9), with the split z stage: the CG form goes from
  // 115 to 95 VGPRs, five waves per SIMD, the 32^3 Q7 mixed PCG iteration -2%.
  static constexpr bool DVL =
      (sizeof(T) == 8 && GEO && ((N == 4 && Q <= 5) || (N == 5 && Q <= 6) || (N == 8 && Q <= 9))) ||
      (sizeof(T) == 4 && GEO && N == 8 && Q <= 9);
  static constexpr int LDS_BYTES = RC * (int)sizeof(T) + (DVL ? RX * 4 : 0);
  // distinct dofs of a block handled per thread in the gather / write-back
  static constexpr int KU = (EB * NNN + NT - 1) / NT;
  static_assert(LDS_BYTES <= 65536, "apply3d: LDS per workgroup above 64 KiB");
  static_assert(NT >= EB * MAXT, "one task per thread per stage");
};

constexpr uint16_t kNoDof = 0xffff;  // local index of a Dirichlet dof

// Reduced-precision (FP32 / bf16) 3D G pairs rows 2m, 2m+1 of the block-
// interleaved layout: entry (k = c*Q + l) of lane t of block b sits at
// bstart*6Q^3 + (k/2)*2*ncb*Q^2 + 2*t + (k%2).
__host__ __device__ inline int64_t g3_pair_index(int64_t bstart, int ncb,
                                                 int64_t cell, int c, int l,
                                                 int r, int Q) {
  const int64_t QQ = (int64_t)Q * Q;
  const int k = c * Q + l;
  return bstart * 6 * QQ * Q + (int64_t)(k / 2) * 2 * ncb * QQ +
         2 * ((cell - bstart) * QQ + r) + (k % 2);
}

// Index of G entry c (0..5) at quadrature point qp = l*Q*Q + r of `cell` in
// the block-interleaved 3D layout: block b (cells bcell[b] .. bcell[b+1]-1,
// ncb of them) stores its G contiguously as [c*Q + l][cell in block][r],
// so the stage-3 thread (cell e, column r) reads one value per (c, l) at a
// fixed stride and every wave-wide load covers contiguous bytes.
__host__ __device__ inline int64_t g3_index(int64_t bstart, int ncb, int64_t cell,
                                            int c, int l, int r, int Q) {
  const int64_t QQ = (int64_t)Q * Q;
  return bstart * 6 * QQ * Q + ((int64_t)(c * Q + l) * ncb + (cell - bstart)) * QQ + r;
}

// 1 / a for a > 0: the hardware reciprocal estimate refined by two Newton
// steps r <- r (2 - a r) (each doubles the correct bits), instead of the
// IEEE division sequence.
__device__ __forceinline__ double recip_pos(double a) {
  double r = __builtin_amdgcn_rcp(a);
  r = fma(r, fma(-a, r, 1.0), r);
  r = fma(r, fma(-a, r, 1.0), r);
  return r;
}
__device__ __forceinline__ float recip_pos(float a) {
  float r = __builtin_amdgcn_rcpf(a);
  return fmaf(r, fmaf(-a, r, 1.0f), r);
}

// The fused-geometry kernels read each cell's multilinear map as its
// coefficient vectors (computed once per plan from the vertices, vertex
// v = b0 + 2 b1 + 4 b2):
//   3D  x(xi, eta, zeta) = V0 + xi a + eta e + zeta p + xi eta b + xi zeta c
//                          + eta zeta g + xi eta zeta d,   stored a e p b c g d
//   2D  x(xi, eta) = V0 + xi a + eta e + xi eta b,         stored a e b
// so the derivatives along a quadrature column are a few FMAs each, not
// bilinear interpolations of vertex differences.
constexpr int kGeoCoef3 = 21, kGeoCoef2 = 6;

// Edge vectors of the 3D map along quadrature column (xi, eta):
// dx/dxi = E0 + zeta dE, dx/deta = F0 + zeta dF, dx/dzeta = H (constant).
template <typename T>
__device__ __forceinline__ void column_geom3(const T* C, T xi, T et, T (&E0)[3], T (&dE)[3],
                                             T (&F0)[3], T (&dF)[3], T (&H)[3]) {
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const T ca = C[a], ce = C[3 + a], cp = C[6 + a], cb = C[9 + a], cc = C[12 + a],
            cg = C[15 + a], cd = C[18 + a];
    E0[a] = fma(et, cb, ca);  // dx/dxi   = a + eta b + zeta (c + eta d)
    dE[a] = fma(et, cd, cc);
    F0[a] = fma(xi, cb, ce);  // dx/deta  = e + xi b + zeta (g + xi d)
    dF[a] = fma(xi, cd, cg);
    H[a] = fma(et, dF[a], fma(xi, cc, cp));  // dx/dzeta = p + xi c + eta (g + xi d)
  }
}

// t = G g at one quadrature point from the columns of J, with G = w / |det J|
// C^T C (C the cofactor matrix; the geometry kernel's formula) applied as
// s C^T (C g): 9 + 9 FMAs and one scaling instead of forming the six G
// entries.
template <typename T>
__device__ __forceinline__ void apply_geom(const T (&j0)[3], const T (&j1)[3],
                                           const T (&j2)[3], T w, T g0, T g1, T g2,
                                           T& t0, T& t1, T& t2) {
  T C[3][3];
  C[0][0] = j1[1] * j2[2] - j2[1] * j1[2];
  C[0][1] = j2[1] * j0[2] - j0[1] * j2[2];
  C[0][2] = j0[1] * j1[2] - j1[1] * j0[2];
  C[1][0] = j2[0] * j1[2] - j1[0] * j2[2];
  C[1][1] = j0[0] * j2[2] - j2[0] * j0[2];
  C[1][2] = j1[0] * j0[2] - j0[0] * j1[2];
  C[2][0] = j1[0] * j2[1] - j2[0] * j1[1];
  C[2][1] = j2[0] * j0[1] - j0[0] * j2[1];
  C[2][2] = j0[0] * j1[1] - j1[0] * j0[1];
  const T det = j0[0] * C[0][0] + j1[0] * C[0][1] + j2[0] * C[0][2];
  const T sc = w * recip_pos(fabs(det));
  const T v0 = C[0][0] * g0 + C[0][1] * g1 + C[0][2] * g2;
  const T v1 = C[1][0] * g0 + C[1][1] * g1 + C[1][2] * g2;
  const T v2 = C[2][0] * g0 + C[2][1] * g1 + C[2][2] * g2;
  t0 = sc * (C[0][0] * v0 + C[1][0] * v1 + C[2][0] * v2);
  t1 = sc * (C[0][1] * v0 + C[1][1] * v1 + C[2][1] * v2);
  t2 = sc * (C[0][2] * v0 + C[1][2] * v1 + C[2][2] * v2);
}

// Block-assembled 3D apply, one launch over all cell blocks.
// tB / tD hold the Q x N table followed by its N x Q transpose: the
// transposed passes read whole rows of the transpose (one scalar load per
// row instead of one per element of a column).
//
// Workgroup w owns block b = xcd_remap(w): cells bcell[b] .. bcell[b+1] - 1,
// at most EB of them, consecutive in the caller's order.  The host listed
// the block's distinct free dofs (bdofs[boff[b] .. boff[b+1]), sorted; the
// bexcl[b] dofs no other block touches come first) and, for every cell dof,
// its index in that list (pidx; kNoDof for Dirichlet dofs).  The block
// gathers its x values once (sorted ids: coalesced), runs the five
// contraction stages, assembles its cells' contributions in LDS and writes
// each distinct dof once: exclusive dofs with plain stores, dofs shared
// with other blocks with one hardware atomic add (consecutive lanes hit
// consecutive ids).  y = A x needs y zeroed first (the caller's fill) unless
// accumulate is set, in which case exclusive dofs are read-modify-written.
//
// GEO: the geometric factor is not read but recomputed at every quadrature
// point from the cell's trilinear map (Xv: its 7 coefficient vectors, 21
// values per cell) and the 1D points / weights (tP, tW): 168 bytes per cell
// instead of 48 q^3.  The z-column thread of quadrature column (i, k) forms
// dx/dxi and dx/deta as linear functions of zeta once (column_geom3), and
// dx/dzeta is constant along its column.
//
// KAP (with GEO): a diffusion coefficient kap[cell][qp] (qp = l Q^2 + k Q
// + i) scales the quadrature weight; it is read in the prologue, 8 (FP64)
// or 4 (FP32) bytes per quadrature point instead of a kappa-scaled G.
//
// DOT (with GEO): wave w of workgroup b also writes dotp[b * NT / 64 + w],
// its share of x.Ax = sum over blocks of x . (the block's contribution):
// in the write-back each lane adds x_d times the assembled value of its
// distinct dofs d (FP32, and FP64 at N = 8 without kappa: x kept in
// registers from the prologue, 147 -> 158 VGPRs there, still 3 waves, the
// 32^3 Q7 PCG iteration -2.6%; otherwise FP64 re-reads x, which costs fewer
// registers).  A CG step then gets p.Ap without
// reading p and q again.
template <int N, int Q, typename T, typename TG, bool GEO, bool DOT, bool KAP = false>
__global__ void __launch_bounds__((Apply3Shape<N, Q, T>::NT))
__attribute__((amdgpu_waves_per_eu((KAP && sizeof(T) == 8 && N == 4) ? 1 : Apply3Shape<N, Q, T, GEO>::WPE)))
    k_apply3d(const T* __restrict__ x, T* __restrict__ y,
              const uint16_t* __restrict__ pidx,
              const int32_t* __restrict__ bdofs,
              const int64_t* __restrict__ boff,
              const int32_t* __restrict__ bexcl,
              const int64_t* __restrict__ bcell, int nblk,
              const TG* __restrict__ G, const T* __restrict__ tB,
              const T* __restrict__ tD, const T* __restrict__ Xv,
              const T* __restrict__ tP, const T* __restrict__ tW, int accumulate,
              double* __restrict__ dotp, const T* __restrict__ kap) {
  using S = Apply3Shape<N, Q, T, GEO>;
  static_assert(GEO || !KAP, "apply3d: kappa per point needs the fused geometry");
  static_assert(S::EB == Apply3Shape<N, Q, T>::EB && S::NT == Apply3Shape<N, Q, T>::NT,
                "apply3d: one block structure for both forms");
  static_assert(S::EB == Apply3Shape<N, Q, double>::EB,
                "apply3d: the host builds the blocks of the FP64 shape for every form");
  __shared__ T smem[S::RC + (GEO ? S::EB * kGeoCoef3 : 0)];
  __shared__ int32_t sdof[S::DVL ? S::RX : 1];  // DVL: the block's dof ids
  T* const regV = smem + S::RC;  // GEO: the block's map coefficients
  T* const regA = smem;                               // x / y^T pass out
  T* const regB = S::LC ? smem : smem + S::RA;        // y pass out, z in place
  T* const regX = S::LC ? smem + S::RA : regB;        // the block's x values
  T* const regY = S::LC ? smem + (S::RC - S::RX) : regB;  // its assembled y
  const int b = xcd_remap(blockIdx.x, nblk);
  const int64_t base = bcell[b];
  const int ncb = (int)(bcell[b + 1] - base);  // cells in this block, <= EB
  const int tid = threadIdx.x;
  const int64_t d0 = boff[b];
  const int nu = (int)(boff[b + 1] - d0);
  const int nexcl = bexcl[b];
  const int32_t* __restrict__ dofs = bdofs + d0;

  // ---- prologue: every global read of the block, issued at once ---------
  // In the order the stages need them (dof ids, x values, local indices,
  // G), so waiting for the first ones leaves the G stream in flight while
  // the gather and the first two contraction stages run.  The block
  // metadata (dof ids, local indices) is read once per apply: non-temporal
  // loads keep it from displacing x lines that neighbouring blocks gather
  // again (measured 1-3% faster at Q3-Q6; the vertex loads below the same,
  // 1-2.5% at Q2-Q4).
  int32_t dv[S::KU];
  T xv[S::KU];
#pragma unroll
  for (int k = 0; k < S::KU; ++k) {
    const int i = tid + k * S::NT;
    dv[k] = i < nu ? __builtin_nontemporal_load(dofs + i) : -1;
  }
#pragma unroll
  for (int k = 0; k < S::KU; ++k) xv[k] = x[dv[k] >= 0 ? dv[k] : 0];

  uint16_t li[N];
  const bool own1 = tid < S::EB * S::NN;
  const int e1 = tid / S::NN, r1 = tid - e1 * S::NN;  // r1 = z*N + y
  const int o1 = (r1 / N) * S::ZA + (r1 % N) * S::YA;  // its x-pencil in A
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
  T gq[GEO ? 1 : 6][GEO ? 1 : Q];
  T kq[KAP ? Q : 1];
  if constexpr (GEO) {
    // the block's map coefficients into LDS (ncb * 21 consecutive values:
    // more than NT when a block holds more than NT / 21 cells, up to 64 at
    // the lowest orders)
    for (int t = tid; t < ncb * kGeoCoef3; t += S::NT)
      regV[t] = __builtin_nontemporal_load(Xv + base * kGeoCoef3 + t);
    if constexpr (KAP) {
      // kappa of column r3 of cell e3 at every z level l
      const T* kp = kap + (base + (act3 ? e3 : 0)) * S::QQQ + (act3 ? r3 : 0);
#pragma unroll
      for (int l = 0; l < Q; ++l)
        kq[l] = act3 ? __builtin_nontemporal_load(kp + l * S::QQ) : T(0);
    }
  } else if constexpr (sizeof(TG) == 8) {
    // block-interleaved G: thread (e3, r3) is lane tid of every [c*Q + l] row
    const TG* g = G + base * 6 * S::QQQ + (act3 ? tid : 0);
    const int kst = ncb * S::QQ;
#pragma unroll
    for (int l = 0; l < Q; ++l)
#pragma unroll
      for (int c = 0; c < 6; ++c)
        gq[c][l] = act3 ? load_g<TG, T>(g + (c * Q + l) * kst) : T(0);
  } else {
    // reduced-precision G, pair-interleaved: rows 2m and 2m+1 sit side by
    // side for every lane, so one 4-byte (bf16) or 8-byte (FP32) load per
    // lane brings two entries
    const TG* g = G + base * 6 * S::QQQ + (act3 ? 2 * tid : 0);
    const int kst = 2 * ncb * S::QQ;
#pragma unroll
    for (int m = 0; m < 3 * Q; ++m) {
      T lo = T(0), hi = T(0);
      if (act3) load_g_pair<TG, T>(g + m * kst, &lo, &hi);
      gq[(2 * m) / Q][(2 * m) % Q] = lo;
      gq[(2 * m + 1) / Q][(2 * m + 1) % Q] = hi;
    }
  }

  // ---- stage 0: the block's distinct x values into LDS --------------------
#pragma unroll
  for (int k = 0; k < S::KU; ++k) {
    const int i = tid + k * S::NT;
    if (i < nu) {
      regX[i] = xv[k];
      if constexpr (S::DVL) sdof[i] = dv[k];
    }
  }
  __syncthreads();

  // ---- stage 1: x-pencils:  Bx u, Dx u ----------------------------------
  // The thread of x-pencil (e1, r1) also scatters it in stage 5.
  if (own1) {
    T u[N];
#pragma unroll
    for (int j = 0; j < N; ++j) u[j] = li[j] != kNoDof ? regX[li[j]] : T(0);
    T* sA = regA + e1 * S::CA + o1;
    TableRows<T, S::TK> t_(tB, tD);
#pragma unroll
    for (int i = 0; i < Q; ++i) {
      t_.at(i);
      T bq = T(0), dq = T(0);
#pragma unroll
      for (int j = 0; j < N; ++j) {
        bq = fma(t_.b[i * N + j], u[j], bq);
        dq = fma(t_.d[i * N + j], u[j], dq);
      }
      sA[i] = bq;
      sA[S::TA + i] = dq;
    }
  }
  __syncthreads();

  // ---- stage 2: y-pencils: (Dx u, Bx u) -> By Dx u, Dy Bx u, By Bx u ------
  // (LC: pencils into registers, barrier, outputs over the A region)
  if (S::LC || tid < S::EB * S::NQ) {  // LC: every thread meets the barrier
    const bool on2 = tid < S::EB * S::NQ;
    const int e = on2 ? tid / S::NQ : 0, r = on2 ? tid - e * S::NQ : 0;
    const int z = r / Q, i = r - z * Q;
    const T* sA = regA + e * S::CA + z * S::ZA + i;
    T* sB = regB + e * S::CB + z * S::ZB + i;
    T a[N], d[N];
#pragma unroll
    for (int j = 0; j < N; ++j) {
      a[j] = sA[j * S::YA];
      d[j] = sA[S::TA + j * S::YA];
    }
    if constexpr (S::LC) __syncthreads();
    TableRows<T, S::TK> t_(tB, tD);
    if (on2)
#pragma unroll
    for (int k = 0; k < Q; ++k) {
      t_.at(k);
      T bb = T(0), bd = T(0), db = T(0);
#pragma unroll
      for (int j = 0; j < N; ++j) {
        bb = fma(t_.b[k * N + j], a[j], bb);
        bd = fma(t_.d[k * N + j], a[j], bd);
        db = fma(t_.b[k * N + j], d[j], db);
      }
      sB[k * Q] = db;                 // x-gradient path
      sB[S::TB + k * Q] = bd;         // y-gradient path
      sB[2 * S::TB + k * Q] = bb;     // z-gradient path
    }
  }
  __syncthreads();

  // ---- stage 3: z-columns: gradients, G, transposed z contraction --------
  if (own3) {
    T* sB = regB + e3 * S::CB + r3;
    T cx[N], cy[N], cz[N];
#pragma unroll
    for (int z = 0; z < N; ++z) {
      cx[z] = sB[z * S::ZB];
      cy[z] = sB[S::TB + z * S::ZB];
      cz[z] = sB[2 * S::TB + z * S::ZB];
    }
    // GEO: edge vectors of the trilinear map for this quadrature column
    // (i, k): dx/dxi = E0 + zeta dE, dx/deta = F0 + zeta dF, dx/dzeta = H;
    // formed before the gradients, or after them (SPLIT)
    T tx[Q], ty[Q], tz[Q];
    T E0[3], dE[3], F0[3], dF[3], H[3], wik = T(0);
    constexpr bool SPLIT = S::SPLIT;
    if constexpr (GEO && !SPLIT) {
      const int qi = r3 % Q, qk = r3 / Q;
      wik = tW[qi] * tW[qk];
      column_geom3<T>(regV + e3 * kGeoCoef3, tP[qi], tP[qk], E0, dE, F0, dF, H);
    }
    TableRows<T, S::TK> t_(tB, tD);
#pragma unroll
    for (int l = 0; l < Q; ++l) {
      t_.at(l);
      T gx = T(0), gy = T(0), gz = T(0);
#pragma unroll
      for (int z = 0; z < N; ++z) {
        gx = fma(t_.b[l * N + z], cx[z], gx);
        gy = fma(t_.b[l * N + z], cy[z], gy);
        gz = fma(t_.d[l * N + z], cz[z], gz);
      }
      if constexpr (SPLIT) {
        tx[l] = gx;
        ty[l] = gy;
        tz[l] = gz;
      } else if constexpr (GEO) {
        const T ze = tP[l];
        T j0[3], j1[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
          j0[a] = fma(ze, dE[a], E0[a]);
          j1[a] = fma(ze, dF[a], F0[a]);
        }
        const T wl = KAP ? wik * tW[l] * kq[l] : wik * tW[l];
        apply_geom<T>(j0, j1, H, wl, gx, gy, gz, tx[l], ty[l], tz[l]);
      } else {
        tx[l] = gq[0][l] * gx + gq[1][l] * gy + gq[2][l] * gz;
        ty[l] = gq[1][l] * gx + gq[3][l] * gy + gq[4][l] * gz;
        tz[l] = gq[2][l] * gx + gq[4][l] * gy + gq[5][l] * gz;
      }
    }
    if constexpr (SPLIT) {
      const int qi = r3 % Q, qk = r3 / Q;
      wik = tW[qi] * tW[qk];
      column_geom3<T>(regV + e3 * kGeoCoef3, tP[qi], tP[qk], E0, dE, F0, dF, H);
#pragma unroll
      for (int l = 0; l < Q; ++l) {
        const T ze = tP[l];
        T j0[3], j1[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
          j0[a] = fma(ze, dE[a], E0[a]);
          j1[a] = fma(ze, dF[a], F0[a]);
        }
        const T wl = KAP ? wik * tW[l] * kq[l] : wik * tW[l];
        apply_geom<T>(j0, j1, H, wl, tx[l], ty[l], tz[l], tx[l], ty[l], tz[l]);
      }
    }
    TableRows<T, S::TK> tt_(tB + Q * N, tD + Q * N);  // transposed tables
#pragma unroll
    for (int z = 0; z < N; ++z) {
      tt_.at(z);
      T rx = T(0), ry = T(0), rz = T(0);
#pragma unroll
      for (int l = 0; l < Q; ++l) {
        rx = fma(tt_.b[z * Q + l], tx[l], rx);
        ry = fma(tt_.b[z * Q + l], ty[l], ry);
        rz = fma(tt_.d[z * Q + l], tz[l], rz);
      }
      sB[z * S::ZB] = rx;
      sB[S::TB + z * S::ZB] = ry;
      sB[2 * S::TB + z * S::ZB] = rz;
    }
  }
  __syncthreads();

  // ---- stage 4: transposed y-pencils (LC: into registers, barrier, over B)
  if (S::LC || tid < S::EB * S::NQ) {
    const bool on4 = tid < S::EB * S::NQ;
    const int e = on4 ? tid / S::NQ : 0, r = on4 ? tid - e * S::NQ : 0;
    const int z = r / Q, i = r - z * Q;
    T* sA = regA + e * S::CA + z * S::ZA + i;
    const T* sB = regB + e * S::CB + z * S::ZB + i;
    T rx[Q], ry[Q], rz[Q];
#pragma unroll
    for (int k = 0; k < Q; ++k) {
      rx[k] = sB[k * Q];
      ry[k] = sB[S::TB + k * Q];
      rz[k] = sB[2 * S::TB + k * Q];
    }
    if constexpr (S::LC) __syncthreads();
    TableRows<T, S::TK> t_(tB + Q * N, tD + Q * N);  // transposed tables
    if (on4)
#pragma unroll
    for (int j = 0; j < N; ++j) {
      t_.at(j);
      T sx = T(0), so = T(0);
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

  // region B (LC: the end of the region, behind the y^T-pass output)
  // accumulates y
#pragma unroll
  for (int k = 0; k < S::KU; ++k) {
    const int i = tid + k * S::NT;
    if (i < nu) regY[i] = T(0);
  }
  __syncthreads();

  // ---- stage 5: transposed x-pencils, assembled in LDS ----------------------
  if (act1) {
    const T* sA = regA + e1 * S::CA + o1;
    T sx[Q], so[Q];
#pragma unroll
    for (int i = 0; i < Q; ++i) {
      sx[i] = sA[i];
      so[i] = sA[S::TA + i];
    }
    TableRows<T, S::TK> t_(tB + Q * N, tD + Q * N);  // transposed tables
#pragma unroll
    for (int m = 0; m < N; ++m) {
      t_.at(m);
      T v = T(0);
#pragma unroll
      for (int i = 0; i < Q; ++i) {
        v = fma(t_.d[m * Q + i], sx[i], v);
        v = fma(t_.b[m * Q + i], so[i], v);
      }
      if (li[m] != kNoDof) lds_atomic_add(regY + li[m], v);
    }
  }
  __syncthreads();

  // ---- write-back: one global access per distinct dof of the block --------
  if constexpr (S::DVL) {
#pragma unroll
    for (int k = 0; k < S::KU; ++k) {
      const int i = tid + k * S::NT;
      dv[k] = i < nu ? sdof[i] : -1;
    }
  }
  T xr[DOT ? S::KU : 1];
  if constexpr (DOT) {
#pragma unroll
    for (int k = 0; k < S::KU; ++k)
      xr[k] = (sizeof(T) == 4 || (N == 8 && !KAP)) ? xv[k] : (dv[k] >= 0 ? x[dv[k]] : T(0));
  }
  double ep = 0.0;
#pragma unroll
  for (int k = 0; k < S::KU; ++k) {
    const int i = tid + k * S::NT;
    if (dv[k] < 0) continue;
    const T v = regY[i];
    if constexpr (DOT) ep += (double)xr[k] * (double)v;
    if (i < nexcl) {
      if (accumulate)
        y[dv[k]] += v;
      else
        y[dv[k]] = v;
    } else {
      atomic_add(y + dv[k], v);
    }
  }
  if constexpr (DOT) {
    ep = wave_sum_f64(ep);
    if ((tid & 63) == 0) dotp[(int64_t)b * (S::NT / 64) + (tid >> 6)] = ep;
  }
}

// ---------------------------------------------------------------------------
// 2D: block-assembled, the design of the 3D kernel in two dimensions.  A
// workgroup takes EB consecutive cells: it gathers the block's distinct x
// values once (sorted ids: coalesced), and per cell one LDS tensor pair
// [2][N][QP] ([c][y][i]) carries the work between three stages:
//   1. x-pencils (one thread per node row y): B u, D u at the Q points;
//   2. y-columns (one thread per quadrature column i): the gradients at the
//      column's Q points, t = G g, the transposed y contraction, in place;
//   3. transposed x-pencils: each thread adds its row's N values into the
//      block's y in LDS (the x values' region, dead after stage 1).
// The write-back stores the dofs exclusive to the block and adds the
// shared ones with one atomic each (accumulate: exclusive dofs
// read-modify-written).
//
// GEO: G is recomputed at every quadrature point from the cell's bilinear
// map (its 3 coefficient vectors, 6 values per cell instead of 24 Q^2
// bytes): along the column of xi = t_i, dx/dxi = a + eta b is linear in eta
// and dx/deta = e + xi b is constant, so a point costs two FMAs for J, its
// determinant, one reciprocal and t = (w / |det J|) C^T (C g) with C the
// cofactor matrix.
// Otherwise G is read, [cell][entry][k Q + i].
template <int N, int Q, typename T>
struct Apply2Shape {
  static constexpr int NN = N * N, QQ = Q * Q;
  // row stride of the [y][i] tensors: FP32 pads an even Q to the next odd
  // length (the rows of consecutive lanes then start in different banks);
  // FP64 keeps Q -- the padding's LDS cost a workgroup per CU at Q4 and
  // Q6, and without it those orders were 8% and 17% faster (2-way bank
  // conflicts in the stage-1 writes; measured A/B/A/B on 1024^2 meshes),
  // while FP32 lost 3.5% at Q4
  static constexpr int QPAD = (Q % 2 == 0) ? Q + 1 : Q;
  static constexpr int QP = sizeof(T) == 8 ? Q : QPAD;
  static constexpr int SA1 = N * QP;                   // one tensor [y][i]
  static constexpr int SE = 2 * SA1;                   // per cell
  static constexpr int MAXT = cmax(N, Q);
  // cells per workgroup: one task per thread in the wider stage, 256 lanes,
  // within 64 KiB of FP64 LDS at the padded stride (the same EB for every
  // T: FP64 and FP32 applies share the blocks)
  static constexpr int EB_T = cmax(1, 256 / MAXT);
  static constexpr int EB_L = 65536 / ((2 * N * QPAD + NN + 8) * 8);
  static constexpr int EB_TL = EB_T < EB_L ? EB_T : EB_L;
  // at (n, q) = (5, 4) and (8, 7) (the reference's 2D benchmark at Q4 and
  // Q7, q = p) a few cells fewer per block give one more FP64 workgroup per
  // CU (160 KiB of LDS): 51 -> 48 and 32 -> 28 cells, FP64 -5% and -7%
  // (A/B/A/B).  At (6, 5) the same step (42 -> 40 cells, 4 -> 5 workgroups)
  // was 4.5% slower, so it is not a general rule.
  static constexpr int CELL_B = (2 * N * Q + NN + kGeoCoef2) * 8;
  static constexpr int W0 = 163840 / (EB_TL * CELL_B);
  static constexpr int EB_O = 163840 / ((W0 + 1) * CELL_B);
  static constexpr bool MORE_WG = (N == 5 && Q == 4) || (N == 8 && Q == 7);
  static constexpr int EB = MORE_WG ? EB_O : EB_TL;
  static constexpr int NT = round_up(EB * MAXT, 64);
  static constexpr int RX = EB * NN;  // distinct dofs of a block, at most
  static constexpr int KU = (RX + NT - 1) / NT;
  static constexpr int TK = 1;  // table rows per scalar-load group
  static constexpr int LDS_BYTES = (EB * SE + RX + EB * kGeoCoef2) * (int)sizeof(T);
  static_assert(LDS_BYTES <= 65536, "apply2d: LDS per workgroup above 64 KiB");
  static_assert(NT >= EB * MAXT, "one task per thread per stage");
};

template <int N, int Q, typename T, typename TG, bool GEO>
__global__ void __launch_bounds__((Apply2Shape<N, Q, T>::NT))
    k_apply2d(const T* __restrict__ x, T* __restrict__ y,
              const uint16_t* __restrict__ pidx,
              const int32_t* __restrict__ bdofs,
              const int64_t* __restrict__ boff,
              const int32_t* __restrict__ bexcl,
              const int64_t* __restrict__ bcell, int nblk,
              const TG* __restrict__ G, const T* __restrict__ tB,
              const T* __restrict__ tD, const T* __restrict__ Xv,
              const T* __restrict__ tP, const T* __restrict__ tW, int accumulate) {
  using S = Apply2Shape<N, Q, T>;
  static_assert(S::EB == Apply2Shape<N, Q, double>::EB,
                "apply2d: the host builds the blocks of the FP64 shape for every form");
  __shared__ T smem[S::EB * S::SE + S::RX + (GEO ? S::EB * kGeoCoef2 : 0)];
  T* const regA = smem;                    // per cell [2][N][QP]
  T* const regX = smem + S::EB * S::SE;    // the block's x values, later y
  T* const regV = regX + S::RX;            // GEO: the block's map coefficients
  const int b = xcd_remap(blockIdx.x, nblk);
  const int64_t base = bcell[b];
  const int ncb = (int)(bcell[b + 1] - base);  // cells in this block, <= EB
  const int tid = threadIdx.x;
  const int64_t d0 = boff[b];
  const int nu = (int)(boff[b + 1] - d0);
  const int nexcl = bexcl[b];
  const int32_t* __restrict__ dofs = bdofs + d0;

  // ---- prologue: every global read of the block, issued at once ---------
  int32_t dv[S::KU];
  T xv[S::KU];
#pragma unroll
  for (int k = 0; k < S::KU; ++k) {
    const int i = tid + k * S::NT;
    dv[k] = i < nu ? __builtin_nontemporal_load(dofs + i) : -1;
  }
#pragma unroll
  for (int k = 0; k < S::KU; ++k) xv[k] = x[dv[k] >= 0 ? dv[k] : 0];

  // x-pencil thread: cell e1, node row r1 (local dofs r1 N .. r1 N + N - 1)
  uint16_t li[N];
  const bool own1 = tid < S::EB * N;
  const int e1 = tid / N, r1 = tid - e1 * N;
  const bool act1 = own1 && e1 < ncb;
  if (act1) {
    const uint16_t* lp = pidx + (base + e1) * S::NN + r1 * N;
#pragma unroll
    for (int j = 0; j < N; ++j) li[j] = __builtin_nontemporal_load(lp + j);
  } else {
#pragma unroll
    for (int j = 0; j < N; ++j) li[j] = kNoDof;
  }
  // y-column thread: cell e2, quadrature column i2 (xi = t_i2)
  const bool own2 = tid < S::EB * Q;
  const int e2 = tid / Q, i2 = tid - e2 * Q;
  const bool act2 = own2 && e2 < ncb;
  T gq[GEO ? 1 : 3][GEO ? 1 : Q];
  if constexpr (GEO) {
    for (int t = tid; t < ncb * kGeoCoef2; t += S::NT)
      regV[t] = __builtin_nontemporal_load(Xv + base * kGeoCoef2 + t);
  } else {
    const TG* g = G + (base + (act2 ? e2 : 0)) * 3 * S::QQ + i2;
#pragma unroll
    for (int k = 0; k < Q; ++k)
#pragma unroll
      for (int c = 0; c < 3; ++c)
        gq[c][k] = act2 ? load_g<TG, T>(g + c * S::QQ + k * Q) : T(0);
  }

  // ---- stage 0: the block's distinct x values into LDS --------------------
#pragma unroll
  for (int k = 0; k < S::KU; ++k) {
    const int i = tid + k * S::NT;
    if (i < nu) regX[i] = xv[k];
  }
  __syncthreads();

  // ---- stage 1: x-pencils: Bx u, Dx u -------------------------------------
  if (own1) {
    T u[N];
#pragma unroll
    for (int j = 0; j < N; ++j) u[j] = li[j] != kNoDof ? regX[li[j]] : T(0);
    T* sA = regA + e1 * S::SE + r1 * S::QP;
    TableRows<T, S::TK> t_(tB, tD);
#pragma unroll
    for (int i = 0; i < Q; ++i) {
      t_.at(i);
      T bq = T(0), dq = T(0);
#pragma unroll
      for (int j = 0; j < N; ++j) {
        bq = fma(t_.b[i * N + j], u[j], bq);
        dq = fma(t_.d[i * N + j], u[j], dq);
      }
      sA[i] = bq;
      sA[S::SA1 + i] = dq;
    }
  }
  __syncthreads();

  // the x values are dead: their region accumulates the block's y
#pragma unroll
  for (int k = 0; k < S::KU; ++k) {
    const int i = tid + k * S::NT;
    if (i < nu) regX[i] = T(0);
  }

  // ---- stage 2: y-columns: gradients, G, transposed y contraction --------
  if (own2) {
    T* sA = regA + e2 * S::SE + i2;
    T a[N], d[N];
#pragma unroll
    for (int j = 0; j < N; ++j) {
      a[j] = sA[j * S::QP];
      d[j] = sA[S::SA1 + j * S::QP];
    }
    // GEO: J[c][0] = dx_c/dxi = E0[c] + eta dE[c] (= a + eta b),
    // J[c][1] = dx_c/deta = H[c] (= e + xi b) on the column xi = t_i2
    T E0[2], dE[2], H[2], wi = T(0);
    if constexpr (GEO) {
      const T xi = tP[i2];
      wi = tW[i2];
      const T* C = regV + (act2 ? e2 : 0) * kGeoCoef2;  // a, e, b
#pragma unroll
      for (int c = 0; c < 2; ++c) {
        E0[c] = C[c];
        dE[c] = C[4 + c];
        H[c] = fma(xi, C[4 + c], C[2 + c]);
      }
    }
    T tx[Q], ty[Q];
    TableRows<T, S::TK> t_(tB, tD);
#pragma unroll
    for (int k = 0; k < Q; ++k) {
      t_.at(k);
      T gx = T(0), gy = T(0);
#pragma unroll
      for (int j = 0; j < N; ++j) {
        gx = fma(t_.b[k * N + j], d[j], gx);  // d/dxi: D in x, B in y
        gy = fma(t_.d[k * N + j], a[j], gy);  // d/deta: B in x, D in y
      }
      if constexpr (GEO) {
        const T et = tP[k];
        const T j00 = fma(et, dE[0], E0[0]), j10 = fma(et, dE[1], E0[1]);
        const T j01 = H[0], j11 = H[1];
        const T det = j00 * j11 - j01 * j10;
        const T sc = act2 ? wi * tW[k] * recip_pos(fabs(det)) : T(0);
        // C = [[j11, -j10], [-j01, j00]]; t = sc C^T (C g)
        const T v0 = j11 * gx - j10 * gy;
        const T v1 = j00 * gy - j01 * gx;
        tx[k] = sc * (j11 * v0 - j01 * v1);
        ty[k] = sc * (j00 * v1 - j10 * v0);
      } else {
        tx[k] = gq[0][k] * gx + gq[1][k] * gy;
        ty[k] = gq[1][k] * gx + gq[2][k] * gy;
      }
    }
    TableRows<T, S::TK> tt_(tB + Q * N, tD + Q * N);  // transposed tables
#pragma unroll
    for (int j = 0; j < N; ++j) {
      tt_.at(j);
      T rx = T(0), ry = T(0);
#pragma unroll
      for (int k = 0; k < Q; ++k) {
        rx = fma(tt_.b[j * Q + k], tx[k], rx);
        ry = fma(tt_.d[j * Q + k], ty[k], ry);
      }
      sA[j * S::QP] = rx;            // then D_x^T
      sA[S::SA1 + j * S::QP] = ry;   // then B_x^T
    }
  }
  __syncthreads();

  // ---- stage 3: transposed x-pencils, assembled in LDS --------------------
  if (act1) {
    const T* sA = regA + e1 * S::SE + r1 * S::QP;
    T sx[Q], so[Q];
#pragma unroll
    for (int i = 0; i < Q; ++i) {
      sx[i] = sA[i];
      so[i] = sA[S::SA1 + i];
    }
    TableRows<T, S::TK> t_(tB + Q * N, tD + Q * N);  // transposed tables
#pragma unroll
    for (int m = 0; m < N; ++m) {
      t_.at(m);
      T v = T(0);
#pragma unroll
      for (int i = 0; i < Q; ++i) {
        v = fma(t_.d[m * Q + i], sx[i], v);
        v = fma(t_.b[m * Q + i], so[i], v);
      }
      if (li[m] != kNoDof) lds_atomic_add(regX + li[m], v);
    }
  }
  __syncthreads();

  // ---- write-back: one global access per distinct dof of the block --------
#pragma unroll
  for (int k = 0; k < S::KU; ++k) {
    const int i = tid + k * S::NT;
    if (dv[k] < 0) continue;
    const T v = regX[i];
    if (i < nexcl) {
      if (accumulate)
        y[dv[k]] += v;
      else
        y[dv[k]] = v;
    } else {
      atomic_add(y + dv[k], v);
    }
  }
}

// ---------------------------------------------------------------------------
// Launch helpers: one (n, q) -> template switch per (T, TG).

template <typename T, typename TG>
struct ApplyArgs {
  const T* x;
  T* y;
  const int32_t* cells;
  const TG* G;
  const T* tB;
  const T* tD;
  int64_t ncell;
  int accumulate;  // 0 -> y zeroed by the caller, 1 -> y += A x
  // cell map coefficients (kGeoCoef3 / kGeoCoef2 per cell), 1D points and
  // weights; set -> G recomputed in the kernel (the G argument is then unused)
  const T* Xv = nullptr;
  const T* tP = nullptr;
  const T* tW = nullptr;
  // 3D with Xv: shares of x.Ax, one per wave (nblk * NT / 64 values), or none
  double* dotp = nullptr;
  // 3D with Xv: diffusion coefficient per quadrature point ([cell][qp]), or
  // none (kappa = 1)
  const T* kap = nullptr;
};

// Host-built block structure of the apply (see k_apply3d, k_apply2d).
struct BlockArgs {
  const uint16_t* pidx;
  const int32_t* bdofs;
  const int64_t* boff;
  const int32_t* bexcl;
  const int64_t* bcell;      // first cell of each block (nblocks + 1)
  int nblk;
};

template <int N, int Q, typename T, typename TG>
hipError_t launch_apply3d_nq(const ApplyArgs<T, TG>& a, const BlockArgs& k,
                             hipStream_t s) {
  using S = Apply3Shape<N, Q, T>;
  if (k.nblk <= 0) return hipSuccess;
  // x.Ax and kappa per point: fused geometry only
  if ((a.dotp || a.kap) && !a.Xv) return hipErrorInvalidValue;
  if constexpr (sizeof(TG) == sizeof(T)) {  // the fused forms read no G
#define SF_GEO_LAUNCH(DOT_, KAP_)                                                        \
  hipLaunchKernelGGL((k_apply3d<N, Q, T, TG, true, DOT_, KAP_>), dim3(k.nblk), dim3(S::NT), \
                     0, s, a.x, a.y, k.pidx, k.bdofs, k.boff, k.bexcl, k.bcell, k.nblk, \
                     a.G, a.tB, a.tD, a.Xv, a.tP, a.tW, a.accumulate, a.dotp, a.kap)
    if (a.Xv) {
      if (a.dotp && a.kap)
        SF_GEO_LAUNCH(true, true);
      else if (a.dotp)
        SF_GEO_LAUNCH(true, false);
      else if (a.kap)
        SF_GEO_LAUNCH(false, true);
      else
        SF_GEO_LAUNCH(false, false);
      return hipGetLastError();
    }
#undef SF_GEO_LAUNCH
  } else {
    if (a.Xv) return hipErrorInvalidValue;
  }
  hipLaunchKernelGGL((k_apply3d<N, Q, T, TG, false, false>), dim3(k.nblk), dim3(S::NT), 0,
                     s, a.x, a.y, k.pidx, k.bdofs, k.boff, k.bexcl, k.bcell,
                     k.nblk, a.G, a.tB, a.tD, a.Xv, a.tP, a.tW, a.accumulate, nullptr,
                     nullptr);
  return hipGetLastError();
}

template <int N, int Q, typename T, typename TG>
hipError_t launch_apply2d_nq(const ApplyArgs<T, TG>& a, const BlockArgs& k, hipStream_t s) {
  using S = Apply2Shape<N, Q, T>;
  if (k.nblk <= 0) return hipSuccess;
  // x.Ax shares and kappa per point: 3D only (a 2D kappa scales the stored G)
  if (a.dotp || a.kap) return hipErrorInvalidValue;
  if constexpr (sizeof(TG) == sizeof(T)) {
    if (a.Xv) {
      hipLaunchKernelGGL((k_apply2d<N, Q, T, TG, true>), dim3(k.nblk), dim3(S::NT), 0, s,
                         a.x, a.y, k.pidx, k.bdofs, k.boff, k.bexcl, k.bcell, k.nblk, a.G,
                         a.tB, a.tD, a.Xv, a.tP, a.tW, a.accumulate);
      return hipGetLastError();
    }
  } else {
    if (a.Xv) return hipErrorInvalidValue;
  }
  hipLaunchKernelGGL((k_apply2d<N, Q, T, TG, false>), dim3(k.nblk), dim3(S::NT), 0, s,
                     a.x, a.y, k.pidx, k.bdofs, k.boff, k.bexcl, k.bcell, k.nblk, a.G,
                     a.tB, a.tD, a.Xv, a.tP, a.tW, a.accumulate);
  return hipGetLastError();
}

// Returns hipErrorNotSupported for an (n, q) pair outside SF_FOR_EACH_NQ.
// Block-assembled in 2D and 3D (k needed).  The shared dofs are added into
// y: the caller zeroes y for y = A x (with accumulate = 0 the dofs
// exclusive to a block are stored, not added).
template <typename T, typename TG>
hipError_t launch_apply(int dim, int n, int q, const ApplyArgs<T, TG>& a,
                        const BlockArgs* k, hipStream_t s) {
  if (dim == 3) {
    if (!k) return hipErrorInvalidValue;
    switch (n * 16 + q) {
#define SF_CASE(N_, Q_) \
  case N_ * 16 + Q_:    \
    return launch_apply3d_nq<N_, Q_, T, TG>(a, *k, s);
      SF_FOR_EACH_NQ(SF_CASE)
#undef SF_CASE
      default:
        return hipErrorNotSupported;
    }
  } else if (dim == 2) {
    if (!k) return hipErrorInvalidValue;
    switch (n * 16 + q) {
#define SF_CASE(N_, Q_) \
  case N_ * 16 + Q_:    \
    return launch_apply2d_nq<N_, Q_, T, TG>(a, *k, s);
      SF_FOR_EACH_NQ(SF_CASE)
#undef SF_CASE
      default:
        return hipErrorNotSupported;
    }
  }
  return hipErrorNotSupported;
}

// Launch shape of the 3D (or 2D) apply for (n, q): threads and cells per
// workgroup (0, 0 if unsupported).
inline void apply_shape(int dim, int n, int q, int* nt, int* eb) {
  *nt = 0;
  *eb = 0;
  if (dim == 3) {
    switch (n * 16 + q) {
#define SF_CASE(N_, Q_)                         \
  case N_ * 16 + Q_:                            \
    *nt = Apply3Shape<N_, Q_, double>::NT;      \
    *eb = Apply3Shape<N_, Q_, double>::EB;      \
    return;
      SF_FOR_EACH_NQ(SF_CASE)
#undef SF_CASE
      default:
        return;
    }
  } else if (dim == 2) {
    switch (n * 16 + q) {
#define SF_CASE(N_, Q_)                         \
  case N_ * 16 + Q_:                            \
    *nt = Apply2Shape<N_, Q_, double>::NT;      \
    *eb = Apply2Shape<N_, Q_, double>::EB;      \
    return;
      SF_FOR_EACH_NQ(SF_CASE)
#undef SF_CASE
      default:
        return;
    }
  }
}

}  // namespace sf
