// sf_apply_mfma.hip -- 3D apply with the tensor contractions on the matrix
// cores (v_mfma_f32_16x16x16_bf16: bf16 operands, FP32 accumulation).
//
// Same operator, block structure and assembly as k_apply3d (sf_apply.hpp);
// FP32 vectors, the geometric factor in bfloat16.  Every 1D contraction of
// a cell is a small GEMM: table (rows = output index, K = contracted index)
// times the cell tensor viewed as K x (the other two indices).  A 16x16x16
// MFMA takes one 16-column tile of that view; tiles of all cells of the
// workgroup are dealt to its four waves.
//
//   forward   X : B, D   (Q x N)  . x      [a][(b,c)]        ->  2 tensors
//             Y : B, D, B         . (Bx, Dx) [b][(i1,c)]      ->  3 tensors
//             Z : B, B, D         . [c][(i1,i2)] -> gradients in registers
//   pointwise    h = G grad  (the MFMA output tile is the next B operand)
//   backward  Z^T: B^T, B^T, D^T  -> in place over the Y output
//             Y^T: B^T, D^T + B^T -> 2 tensors [i1][(b,c)]
//             X^T: D^T + B^T      -> y, assembled in LDS as in k_apply3d
//
// Operand precision.  SPLIT = 1 rounds every operand to bf16 (8 significant
// bits).  SPLIT = 3 writes each FP32 operand as the sum of three bf16
// terms (hi + mid + lo, each the rounded remainder of the previous) and
// keeps the six partial products of order >= 2^-24, so the contraction is
// as accurate as FP32 arithmetic at six MFMAs per tile.  SPLIT = 2 (hi +
// lo, three products) carries about 16 bits.  SPLIT = 0 keeps the operands
// in FP32 and runs each tile as four v_mfma_f32_16x16x4f32 (no operand
// conversion at all); G is then read in FP32 too.
//
// MFMA 16x16x4 f32 lane layout: A[m][k]: m = l % 16, k = g; B[k][n]:
// n = l % 16, k = g; C as above.  Instruction j of a tile takes value j of
// every lane, i.e. the contraction indices k = 4 g + j -- the same four
// values per lane as one 16x16x16 bf16 instruction, so both forms share
// every load and store below.
//
// MFMA 16x16x16 bf16 lane layout (wave64, lane l, g = l / 16, j = 0..3):
//   A[m][k]: m = l % 16, k = 4 g + j      B[k][n]: n = l % 16, k = 4 g + j
//   C[m][n]: n = l % 16, m = 4 g + j
// so a C tile is, lane for lane, the B operand of a contraction over m.

#include "sf_apply.hpp"
#include "sf_kernels.h"

namespace sf {
namespace {

typedef short s4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf2 __attribute__((ext_vector_type(2)));
typedef unsigned u2 __attribute__((ext_vector_type(2)));

template <int SPLIT>
struct Op {
  s4 t[SPLIT];
};
template <>
struct Op<0> {  // FP32 operands
  f4 v;
};

// Two FP32 values -> one packed bf16 pair (v_cvt_pk_bf16_f32, round to
// nearest).  The empty asm keeps the pair opaque, so the remainders below
// reuse it instead of converting each value again.
__device__ __forceinline__ unsigned pk_bf16(float a, float b) {
  bf2 h;
  h.x = (__bf16)a;
  h.y = (__bf16)b;
  unsigned u = __builtin_bit_cast(unsigned, h);
  asm volatile("" : "+v"(u));
  return u;
}

// FP32 -> SPLIT bf16 terms (round to nearest; each remainder is exact).
template <int SPLIT>
__device__ __forceinline__ Op<SPLIT> split4(f4 v) {
  Op<SPLIT> o;
  if constexpr (SPLIT == 0) {
    o.v = v;
    return o;
  } else {
#pragma unroll
  for (int s = 0; s < SPLIT; ++s) {
    const unsigned h0 = pk_bf16(v.x, v.y), h1 = pk_bf16(v.z, v.w);
    const u2 hh = {h0, h1};
    o.t[s] = __builtin_bit_cast(s4, hh);
    if (s + 1 < SPLIT) {
      v.x -= __uint_as_float(h0 << 16);
      v.y -= __uint_as_float(h0 & 0xffff0000u);
      v.z -= __uint_as_float(h1 << 16);
      v.w -= __uint_as_float(h1 & 0xffff0000u);
    }
  }
  return o;
  }
}

// Branch-free masked accesses.  A read is always issued (an out-of-range
// slot reads element 0 of the same region) and its value replaced by zero,
// so the reads of a tile go out back to back and share one wait; a masked
// write goes to the lane's own dummy word.  (A conditional access compiles
// to a branch per element, each read then waiting for its own result.)
__device__ __forceinline__ float ld_or0(const float* base, int idx, bool ok) {
  const float v = base[ok ? idx : 0];
  return ok ? v : 0.f;
}
__device__ __forceinline__ void st_or_dummy(float* base, int idx, bool ok, float v,
                                            float* dummy) {
  float* p = ok ? base + idx : dummy;
  *p = v;
}

__device__ __forceinline__ f4 mfma(s4 a, s4 b, f4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// c += a . b with split operands, small terms first.
template <int SPLIT>
__device__ __forceinline__ f4 mma(const Op<SPLIT>& a, const Op<SPLIT>& b, f4 c) {
  if constexpr (SPLIT == 0) {
#pragma unroll
    for (int j = 0; j < 4; ++j) c = __builtin_amdgcn_mfma_f32_16x16x4f32(a.v[j], b.v[j], c, 0, 0, 0);
    return c;
  } else {
    if constexpr (SPLIT >= 3) {
      c = mfma(a.t[2], b.t[0], c);
      c = mfma(a.t[1], b.t[1], c);
      c = mfma(a.t[0], b.t[2], c);
    }
    if constexpr (SPLIT >= 2) {
      c = mfma(a.t[1], b.t[0], c);
      c = mfma(a.t[0], b.t[1], c);
    }
    return mfma(a.t[0], b.t[0], c);
  }
}

template <int N, int Q, int SPLIT>
struct MfmaShape {
  static constexpr int EB = Apply3Shape<N, Q, double>::EB;  // block structure
  static constexpr int NN = N * N, NQ = N * Q, QQ = Q * Q;
  static constexpr int NNN = N * N * N, QQQ = Q * Q * Q;
  static constexpr int TX = (NN + 15) / 16, TY = (NQ + 15) / 16, TZ = (QQ + 15) / 16;
  static constexpr int W = 4;  // waves per workgroup (more measured slower)
  static constexpr int NT = 64 * W;
  static constexpr int UX = (EB * TX + W - 1) / W;  // tiles per wave, X and X^T
  static constexpr int UY = (EB * TY + W - 1) / W;
  static constexpr int UZ = (EB * TZ + W - 1) / W;
  static constexpr int S1 = N * Q * N;  // one X-pass (or Y^T-pass) tensor
  static constexpr int S2 = Q * Q * N;  // one Y-pass tensor
  static constexpr int RX = EB * NNN;   // the block's x values, later its y
  static constexpr int LDS_FLOATS = RX + EB * (2 * S1 + 3 * S2) + 64;  // + dummies
  static constexpr int KU = (EB * NNN + NT - 1) / NT;
  // G of every z tile loaded up front (bf16 pairs: 12 registers a tile)
  static constexpr bool GPRE = SPLIT > 0 && UZ <= 4;
  static_assert(N <= 16 && Q <= 16, "one MFMA tile per contraction");
  static_assert(LDS_FLOATS * 4 <= 65536, "apply3d_mfma: LDS above 64 KiB");
};

// G of one Z tile for lane (g, cl): rows i3 = 4g + j, column col; the pair
// p of entries (2p, 2p+1) of point (i3, col) sits at
// cell * 6Q^3 + ((i3 * 3 + p) * Q^2 + col) * 2  (bf16 units).
template <int N, int Q>
__device__ __forceinline__ void load_gz(const uint32_t* __restrict__ Gm,
                                        int64_t cell, int col, int g, bool ok,
                                        uint32_t gv[12]) {
  constexpr int QQ = Q * Q;
  const uint32_t* p = Gm + (ok ? cell * 3 * QQ * Q : 0);
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int i3 = 4 * g + j;
    const bool v = ok && i3 < Q && col < QQ;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const uint32_t w = __builtin_nontemporal_load(p + (v ? (i3 * 3 + k) * QQ + col : 0));
      gv[j * 3 + k] = v ? w : 0u;
    }
  }
}

// FP32 G of one Z tile (SPLIT = 0): the six entries of point (i3, col) at
// cell * 6Q^3 + (i3 * 6 + k) * Q^2 + col.
template <int N, int Q>
__device__ __forceinline__ void load_gz32(const float* __restrict__ Gf, int64_t cell, int col,
                                          int g, bool ok, float gv[24]) {
  constexpr int QQ = Q * Q;
  const float* p = Gf + (ok ? cell * 6 * QQ * Q : 0);
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int i3 = 4 * g + j;
    const bool v = ok && i3 < Q && col < QQ;
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      const float w = __builtin_nontemporal_load(p + (v ? (i3 * 6 + k) * QQ + col : 0));
      gv[j * 6 + k] = v ? w : 0.f;
    }
  }
}

template <int N, int Q, int SPLIT>
__global__ void __launch_bounds__((MfmaShape<N, Q, SPLIT>::NT))
    k_apply3d_mfma(const float* __restrict__ x, float* __restrict__ y,
                   const uint16_t* __restrict__ pidx,
                   const int32_t* __restrict__ bdofs,
                   const int64_t* __restrict__ boff,
                   const int32_t* __restrict__ bexcl,
                   const int64_t* __restrict__ bcell, int nblk,
                   const void* __restrict__ Gv,
                   const float* __restrict__ tB, const float* __restrict__ tD) {
  using M = MfmaShape<N, Q, SPLIT>;
  const uint32_t* __restrict__ Gm = (const uint32_t*)Gv;  // bf16 pairs (SPLIT > 0)
  const float* __restrict__ Gf = (const float*)Gv;        // FP32 (SPLIT = 0)
  constexpr int NN = M::NN, NQ = M::NQ, QQ = M::QQ, NNN = M::NNN;
  constexpr int S1 = M::S1, S2 = M::S2, W = M::W;
  __shared__ float smem[M::LDS_FLOATS];
  float* const bx = smem;
  float* const T1 = smem + M::RX;
  float* const T2 = T1 + M::EB * 2 * S1;
  float* const dummy = smem + (M::LDS_FLOATS - 64) + (threadIdx.x & 63);

  const int b = xcd_remap(blockIdx.x, nblk);
  const int64_t base = bcell[b];
  const int ncb = (int)(bcell[b + 1] - base);
  const int tid = threadIdx.x;
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int lane = tid & 63, g = lane >> 4, cl = lane & 15;
  const int64_t d0 = boff[b];
  const int nu = (int)(boff[b + 1] - d0);
  const int nexcl = bexcl[b];
  const int32_t* __restrict__ dofs = bdofs + d0;

  // ---- prologue: dof ids, x, local indices, G ------------------------------
  int32_t dv[M::KU];
  float xv[M::KU];
#pragma unroll
  for (int k = 0; k < M::KU; ++k) {
    const int i = tid + k * M::NT;
    dv[k] = i < nu ? dofs[i] : -1;
  }
#pragma unroll
  for (int k = 0; k < M::KU; ++k) xv[k] = x[dv[k] >= 0 ? dv[k] : 0];

  // X-tile u = w + W i: cell u / TX, tile u % TX; lane value j is node
  // (a = 4g + j, col = b + N c) -- the same slots in X (B operand) and X^T
  // (C tile), so one set of local indices serves both.
  uint16_t pv[M::UX][4];
#pragma unroll
  for (int i = 0; i < M::UX; ++i) {
    const int u = w + W * i;
    const int e = u / M::TX, col = (u - e * M::TX) * 16 + cl;
    const bool ok = u < M::EB * M::TX && e < ncb && col < NN;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int a = 4 * g + j;
      const bool v = ok && a < N;
      const uint16_t l = pidx[v ? (base + e) * NNN + a + N * col : 0];
      pv[i][j] = v ? l : kNoDof;
    }
  }

  uint32_t gv[M::GPRE ? M::UZ : 1][12];
  if constexpr (M::GPRE) {
#pragma unroll
    for (int i = 0; i < M::UZ; ++i) {
      const int u = w + W * i;
      const int e = u / M::TZ, col = (u - e * M::TZ) * 16 + cl;
      load_gz<N, Q>(Gm, base + e, col, g, u < M::EB * M::TZ && e < ncb, gv[i]);
    }
  }

  // tables as A operands: forward A[m][k] = T[m][k] (m < Q, k < N),
  // backward A[m][k] = T[k][m] (m < N, k < Q); zero outside
  Op<SPLIT> aB, aD, aBt, aDt;
  {
    f4 vb, vd, vbt, vdt;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int k = 4 * g + j;
      const bool f = cl < Q && k < N, r = cl < N && k < Q;
      vb[j] = ld_or0(tB, cl * N + k, f);
      vd[j] = ld_or0(tD, cl * N + k, f);
      vbt[j] = ld_or0(tB, k * N + cl, r);
      vdt[j] = ld_or0(tD, k * N + cl, r);
    }
    aB = split4<SPLIT>(vb);
    aD = split4<SPLIT>(vd);
    aBt = split4<SPLIT>(vbt);
    aDt = split4<SPLIT>(vdt);
  }

#pragma unroll
  for (int k = 0; k < M::KU; ++k) {
    const int i = tid + k * M::NT;
    if (i < nu) bx[i] = xv[k];
  }
  __syncthreads();

  const f4 zero = {0.f, 0.f, 0.f, 0.f};

  // ---- X: [a][(b,c)] -> B x, D x stored [b][(i1 + Q c)] ---------------------
#pragma unroll
  for (int i = 0; i < M::UX; ++i) {
    const int u = w + W * i;
    if (u >= M::EB * M::TX) break;
    const int e = u / M::TX;
    if (e >= ncb) break;
    const int col = (u - e * M::TX) * 16 + cl;
    f4 v;
#pragma unroll
    for (int j = 0; j < 4; ++j) v[j] = ld_or0(bx, pv[i][j], pv[i][j] != kNoDof);
    const Op<SPLIT> o = split4<SPLIT>(v);
    const f4 cb = mma(aB, o, zero), cd = mma(aD, o, zero);
    const int bb = col % N, c = col / N;
    float* t = T1 + e * 2 * S1;
    const int o0 = bb * NQ + Q * c;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int i1 = 4 * g + j;
      const bool ok = col < NN && i1 < Q;
      st_or_dummy(t, o0 + i1, ok, cb[j], dummy);
      st_or_dummy(t, S1 + o0 + i1, ok, cd[j], dummy);
    }
  }
  __syncthreads();

  // the x values are consumed: region bx becomes the block's y accumulator
#pragma unroll
  for (int k = 0; k < M::KU; ++k) {
    const int i = tid + k * M::NT;
    if (i < nu) bx[i] = 0.f;
  }

  // ---- Y: K = b -> (B.Bx, D.Bx, B.Dx) stored [c][(i1 + Q i2)] ---------------
#pragma unroll
  for (int i = 0; i < M::UY; ++i) {
    const int u = w + W * i;
    if (u >= M::EB * M::TY) break;
    const int e = u / M::TY;
    if (e >= ncb) break;
    const int col = (u - e * M::TY) * 16 + cl;  // i1 + Q c
    const float* t = T1 + e * 2 * S1;
    f4 vx, vd;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int bb = 4 * g + j;
      const bool ok = bb < N && col < NQ;
      vx[j] = ld_or0(t, bb * NQ + col, ok);
      vd[j] = ld_or0(t, S1 + bb * NQ + col, ok);
    }
    const Op<SPLIT> ox = split4<SPLIT>(vx), od = split4<SPLIT>(vd);
    const f4 cbb = mma(aB, ox, zero), cbd = mma(aD, ox, zero), cdb = mma(aB, od, zero);
    const int i1 = col % Q, c = col / Q;
    float* o = T2 + e * 3 * S2;
    const int o0 = c * QQ + i1;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int i2 = 4 * g + j;
      const bool ok = col < NQ && i2 < Q;
      st_or_dummy(o, o0 + Q * i2, ok, cdb[j], dummy);           // x-gradient path
      st_or_dummy(o, S2 + o0 + Q * i2, ok, cbd[j], dummy);      // y-gradient path
      st_or_dummy(o, 2 * S2 + o0 + Q * i2, ok, cbb[j], dummy);  // z-gradient path
    }
  }
  __syncthreads();

  // ---- Z, G, Z^T: K = c; every lane rewrites exactly the words it read ------
#pragma unroll
  for (int i = 0; i < M::UZ; ++i) {
    const int u = w + W * i;
    if (u >= M::EB * M::TZ) break;
    const int e = u / M::TZ;
    if (e >= ncb) break;
    const int col = (u - e * M::TZ) * 16 + cl;  // i1 + Q i2
    float* t = T2 + e * 3 * S2;
    f4 v0, v1, v2;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int c = 4 * g + j;
      const bool ok = c < N && col < QQ;
      v0[j] = ld_or0(t, c * QQ + col, ok);
      v1[j] = ld_or0(t, S2 + c * QQ + col, ok);
      v2[j] = ld_or0(t, 2 * S2 + c * QQ + col, ok);
    }
    const f4 gx = mma(aB, split4<SPLIT>(v0), zero);
    const f4 gy = mma(aB, split4<SPLIT>(v1), zero);
    const f4 gz = mma(aD, split4<SPLIT>(v2), zero);
    f4 h0, h1, h2;
    if constexpr (SPLIT == 0) {
      float gf[24];
      load_gz32<N, Q>(Gf, base + e, col, g, true, gf);
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const float* q6 = gf + 6 * j;  // xx xy xz yy yz zz
        h0[j] = q6[0] * gx[j] + q6[1] * gy[j] + q6[2] * gz[j];
        h1[j] = q6[1] * gx[j] + q6[3] * gy[j] + q6[4] * gz[j];
        h2[j] = q6[2] * gx[j] + q6[4] * gy[j] + q6[5] * gz[j];
      }
    } else {
      uint32_t gl[12];
      if constexpr (M::GPRE) {
#pragma unroll
        for (int k = 0; k < 12; ++k) gl[k] = gv[i][k];
      } else {
        load_gz<N, Q>(Gm, base + e, col, g, true, gl);
      }
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        // entries xx, xy | xz, yy | yz, zz (zero outside the Q^3 points)
        const float g0 = __uint_as_float(gl[3 * j] << 16);
        const float g1 = __uint_as_float(gl[3 * j] & 0xffff0000u);
        const float g2 = __uint_as_float(gl[3 * j + 1] << 16);
        const float g3 = __uint_as_float(gl[3 * j + 1] & 0xffff0000u);
        const float g4 = __uint_as_float(gl[3 * j + 2] << 16);
        const float g5 = __uint_as_float(gl[3 * j + 2] & 0xffff0000u);
        h0[j] = g0 * gx[j] + g1 * gy[j] + g2 * gz[j];
        h1[j] = g1 * gx[j] + g3 * gy[j] + g4 * gz[j];
        h2[j] = g2 * gx[j] + g4 * gy[j] + g5 * gz[j];
      }
    }
    const f4 ax = mma(aBt, split4<SPLIT>(h0), zero);
    const f4 ay = mma(aBt, split4<SPLIT>(h1), zero);
    const f4 az = mma(aDt, split4<SPLIT>(h2), zero);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int c = 4 * g + j;
      const bool ok = col < QQ && c < N;
      st_or_dummy(t, c * QQ + col, ok, ax[j], dummy);
      st_or_dummy(t, S2 + c * QQ + col, ok, ay[j], dummy);
      st_or_dummy(t, 2 * S2 + c * QQ + col, ok, az[j], dummy);
    }
  }
  __syncthreads();

  // ---- Y^T: K = i2 -> B^T (x path), D^T (y) + B^T (z), stored [i1][(b,c)] ---
#pragma unroll
  for (int i = 0; i < M::UY; ++i) {
    const int u = w + W * i;
    if (u >= M::EB * M::TY) break;
    const int e = u / M::TY;
    if (e >= ncb) break;
    const int col = (u - e * M::TY) * 16 + cl;  // i1 + Q c
    const int i1 = col % Q, c = col / Q;
    const float* t = T2 + e * 3 * S2;
    const int o0 = c * QQ + i1;
    f4 v0, v1, v2;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int i2 = 4 * g + j;
      const bool ok = i2 < Q && col < NQ;
      v0[j] = ld_or0(t, o0 + Q * i2, ok);
      v1[j] = ld_or0(t, S2 + o0 + Q * i2, ok);
      v2[j] = ld_or0(t, 2 * S2 + o0 + Q * i2, ok);
    }
    const f4 b0 = mma(aBt, split4<SPLIT>(v0), zero);
    f4 b12 = mma(aDt, split4<SPLIT>(v1), zero);
    b12 = mma(aBt, split4<SPLIT>(v2), b12);
    float* o = T1 + e * 2 * S1;
    const int o1 = i1 * NN + N * c;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int bb = 4 * g + j;
      const bool ok = col < NQ && bb < N;
      st_or_dummy(o, o1 + bb, ok, b0[j], dummy);        // then D_x^T
      st_or_dummy(o, S1 + o1 + bb, ok, b12[j], dummy);  // then B_x^T
    }
  }
  __syncthreads();

  // ---- X^T: K = i1 -> y of node (a, b, c), assembled in LDS ----------------
#pragma unroll
  for (int i = 0; i < M::UX; ++i) {
    const int u = w + W * i;
    if (u >= M::EB * M::TX) break;
    const int e = u / M::TX;
    if (e >= ncb) break;
    const int col = (u - e * M::TX) * 16 + cl;  // b + N c
    const float* t = T1 + e * 2 * S1;
    f4 v0, v1;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int i1 = 4 * g + j;
      const bool ok = i1 < Q && col < NN;
      v0[j] = ld_or0(t, i1 * NN + col, ok);
      v1[j] = ld_or0(t, S1 + i1 * NN + col, ok);
    }
    f4 yv = mma(aDt, split4<SPLIT>(v0), zero);
    yv = mma(aBt, split4<SPLIT>(v1), yv);
#pragma unroll
    for (int j = 0; j < 4; ++j)
      lds_atomic_add(pv[i][j] != kNoDof ? bx + pv[i][j] : dummy, yv[j]);
  }
  __syncthreads();

  // ---- write-back: exclusive dofs stored, shared dofs added ----------------
#pragma unroll
  for (int k = 0; k < M::KU; ++k) {
    const int i = tid + k * M::NT;
    if (dv[k] < 0) continue;
    const float v = bx[i];
    if (i < nexcl)
      y[dv[k]] = v;
    else
      atomic_add(y + dv[k], v);
  }
}

template <int N, int Q, int SPLIT>
hipError_t launch_mfma_nq(const float* x, float* y, const void* Gm,
                          const float* tB, const float* tD, const BlockArgs& k,
                          hipStream_t s) {
  if (k.nblk <= 0) return hipSuccess;
  hipLaunchKernelGGL((k_apply3d_mfma<N, Q, SPLIT>), dim3(k.nblk),
                     dim3(MfmaShape<N, Q, SPLIT>::NT), 0, s,
                     x, y, k.pidx, k.bdofs, k.boff, k.bexcl, k.bcell, k.nblk,
                     Gm, tB, tD);
  return hipGetLastError();
}

// G (FP64, block-interleaved) -> FP32 in the layout of load_gz32.
__global__ void k_g3_to_mfma32(const double* __restrict__ G, float* __restrict__ out, int q,
                               int64_t ncell, const int32_t* __restrict__ cell_block,
                               const int64_t* __restrict__ bcell) {
  const int64_t QQ = (int64_t)q * q, QQQ = QQ * q;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ncell * QQQ) return;
  const int64_t cell = t / QQQ;
  const int qp = (int)(t - cell * QQQ);
  const int i3 = qp / (int)QQ, col = qp - i3 * (int)QQ;
  const int blk = cell_block[cell];
  const int64_t b0 = bcell[blk];
  const int ncb = (int)(bcell[blk + 1] - b0);
  float* o = out + cell * 6 * QQQ;
  for (int k = 0; k < 6; ++k)
    o[(i3 * 6 + k) * QQ + col] = (float)G[g3_index(b0, ncb, cell, k, i3, col, q)];
}

// G (FP64, block-interleaved) -> bf16 pairs in the MFMA layout of load_gz.
__global__ void k_g3_to_mfma(const double* __restrict__ G, uint32_t* __restrict__ out,
                             int q, int64_t ncell, const int32_t* __restrict__ cell_block,
                             const int64_t* __restrict__ bcell) {
  const int64_t QQ = (int64_t)q * q, QQQ = QQ * q;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ncell * QQQ) return;
  const int64_t cell = t / QQQ;
  const int qp = (int)(t - cell * QQQ);
  const int i3 = qp / (int)QQ, col = qp - i3 * (int)QQ;
  const int blk = cell_block[cell];
  const int64_t b0 = bcell[blk];
  const int ncb = (int)(bcell[blk + 1] - b0);
  uint32_t* o = out + cell * 3 * QQQ;
  for (int p = 0; p < 3; ++p) {
    const float lo = (float)G[g3_index(b0, ncb, cell, 2 * p, i3, col, q)];
    const float hi = (float)G[g3_index(b0, ncb, cell, 2 * p + 1, i3, col, q)];
    o[(i3 * 3 + p) * QQ + col] =
        (uint32_t)f32_to_bf16_bits(lo) | ((uint32_t)f32_to_bf16_bits(hi) << 16);
  }
}

}  // namespace

hipError_t apply_mfma(int n, int q, int split, const float* x, float* y,
                      const void* Gm, const float* tB, const float* tD,
                      const BlockArgs* k, hipStream_t s) {
  if (!k) return hipErrorInvalidValue;
  if (split < 0 || split > 3) return hipErrorNotSupported;
  switch (n * 16 + q) {
#define SF_CASE(N_, Q_)                                                      \
  case N_ * 16 + Q_:                                                         \
    return split == 0   ? launch_mfma_nq<N_, Q_, 0>(x, y, Gm, tB, tD, *k, s) \
           : split == 1 ? launch_mfma_nq<N_, Q_, 1>(x, y, Gm, tB, tD, *k, s) \
           : split == 2 ? launch_mfma_nq<N_, Q_, 2>(x, y, Gm, tB, tD, *k, s) \
                        : launch_mfma_nq<N_, Q_, 3>(x, y, Gm, tB, tD, *k, s);
    SF_FOR_EACH_NQ(SF_CASE)
#undef SF_CASE
    default:
      return hipErrorNotSupported;
  }
}

hipError_t launch_g3_to_mfma32(const double* G, float* out, int q, int64_t ncell,
                               const int32_t* cell_block, const int64_t* bcell,
                               hipStream_t s) {
  const int64_t n = ncell * (int64_t)q * q * q;
  if (n <= 0) return hipSuccess;
  const int64_t nb = (n + 255) / 256;
  if (nb > 0x7fffffffLL) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_g3_to_mfma32, dim3((unsigned)nb), dim3(256), 0, s, G, out, q, ncell,
                     cell_block, bcell);
  return hipGetLastError();
}

hipError_t launch_g3_to_mfma(const double* G, uint32_t* out, int q, int64_t ncell,
                             const int32_t* cell_block, const int64_t* bcell,
                             hipStream_t s) {
  const int64_t n = ncell * (int64_t)q * q * q;
  if (n <= 0) return hipSuccess;
  const int64_t nb = (n + 255) / 256;
  if (nb > 0x7fffffffLL) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_g3_to_mfma, dim3((unsigned)nb), dim3(256), 0, s, G, out, q,
                     ncell, cell_block, bcell);
  return hipGetLastError();
}

}  // namespace sf
