// sf_kernels.h -- internal launch entry points shared by the translation
// units of libsfhip (not part of the C ABI).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sf {

struct bf16s;
struct BlockArgs;  // host-built block structure of the 3D apply (sf_apply.hpp)
struct CGScalars;  // device scalars of the CG solvers (below)

// ---- geometry (sf_geom.hip) ----
// 3D writes G in the block-interleaved layout (cell_block, bcell); 2D in
// the caller's layout [cell][entry][qp].
hipError_t launch_geometry(int dim, const double* X, const double* pts,
                           const double* w1d, int q, int64_t ncell,
                           const int32_t* cell_block, const int64_t* bcell,
                           double* G, int* bad, hipStream_t s);
hipError_t launch_g3_relayout(const double* src, double* dst, int q,
                              int64_t ncell, const int32_t* cell_block,
                              const int64_t* bcell, int to_int, hipStream_t s);
hipError_t launch_scale_kappa3(const double* G, const double* kappa, int q,
                               int64_t ncell, const int32_t* cell_block,
                               const int64_t* bcell, double* Gk, hipStream_t s);
hipError_t launch_scale_kappa(const double* G, const double* kappa, int ng,
                              int64_t nqc, int64_t ncell, double* Gk,
                              hipStream_t s);
hipError_t launch_check_positive(const double* v, int64_t n, int* bad,
                                 hipStream_t s);
hipError_t launch_to_f32(const double* a, int64_t n, float* b, hipStream_t s);
// 3D reduced-precision G in the pair-interleaved layout (g3_pair_index)
hipError_t launch_g3_to_f32_pairs(const double* G, float* out, int q,
                                  int64_t ncell, const int32_t* cell_block,
                                  const int64_t* bcell, hipStream_t s);
hipError_t launch_g3_to_bf16_pairs(const double* G, bf16s* out, int q,
                                   int64_t ncell, const int32_t* cell_block,
                                   const int64_t* bcell, hipStream_t s);
hipError_t launch_to_bf16(const double* a, int64_t n, bf16s* b, hipStream_t s);

// ---- applies (sf_apply_*.hip) ----
// 3D (k != nullptr): y = A x, or y += A x when accumulate; no fill needed.
// 2D (k == nullptr): y += A x by atomics (the caller zeroes y for y = A x).
// 3D with Xv (cell vertices, 24 per cell) and the 1D points / weights tP,
// tW: the geometric factor is recomputed in the kernel instead of read.
// 3D with dotp (and Xv): shares of x.Ax, one per wave, into
// dotp[0 .. k->nblk * NT / 64) (NT from apply_shape_f64).
// 3D with kap (and Xv): kappa per quadrature point, [cell][qp], scales the
// weights (G is then unused).
hipError_t apply_f64(int dim, int n, int q, const double* x, double* y,
                     const int32_t* cells, const double* G, const double* tB,
                     const double* tD, int64_t ncell, const BlockArgs* k,
                     int accumulate, hipStream_t s, const double* Xv = nullptr,
                     const double* tP = nullptr, const double* tW = nullptr,
                     double* dotp = nullptr, const double* kap = nullptr);
hipError_t apply_f32(int dim, int n, int q, const float* x, float* y,
                     const int32_t* cells, const float* G, const float* tB,
                     const float* tD, int64_t ncell, const BlockArgs* k,
                     int accumulate, hipStream_t s, const float* Xv = nullptr,
                     const float* tP = nullptr, const float* tW = nullptr,
                     double* dotp = nullptr, const float* kap = nullptr);
hipError_t apply_f32_gbf16(int dim, int n, int q, const float* x, float* y,
                           const int32_t* cells, const bf16s* G,
                           const float* tB, const float* tD, int64_t ncell,
                           const BlockArgs* k, int accumulate, hipStream_t s);
void apply_shape_f64(int dim, int n, int q, int* nt, int* eb);
// 3D, matrix cores: y = A x over the block structure k (y zeroed by the
// caller); split = 1 / 2 / 3 (bf16 operands split in that many terms, G as
// bf16 pairs) or 0 (FP32 operands and G, v_mfma_f32_16x16x4f32).
hipError_t apply_mfma(int n, int q, int split, const float* x, float* y,
                      const void* Gm, const float* tB, const float* tD,
                      const BlockArgs* k, hipStream_t s);
// G (FP64, block-interleaved) -> FP32 in the layout of the split = 0 form
hipError_t launch_g3_to_mfma32(const double* G, float* out, int q, int64_t ncell,
                               const int32_t* cell_block, const int64_t* bcell,
                               hipStream_t s);
// G (FP64, block-interleaved) -> the bf16-pair layout of apply_mfma
hipError_t launch_g3_to_mfma(const double* G, uint32_t* out, int q,
                             int64_t ncell, const int32_t* cell_block,
                             const int64_t* bcell, hipStream_t s);

// Scales of the FP16 conversions of apply_gemm (powers of two; each
// brings the bound of the converted value below 1, per table).
struct GemmScales {
  float pB, pD;      // forward contraction outputs (1 / row-sum bound)
  float tB, tD;      // transposed contraction outputs (1 / column-sum bound)
  float cx, cy, cz;  // gradient components to the flux's scale
  float hB, hD;      // the two terms of the transposed y pass to one scale
  float bx;          // the x pass's B path to the scale of its D path
  float vD;          // that scale back to the units of x
};
// 3D at n = 8, q in [8, 10]: the contractions as FP16 GEMMs on the matrix
// cores over pencils batched across cells (v_mfma_f32_32x32x8f16), G as
// the bf16 pairs of apply_mfma; y zeroed by the caller.
bool gemm_supported(int n, int q);
hipError_t apply_gemm(int n, int q, const float* x, float* y, const uint32_t* Gm,
                      const float* tB, const float* tD, const BlockArgs* k,
                      const GemmScales& sc, hipStream_t s);
// *out = bit pattern of max |v| over n bf16 pairs (device, stream ordered)
hipError_t launch_absmax_bf16pairs(const uint32_t* a, int64_t n, unsigned* out, hipStream_t s);

// ---- adjoint (sf_adjoint.hip) ----
// dJdk[cell][qp] = grad(lam) . G grad(u);  ybar += A(kappa) lam
// 3D reads G through the block-interleaved layout (cell_block, bcell).
// corig (3D, may be null): the caller's index of each plan cell, the row
// dJdk is written to.
hipError_t adjoint_f64(int dim, int n, int q, const double* u,
                       const double* lam, double* dJdk, double* ybar,
                       const int32_t* cells, const double* G,
                       const double* kappa, const double* tB,
                       const double* tD, int64_t ncell,
                       const int32_t* cell_block, const int64_t* bcell,
                       const int32_t* corig, hipStream_t s);

// 3D, block-assembled, geometry recomputed from the cell vertices Xv (24 per
// cell) with the 1D points / weights tP, tW: dJdk[cell][qp] (required) and,
// if ybar, the dofs of A(kappa) lam (exclusive dofs stored, shared ones
// added: the caller zeroes ybar).  kappa may be null (kappa = 1).
hipError_t adjoint3d_fused_f64(int n, int q, const double* u, const double* lam,
                               double* dJdk, double* ybar, const BlockArgs* k,
                               const double* tB, const double* tD, const double* Xv,
                               const double* tP, const double* tW, const double* kappa,
                               const int32_t* corig, hipStream_t s);

// ---- diagonal of the operator (sf_diag.hip) ----
// diag += the element-matrix diagonals of every cell (the caller zeroes
// diag): 3D reads G in the block-interleaved layout (cell_block, bcell),
// 2D in the caller's [cell][entry][qp]; cells are the plan's 0-based
// tensor-ordered ids (negative = Dirichlet).
hipError_t launch_diagonal(int dim, int n, int q, const double* G, const double* tB,
                           const double* tD, const int32_t* cells,
                           const int32_t* cell_block, const int64_t* bcell,
                           int64_t ncell, double* diag, hipStream_t s);
// inv = 1 / d (FP32); *bad |= 1 for a non-positive or non-finite d
hipError_t launch_inv_diag(const double* d, float* inv, int64_t n, int* bad, hipStream_t s);

// ---- Gauss-Lobatto-basis preconditioner (sf_precond.hip, 3D) ----
// minv = 1 / (number of cells holding each free dof); minv zeroed by the
// caller; cells: n_cells * n^3 plan ids
hipError_t launch_mult_inverse(const int32_t* cells, int64_t ncell_dofs, float* minv,
                               int64_t nfree, hipStream_t s);
// z = T (dginv . T^T (minv . r)) over the block structure k (eb cells per
// block at most, max_nu distinct dofs); w: scratch vector of nfree values;
// T1 / T1t: the n x n GLL-to-equispaced table and its transpose; sc (may
// be null): nothing runs once sc->done is set
template <typename TR, typename TZ>
hipError_t launch_pc_gll(int n, const TR* r, TZ* z, TZ* w, int64_t nfree, const float* minv,
                         const float* dginv, const double* T1, const double* T1t,
                         const BlockArgs& k, int eb, int max_nu, const CGScalars* sc,
                         hipStream_t s);
// One pass: trans: out = T^T (scale . in) (out zeroed first, shared dofs
// added); else out = M (scale . in) with M = T1 or T1^{-1} (unit vertex
// rows).  scale may be null (1).  FP64.
hipError_t launch_gll_pass(int n, bool trans, const double* in, const float* scale, double* out,
                           const double* M, int64_t nfree, const BlockArgs& k, int eb,
                           int max_nu, hipStream_t s);
// z = dinv . r (the Jacobi preconditioner as a stand-alone product)
hipError_t launch_scale_vec(const double* r, const float* dinv, double* z, int64_t n,
                            hipStream_t s);

// ---- BLAS-1 for the solvers (sf_blas.hip) ----
struct CGScalars {
  double rr;       // current r.z (r.r without a preconditioner)
  double pq;       // p.Ap
  double alpha;
  double beta;
  double rr_new;
  double stop_rr;  // stop when res2 <= stop_rr
  double res2;     // current r.r (the stop test and progress reports)
  int done;        // 1 once rr <= stop_rr (later updates are no-ops)
  int iters;       // iterations taken while not done
  int nonfinite;   // a NaN / Inf was met
  int xpend;       // fused step: x += alpha p still to apply (set by the
                   // residual update, cleared by the next alpha)
};

struct Reducer {
  double* part;        // kPartSlots partial sums
  unsigned* counter;   // zero at rest, reset by the last block
};

constexpr int kRedBlocks = 1024;     // grid of the single-pass reductions
constexpr int kStreamBlocks = 4096;  // grid of the streaming kernels
constexpr int kPartSlots = kStreamBlocks;
constexpr int kRedThreads = 256;

template <typename V>
hipError_t launch_fill(V* a, int64_t n, V v, hipStream_t s);

// y[idx[i]] = 0 for i < n
template <typename V>
hipError_t launch_zero_list(V* y, const int32_t* idx, int64_t n, hipStream_t s);

// out = a . b  (FP64 accumulation); result in *out (device)
// *out = ||b - t||^2
hipError_t launch_diff_norm2(const double* b, const double* t, int64_t n, Reducer red,
                            double* out, hipStream_t s);
template <typename V>
hipError_t launch_dot(const V* a, const V* b, int64_t n, Reducer red,
                      double* out, hipStream_t s);
// out = a . (w a), w FP32 (r.z of the Jacobi preconditioner)
template <typename V>
hipError_t launch_wdot(const V* a, const float* w, int64_t n, Reducer red,
                       double* out, hipStream_t s);

// CG pieces on device scalars sc.  dinv (may be null): the Jacobi
// preconditioner, z = dinv r (FP32 inverse diagonal); null: z = r.
//   cg_start : r = b - Ax, p = z, q = 0, res2 = r.r, rr = r.z,
//              done = res2 <= stop_rr
//   cg_alpha : pq = p.q, alpha = rr / pq
//   cg_update: x += alpha p, r -= alpha q, res2 = r.r, rr_new = r.z,
//              beta = rr_new / rr, rr = rr_new, iters++, done = res2 <= stop_rr
//   cg_dir   : p = z + beta p, q = 0 (ready for the next q += A p)
// Once done is set every later piece returns at once.  The reductions use
// 2 * kStreamBlocks slots of red.part.
template <typename V>
hipError_t launch_cg_start(const V* b, const V* Ax, V* r, V* p, V* q,
                           int64_t n, Reducer red, CGScalars* sc,
                           double stop_rr, hipStream_t s, const float* dinv = nullptr);
template <typename V>
hipError_t launch_cg_alpha(const V* p, const V* q, int64_t n, Reducer red,
                           CGScalars* sc, hipStream_t s);
template <typename V>
hipError_t launch_cg_update(V* x, V* r, const V* p, const V* q, int64_t n,
                            Reducer red, CGScalars* sc, hipStream_t s,
                            const float* dinv = nullptr);
template <typename V>
hipError_t launch_cg_dir(V* p, const V* r, V* q, int64_t n, CGScalars* sc,
                         hipStream_t s, const float* dinv = nullptr);

// IR helpers (FP64 outer, FP32 inner)
// r = b - Ax (FP64)
hipError_t launch_residual(const double* b, const double* Ax, double* r,
                           int64_t n, hipStream_t s);
// r32 = r * inv_scale
hipError_t launch_scale_to_f32(const double* r, double inv_scale, float* r32,
                               int64_t n, hipStream_t s);
// Fused CG step (3D; the apply writes per-wave shares of p.Ap):
//   alpha_parts: alpha = rr / sum(part[0 .. np)), clears xpend
//   resid      : r -= alpha q, res2 = r.r, rr_new = r.z, beta, done; sets xpend
//   xdir       : if xpend: x += alpha p; unless done: p = z + beta p, q = 0
hipError_t launch_cg_alpha_parts(const double* part, int64_t np, Reducer red,
                                 CGScalars* sc, hipStream_t s);
template <typename V>
hipError_t launch_cg_resid(V* r, const V* q, int64_t n, Reducer red, CGScalars* sc,
                           hipStream_t s, const float* dinv = nullptr);
template <typename V>
hipError_t launch_cg_xdir(V* x, V* p, const V* r, V* q, int64_t n, CGScalars* sc,
                          hipStream_t s, const float* dinv = nullptr);
// ... with FP64 vectors and an FP32 operator (q32 = A32 p32; xdir also
// refreshes p32 = (float)p)
hipError_t launch_cg_resid_mixed(double* r, const float* q, int64_t n, Reducer red,
                                 CGScalars* sc, hipStream_t s, const float* dinv = nullptr);
hipError_t launch_cg_xdir_mixed(double* x, double* p, float* p32, const double* r,
                                float* q32, int64_t n, CGScalars* sc, hipStream_t s,
                                const float* dinv = nullptr);
// ... with the search direction only in FP32: x += alpha p32,
// p32 = z + beta p32, q32 = 0; and the host-coefficient form p32 = z + c p32
hipError_t launch_cg_xdir_p32(double* x, float* p32, const double* r, float* q32,
                              int64_t n, const CGScalars* sc, hipStream_t s,
                              const float* dinv = nullptr);
hipError_t launch_dir_coef_p32(float* p32, const double* r, float* q32, double c,
                               int64_t n, hipStream_t s, const float* dinv = nullptr);
// FP64 CG vectors with an FP32 operator (q = A32 p32):
//   alpha_mixed : pq = p . q32, alpha = rr / pq
//   update_mixed: x += alpha p, r -= alpha q32, res2, rr_new = r.z, beta, done
//   dir_mixed   : p = z + beta p, p32 = (float)p, q32 = 0
//   dir_coef_mixed: the same with the coefficient c given by the host
hipError_t launch_cg_alpha_mixed(const double* p, const float* q, int64_t n,
                                 Reducer red, CGScalars* sc, hipStream_t s);
hipError_t launch_cg_update_mixed(double* x, double* r, const double* p, const float* q,
                                  int64_t n, Reducer red, CGScalars* sc, hipStream_t s,
                                  const float* dinv = nullptr);
hipError_t launch_cg_dir_mixed(double* p, float* p32, const double* r, float* q32,
                               int64_t n, CGScalars* sc, hipStream_t s,
                               const float* dinv = nullptr);
hipError_t launch_dir_coef_mixed(double* p, float* p32, const double* r, float* q32,
                                 double c, int64_t n, hipStream_t s,
                                 const float* dinv = nullptr);
// p = z + c p, q = 0 (FP32)
hipError_t launch_dir_coef(float* p, const float* r, float* q, float c,
                           int64_t n, hipStream_t s, const float* dinv = nullptr);
// x += scale * d32
hipError_t launch_axpy_f32_to_f64(double* x, const float* d32, double scale,
                                  int64_t n, hipStream_t s);

}  // namespace sf
