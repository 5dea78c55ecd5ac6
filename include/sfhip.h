/*
 * sfhip.h -- C ABI of the MI355X (gfx950) matrix-free, sum-factorized
 * high-order finite-element Laplacian  y = A u.
 *
 * Every entry point takes plain C types (pointers, sizes, ints), returns an
 * int status (SF_OK = 0) and never lets a C++ exception cross the boundary.
 * sf_last_error() returns a thread-local message for the last failure.
 *
 * Conventions
 *   - D in {2, 3}; n = p + 1 nodes and q Gauss points per direction
 *     (isotropic, as the reference's GPU path requires:
 *     src/SumFactorizationMaps.jl:25-32 packs one q x n matrix per direction).
 *   - 1D tables are row-major q x n:  B[i*n + j] = phi_j(xi_i),
 *     Dm[i*n + j] = phi_j'(xi_i)  (reference: A0 / B0 in
 *     src/SumFactorizationMaps.jl:54-55).
 *   - Cell-local arrays are tensor-lexicographic with direction 1 (x) fastest,
 *     the same order as the reference's Z tensors and its linear quadrature
 *     index (src/SumFactorizationKernels.jl:174-216).
 *   - Geometric factor per quadrature point: the symmetric
 *     G = w_q |det J| (J^T J)^{-1}, stored as D(D+1)/2 entries
 *     (2D: G00 G01 G11; 3D: G00 G01 G02 G11 G12 G22), layout
 *     [cell][entry][quadrature point].  This is the product the reference
 *     forms on the fly in _sumfac_apply_weights! from wq, jq = J^{-T} and
 *     djq = |det J| (src/SumFactorizationKernels.jl:191-210).
 *   - Device-pointer entry points are stream ordered on `stream`
 *     (a hipStream_t; NULL = the null stream) and return before the work ends.
 *   - A plan is not re-entrant: one host thread per plan at a time.
 */
#ifndef SFHIP_H
#define SFHIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum sf_status {
  SF_OK = 0,
  SF_ERR_ARG = 1,           /* bad dimension / order / size / null pointer   */
  SF_ERR_UNSUPPORTED = 2,   /* (n, q) pair or precision without a kernel     */
  SF_ERR_DOF_RANGE = 3,     /* a cell dof id outside the free-dof range      */
  SF_ERR_JACOBIAN = 4,      /* det J <= 0 (or non-finite) at a quad. point   */
  SF_ERR_HIP = 5,           /* HIP runtime error                             */
  SF_ERR_NOT_READY = 6,     /* plan has no cells or no geometry yet          */
  SF_ERR_NOT_CONVERGED = 7, /* solver hit its iteration limit                */
  SF_ERR_NONFINITE = 8      /* NaN / Inf met in a solver                     */
};

/* Arithmetic of an operator application. */
enum sf_precision {
  SF_PREC_F64 = 0,       /* FP64 everywhere                                       */
  SF_PREC_F32 = 1,       /* FP32 vectors, tables, G and contractions (vector ALU) */
  SF_PREC_F32_GBF16 = 2, /* as F32, but G stored in bfloat16 (half the G bytes)   */
  SF_PREC_MFMA_BF16 = 3, /* 3D: contractions as bf16 MFMA GEMMs (operands rounded
                            to bf16), FP32 accumulation, G in bf16               */
  SF_PREC_MFMA_BF16X3 = 4, /* 3D: as MFMA_BF16 with every operand split into three
                            bf16 terms (six MFMAs per tile): FP32-level accuracy */
  SF_PREC_MFMA_BF16X2 = 5, /* 3D: two bf16 terms per operand (three MFMAs per
                            tile): about 16 significant bits                     */
  SF_PREC_MFMA_F32 = 6,    /* 3D: FP32 operands on the matrix cores (four
                            v_mfma_f32_16x16x4f32 per tile), G in FP32           */
  SF_PREC_MFMA_F16 = 7     /* 3D, n = 8 (Q7), q = 8..10: every contraction an
                            FP16 GEMM over 32 pencils batched across cells
                            (v_mfma_f32_32x32x8f16, FP32 accumulation), G in
                            bf16; about 11 significant bits                      */
};

/* Preconditioner of the device-resident solvers (sf_cg, sf_cg_mixed,
 * sf_cg_ir), selected per plan with sf_plan_set_preconditioner. */
enum sf_preconditioner {
  SF_PC_NONE = 0,   /* unpreconditioned CG, as the reference's cg! call       */
  SF_PC_JACOBI = 1, /* z = diag(A)^{-1} r: the matrix-free diagonal of A(kappa),
                       inverted once in FP32 and used by every solver form     */
  SF_PC_GLL = 2     /* 3D: Jacobi in the Gauss-Lobatto nodal basis of the same
                       space, M^{-1} = T diag(T^T A T)^{-1} T^T with T the
                       tensor-product interpolation from the Gauss-Lobatto to
                       the equispaced nodes.  The solvers run that PCG as
                       Jacobi CG on T^T A T (the same apply on the tables
                       B T1, D T1), so it costs no pass per iteration; only
                       T^T b, T^{-1} x0 and x = T y are formed            */
};

typedef struct sf_plan sf_plan; /* opaque */

typedef struct sf_plan_info {
  int dim, order, n1d, q1d;
  int64_t n_cells, n_free;
  int64_t n_qp;            /* n_cells * q^D                                   */
  int has_cells, has_geometry, has_kappa;
  int threads_per_block;   /* launch shape of the FP64 apply kernel           */
  int cells_per_block;
  int64_t device_bytes;    /* device memory the plan holds                   */
  int n_blocks;            /* 3D: cell blocks (<= cells_per_block cells each) */
  int launches_per_apply;  /* kernel launches of one apply (besides the fill) */
  int cells_reordered;     /* 3D: 1 when the plan keeps its cells in an order
                              built from the connectivity (the caller's order
                              kept fewer neighbours in consecutive cells);
                              per-cell arrays still use the caller's order
                              at the boundary                               */
} sf_plan_info;

typedef struct sf_solve_stats {
  int iterations;        /* CG iterations (IR: total inner iterations)      */
  int outer_iterations;  /* IR: outer (FP64 residual) steps; mixed CG: the
                            residual replacements; CG: how often the true
                            residual, checked when the recurrence met its
                            target, sent it on to a lower one (0 when the
                            first target sufficed)                          */
  int converged;         /* 1 when ||b - A x|| / ||b|| <= rtol               */
  double rel_residual;   /* final FP64 ||b - A x|| / ||b||                   */
  double seconds;        /* wall time of the solve (device synchronised)     */
} sf_solve_stats;

/* ---- diagnostics ------------------------------------------------------- */

/* Message of the last failed call on this thread ("" if none). */
const char* sf_last_error(void);

/* Library version string. */
const char* sf_version(void);

/* ---- 1D tables (host only, no device needed) -------------------------- */

/* Lagrange basis of order p on the 1D nodes in the reference's node order
 * [0, 1, 1/p, 2/p, ..., (p-1)/p] evaluated at the q Gauss-Legendre points of
 * [0, 1] (ascending), plus the points and weights.  Any output may be NULL.
 * Replaces the table construction of SumFactorizationMap(D, fe_orders,
 * quad_orders), src/SumFactorizationMaps.jl:47-65 (A0 :54, B0 :55). */
int sf_tables_1d(int order, int nquad, double* B, double* Dm, double* points,
                 double* weights);

/* Number of Gauss points Gridap uses for a quadrature degree:
 * ceil((degree + 1) / 2) (Quadrature(SEGMENT, degree), used at
 * src/SumFactorizationMaps.jl:51). */
int sf_gauss_points_for_degree(int degree);

/* Analytic FLOP model of one apply, restating
 * count_manual_flops_poisson_matrix_free(ncells, D, SQ, SB;
 * count_only_fma_ops, structured_mesh_opt), src/Utils.jl:33-72. */
int64_t sf_flops_model(int64_t n_cells, int dim, int nquad, int n1d,
                       int count_only_fma_ops, int structured_mesh_opt);

/* ---- plan life cycle ---------------------------------------------------- */

/* Create a plan on the current HIP device.  Replaces the setup of
 * LazyMatrix(m, U, V, dOmega), src/LazyMatrices.jl:30-50, together with the
 * descriptor SumFactorizationMap{D,SB,SQ}, src/SumFactorizationMaps.jl:2-10.
 * Default tables (sf_tables_1d) are installed; flags is reserved (0). */
int sf_plan_create(sf_plan** out, int dim, int order, int nquad,
                   int64_t n_cells, int64_t n_free, int flags);

/* Release every device buffer of the plan.  NULL is accepted. */
int sf_plan_destroy(sf_plan* plan);

int sf_plan_get_info(const sf_plan* plan, sf_plan_info* info);

/* Kernel timing of the applies: with enable = 1 the plan records HIP
 * events on the launch stream around every apply kernel (FP64 and reduced
 * precision, in sf_apply, sf_apply_add and the solvers; not the fill of y),
 * restarting the record; 0 stops recording.  sf_plan_read_timing waits for
 * the recorded events, writes up to max per-launch durations in ms (oldest
 * first), sets *count to the number recorded and clears the record.  A
 * measurement aid that changes no result; the reference times its kernels
 * with CUDA.@profile (test/GPU_v3.jl:96-100).  The library also marks the
 * apply, the solves, their chunks of iterations and the preconditioner
 * set-up as ranges for the profiler's marker trace. */
int sf_plan_set_timing(sf_plan* plan, int enable);
int sf_plan_read_timing(sf_plan* plan, double* ms, int max, int* count);

/* Replace the 1D tables (host arrays, q x n row-major) -- e.g. Gridap's own
 * A0 / B0 so the device uses bit-identical tables.
 * Reference: m.mats = [evalMats, gradMats], src/SumFactorizationMaps.jl:57-58. */
int sf_plan_set_tables(sf_plan* plan, const double* B, const double* Dm,
                       const double* weights);

/* Cell-to-dof map, host array of n_cells * n^D int32.
 * 3D: the apply works on blocks of consecutive cells that share dofs.  When
 * the caller's cell order yields more blocks than an order built from the
 * face connectivity (rows of face neighbours), the plan keeps its cells in
 * the latter; every per-cell array at the boundary (vertices, G, kappa,
 * dJ/dkappa) stays in the caller's order.
 * index_base = 1: Gridap convention (free ids 1..n_free, ids <= 0 are
 *                 Dirichlet), i.e. get_cell_dof_ids(U).data as used by
 *                 src/LazyMatrices.jl:47 and test/Demo_StepByStep.jl:57-69;
 * index_base = 0: free ids 0..n_free-1, negative ids are Dirichlet.
 * local_to_tensor (may be NULL): n^D entries, local dof k of every cell goes
 * to tensor-lexicographic slot local_to_tensor[k] (0-based) -- the
 * reference's dof_map (_get_dof_map, src/SumFactorizationMaps.jl:40-45)
 * linearised.  NULL means the ids are already tensor-lexicographic.
 * Ids are range-checked; SF_ERR_DOF_RANGE on failure. */
int sf_plan_set_cells(sf_plan* plan, const int32_t* cell_dofs, int index_base,
                      const int32_t* local_to_tensor);

/* Geometry from cell vertices: host array [cell][2^D vertices][D coords],
 * (3D: after sf_plan_set_cells; setting new cells drops the geometry),
 * vertices tensor-lexicographic (x fastest), trilinear (bilinear) map.
 * Computes G on the device in FP64 once (J, det J, (J^T J)^{-1}).
 * Replaces the lazy cell_jq / cell_djq of src/LazyMatrices.jl:37-41.
 * SF_ERR_JACOBIAN if det J <= 0 anywhere. */
int sf_plan_set_vertices(sf_plan* plan, const double* vertices);

/* Geometry given directly: host array [cell][D(D+1)/2][q^D] (see header). */
int sf_plan_set_geometry(sf_plan* plan, const double* G);

/* Copy the plan's FP64 geometric factor back to a host array of the layout
 * above (for checking). */
int sf_plan_get_geometry(const sf_plan* plan, double* G);

/* Variable diffusion coefficient kappa at quadrature points, host array
 * [cell][q^D] (positive).  Afterwards every apply uses kappa * G (3D plans
 * built from vertices read kappa per point and recompute G; the others,
 * and the reduced-precision G forms, store kappa * G).
 * No reference counterpart (the reference's operator has kappa = 1). */
int sf_plan_set_kappa(sf_plan* plan, const double* kappa);

/* Build the reduced-precision operator data (FP32 / bf16 G, FP32 tables)
 * ahead of time; otherwise done on the first call that needs it. */
int sf_plan_prepare(sf_plan* plan, int precision);

/* ---- the hot path ------------------------------------------------------- */

/* y = A x on free dofs, FP64, device pointers, y overwritten (not
 * accumulated).  Replaces LinearAlgebra.mul!(y, A::LazyMatrix, x),
 * src/LazyMatrices.jl:52-61, whose per-cell work is
 * Arrays.evaluate!(cache, m::SumFactorizationMap, x, wq, jq, djq),
 * src/SumFactorizationMaps.jl:78-89 (gather, dof->quad contractions,
 * weights + Jacobian, quad->dof contractions, scatter). */
int sf_apply(sf_plan* plan, double* y, const double* x, void* stream);

/* y += A x (FP64, device pointers): the accumulate convention of the
 * reference's GPU kernels, whose scatter adds into a y the caller zeroed
 * (gpu_mul_v0! .. gpu_mul_v4!, src/GPUKernels.jl:77-89; launched as
 * kernel(args...; threads, blocks, shmem), test/GPU_v3.jl:77-84).  One
 * kernel launch, no fill. */
int sf_apply_add(sf_plan* plan, double* y, const double* x, void* stream);

/* Same, with host pointers: stages x / y through device buffers and returns
 * when y is complete.  The host-vector form of the reference's mul!. */
int sf_apply_host(sf_plan* plan, double* y, const double* x);

/* y = A x with FP32 vectors (device pointers).  precision is SF_PREC_F32,
 * SF_PREC_F32_GBF16, SF_PREC_MFMA_BF16, SF_PREC_MFMA_BF16X2,
 * SF_PREC_MFMA_BF16X3, SF_PREC_MFMA_F32 or SF_PREC_MFMA_F16 (the MFMA forms are 3D only:
 * SF_ERR_UNSUPPORTED in 2D).  No reference counterpart:
 * the reduced-precision inner operator of the mixed-precision solver. */
int sf_apply_f32(sf_plan* plan, float* y, const float* x, int precision,
                 void* stream);

/* ---- differentiation ---------------------------------------------------- */

/* Pullback of y = A(kappa) u for a cotangent lam (device pointers, FP64):
 *   dJdkappa[cell][qp] = grad(lam)_q . (G_q grad(u)_q)      (kappa-free G),
 *   dJdu               = A(kappa)^T lam = A(kappa) lam       (overwritten).
 * Gradients at quadrature points are recomputed, never stored; 3D plans
 * built from vertices also recompute the geometric factor (one
 * block-assembled kernel).  Either output may be NULL.  No reference
 * counterpart (new kernel). */
int sf_adjoint(sf_plan* plan, const double* u, const double* lam,
               double* dJdkappa, double* dJdu, void* stream);

/* ---- solvers (device pointers, FP64 b and x) --------------------------- */

/* diag = the diagonal of the assembled operator A(kappa) on the free dofs
 * (FP64, device pointer, overwritten, stream ordered), formed matrix-free
 * from the sum-factorised element-matrix diagonals.  The diag(A) a caller
 * of IterativeSolvers.cg!(x, A, b; Pl = ...) would take from the assembled
 * matrix (profiling/drivers.jl:7-26); the reference's lazy operator has no
 * counterpart. */
int sf_diagonal(sf_plan* plan, double* diag, void* stream);

/* Preconditioner of sf_cg, sf_cg_mixed and sf_cg_ir on this plan
 * (SF_PC_NONE, SF_PC_JACOBI, or SF_PC_GLL on 3D plans): the Pl keyword of
 * IterativeSolvers.cg!, which the reference's compute_solution leaves at
 * the identity (profiling/drivers.jl:85-95).  Its data (the inverse
 * diagonal; for GLL also the basis tables) is built at the next solve or
 * sf_precond_apply and again after any change of the operator (tables,
 * cells, geometry, kappa). */
int sf_plan_set_preconditioner(sf_plan* plan, int kind);

/* z = M^{-1} r for the plan's preconditioner (FP64 device vectors, z
 * overwritten, stream ordered): r itself without one, diag(A)^{-1} r for
 * Jacobi, T diag(T^T A T)^{-1} T^T r for SF_PC_GLL.  The ldiv!(z, Pl, r)
 * that IterativeSolvers.cg! applies to its Pl in every iteration. */
int sf_precond_apply(sf_plan* plan, double* z, const double* r, void* stream);

/* CG on the FP64 operator, x used as the initial guess, stop when the FP64
 * ||b - A x|| <= rtol ||b||.  Unpreconditioned unless the plan has a
 * preconditioner (sf_plan_set_preconditioner).  Replaces
 * IterativeSolvers.cg!(x, A, b) as called by compute_solution,
 * profiling/drivers.jl:85-95 (with Pl: cg!(x, A, b; Pl = ...)). */
int sf_cg(sf_plan* plan, const double* b, double* x, double rtol, int maxit,
          sf_solve_stats* stats, void* stream);

/* Size of the device array behind sf_cg_noalloc, per GPU (128 MiB). */
#define SF_NOALLOC_SCRATCH_BYTES (128u << 20)

/* sf_cg without any device allocation: no counterpart in the reference.
 * For callers that cannot allocate device memory (serving loops, code
 * under stream capture, allocation-free harnesses).  The CG workspace
 * (4 vectors of n_free doubles, 6 with SF_PC_GLL; reduction shares; the
 * scalars) lives in one fixed device array of SF_NOALLOC_SCRATCH_BYTES per
 * GPU, which the library holds until the process ends: about 4M free dofs
 * unpreconditioned or with Jacobi, 2.7M with SF_PC_GLL.  A plan whose
 * workspace does not fit is refused (SF_ERR_ARG); so is a plan whose
 * preconditioner is selected but not built yet (SF_ERR_NOT_READY: one
 * sf_cg or sf_precond_apply call builds it).  Same arguments, results and
 * errors as sf_cg otherwise.  Every call on a GPU shares that GPU's array
 * and keeps nothing from the previous one: calls that overlap on one GPU
 * (two host threads, two streams) race. */
int sf_cg_noalloc(sf_plan* plan, const double* b, double* x, double rtol, int maxit,
                  sf_solve_stats* stats, void* stream);

/* Mixed-precision iterative refinement: FP64 outer residual r = b - A x,
 * FP64 norms and dot products, inner CG with the reduced-precision operator
 * `inner_precision` (FP32 vectors) to relative tolerance inner_rtol or
 * inner_maxit iterations, FP64 correction x += d.  Stops when the FP64
 * ||b - A x|| <= rtol ||b|| or after max_outer steps.  x is the initial
 * guess; the plan's preconditioner applies to the inner CG.  No reference
 * counterpart. */
int sf_cg_ir(sf_plan* plan, const double* b, double* x, double rtol,
             int max_outer, int inner_precision, double inner_rtol,
             int inner_maxit, sf_solve_stats* stats, void* stream);

/* Mixed-precision CG with reliable updates: one CG recurrence on the
 * reduced-precision operator `inner_precision`.  fp64_recurrence = 0: FP32
 * vectors (dot products accumulated in FP64), x (FP64) absorbing the FP32
 * partial solution at each update; fp64_recurrence = 1: FP64 vectors and
 * updates, only the operator in reduced precision (an FP32 copy of the
 * direction goes into it); fp64_recurrence = 2: as 1, but the search
 * direction is held only in FP32 (3D plans with the fused-geometry FP32
 * operator; elsewhere it runs as 1).  Whenever the recurred residual has fallen by
 * the factor delta (0 < delta < 1) since the last update, the residual is
 * replaced by the FP64 b - A x; the search direction is kept, so no Krylov
 * information is lost (unlike sf_cg_ir's restarts).  Stops when the FP64
 * ||b - A x|| <= rtol ||b|| or after maxit iterations;
 * stats->outer_iterations counts the replacements.  x is the initial guess;
 * the plan's preconditioner applies.  No reference counterpart. */
int sf_cg_mixed(sf_plan* plan, const double* b, double* x, double rtol,
                int maxit, int inner_precision, double delta,
                int fp64_recurrence, sf_solve_stats* stats, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* SFHIP_H */
