/*
 * sumfac_oracle.c -- TEST INFRASTRUCTURE ONLY.  A CPU restatement of the
 * reference's matrix-free Laplacian apply, used as the checker in tests/,
 * by __graft_entry__.smoke() and as the timed CPU baseline ("port") in
 * bench.py.  Nothing in the product library links or calls this code.
 *
 * It follows the reference's own algorithm step by step:
 *   - gather with Dirichlet dofs read as 0 (scatter_free_and_dirichlet_values
 *     with dir_zeros, src/LazyMatrices.jl:53; test/Demo_StepByStep.jl:57-69);
 *   - _sumfac_array2tensor!: one copy of the cell dofs per gradient
 *     component (src/SumFactorizationKernels.jl:135-151);
 *   - _sumfac_dof2quad! matrix form: for d = 1..D and component k,
 *     Z[d+1][k] = mats[d,k] contracted along direction d of Z[d][k], with
 *     mats[d,k] = (d == k) ? B0 : A0 (src/SumFactorizationKernels.jl:40-65,
 *     src/SumFactorizationMaps.jl:58);
 *   - _sumfac_apply_weights!: per quadrature point
 *     c_r = sum_k (sum_l jq[l,r] jq[l,k]) Z[k],  Z[r] = c_r * wq * djq
 *     (src/SumFactorizationKernels.jl:174-216), with jq = inv(grad(cell_map))
 *     and djq = meas(grad(cell_map)) recomputed per cell from the cell's
 *     vertices, as Gridap's lazy arrays do (src/LazyMatrices.jl:37-41);
 *     grad(cell_map)[i][j] = d x_j / d xi_i (Gridap's gradient layout);
 *     an optional kappa multiplies the quadrature-point factor;
 *   - _sumfac_quad2dof! matrix form, d = D..1 (src/SumFactorizationKernels.jl:
 *     105-130);
 *   - _sumfac_tensor2array!: y_c = sum over components (:156-172);
 *   - assemble_vector!: y zeroed, then y[id] += y_c for free ids
 *     (src/LazyMatrices.jl:57-58; test/Demo_StepByStep.jl:141-145).
 * Contractions sum in the reference's order: 0.0 + m[i,1] z[1] + m[i,2] z[2]..
 *
 * Cell dofs are tensor-lexicographic (x fastest), 0-based, negative =
 * Dirichlet; vertices are [cell][2^D][D], tensor-lexicographic.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* out = M (x) along axis `ax` of a tensor with dims[3] (x fastest).
 * transpose = 0: M is rows x cols, in has dims[ax] == cols, out dims[ax] = rows
 * transpose = 1: uses M^T, in has dims[ax] == rows, out dims[ax] = cols. */
static void contract(const double* in, double* out, const int dims[3], int ax,
                     const double* M, int rows, int cols, int transpose) {
  int din[3] = {dims[0], dims[1], dims[2]};
  int dout[3] = {dims[0], dims[1], dims[2]};
  const int nin = transpose ? rows : cols;
  const int nout = transpose ? cols : rows;
  din[ax] = nin;
  dout[ax] = nout;
  for (int c = 0; c < dout[2]; ++c)
    for (int b = 0; b < dout[1]; ++b)
      for (int a = 0; a < dout[0]; ++a) {
        int o[3] = {a, b, c};
        const int i = o[ax];
        double s = 0.0;
        for (int j = 0; j < nin; ++j) {
          int src[3] = {a, b, c};
          src[ax] = j;
          const double m = transpose ? M[j * cols + i] : M[i * cols + j];
          s += m * in[(src[2] * din[1] + src[1]) * din[0] + src[0]];
        }
        out[(c * dout[1] + b) * dout[0] + a] = s;
      }
}

/* grad(cell_map) at reference point xi: Jt[i][j] = d x_j / d xi_i. */
static void cell_map_gradient(int dim, const double* Xe, const double* xi,
                              double Jt[3][3]) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) Jt[i][j] = (i == j && i >= dim) ? 1.0 : 0.0;
  const int nv = 1 << dim;
  for (int v = 0; v < nv; ++v) {
    double s[3], ds[3];
    for (int d = 0; d < dim; ++d) {
      const int bit = (v >> d) & 1;
      s[d] = bit ? xi[d] : 1.0 - xi[d];
      ds[d] = bit ? 1.0 : -1.0;
    }
    for (int i = 0; i < dim; ++i) {
      double dn = ds[i];
      for (int d = 0; d < dim; ++d)
        if (d != i) dn *= s[d];
      for (int j = 0; j < dim; ++j) Jt[i][j] += dn * Xe[v * dim + j];
    }
  }
}

/* inverse and |det| of a dim x dim matrix (explicit formulae). */
static int inv_meas(int dim, double A[3][3], double Ai[3][3], double* meas) {
  if (dim == 2) {
    const double det = A[0][0] * A[1][1] - A[0][1] * A[1][0];
    if (det == 0.0 || !isfinite(det)) return 1;
    Ai[0][0] = A[1][1] / det;
    Ai[0][1] = -A[0][1] / det;
    Ai[1][0] = -A[1][0] / det;
    Ai[1][1] = A[0][0] / det;
    *meas = fabs(det);
    return 0;
  }
  const double c00 = A[1][1] * A[2][2] - A[1][2] * A[2][1];
  const double c01 = A[1][2] * A[2][0] - A[1][0] * A[2][2];
  const double c02 = A[1][0] * A[2][1] - A[1][1] * A[2][0];
  const double det = A[0][0] * c00 + A[0][1] * c01 + A[0][2] * c02;
  if (det == 0.0 || !isfinite(det)) return 1;
  Ai[0][0] = c00 / det;
  Ai[1][0] = c01 / det;
  Ai[2][0] = c02 / det;
  Ai[0][1] = (A[0][2] * A[2][1] - A[0][1] * A[2][2]) / det;
  Ai[1][1] = (A[0][0] * A[2][2] - A[0][2] * A[2][0]) / det;
  Ai[2][1] = (A[0][1] * A[2][0] - A[0][0] * A[2][1]) / det;
  Ai[0][2] = (A[0][1] * A[1][2] - A[0][2] * A[1][1]) / det;
  Ai[1][2] = (A[0][2] * A[1][0] - A[0][0] * A[1][2]) / det;
  Ai[2][2] = (A[0][0] * A[1][1] - A[0][1] * A[1][0]) / det;
  *meas = fabs(det);
  return 0;
}

/* Quadrature-point factor of one cell: M[qp][r][k] = wq * djq * kappa *
 * sum_l jq[l][r] jq[l][k]. */
static int cell_metric(int dim, int q, const double* pts, const double* w1d,
                       const double* Xe, const double* kap, double* M) {
  const int nq = dim == 3 ? q * q * q : q * q;
  for (int t = 0; t < nq; ++t) {
    const int i = t % q, k = (t / q) % q, l = t / (q * q);
    double xi[3] = {pts[i], pts[k], dim == 3 ? pts[l] : 0.0};
    double wq = w1d[i] * w1d[k] * (dim == 3 ? w1d[l] : 1.0);
    double Jt[3][3], jq[3][3], djq;
    cell_map_gradient(dim, Xe, xi, Jt);
    if (inv_meas(dim, Jt, jq, &djq)) return 1;
    const double f = wq * djq * (kap ? kap[t] : 1.0);
    for (int r = 0; r < dim; ++r)
      for (int kk = 0; kk < dim; ++kk) {
        double s = 0.0;
        for (int ll = 0; ll < dim; ++ll) s += jq[ll][r] * jq[ll][kk];
        M[(t * dim + r) * dim + kk] = s * f;
      }
  }
  return 0;
}

/* Reference gradients of the cell function xc (tensor order) at the
 * quadrature points: Zq[k][qp], k = component, via the D x D contractions. */
static void dof2quad(int dim, int n, int q, const double* B, const double* Dm,
                     const double* xc, double* Zq, double* w1, double* w2) {
  const int nl = dim == 3 ? n * n * n : n * n;
  const int nq = dim == 3 ? q * q * q : q * q;
  for (int k = 0; k < dim; ++k) {
    /* Z[1][k] = x (array2tensor) */
    memcpy(w1, xc, sizeof(double) * nl);
    int dims[3] = {n, n, dim == 3 ? n : 1};
    double* cur = w1;
    double* nxt = w2;
    for (int d = 0; d < dim; ++d) {
      const double* M = (d == k) ? Dm : B;
      contract(cur, nxt, dims, d, M, q, n, 0);
      dims[d] = q;
      double* t = cur;
      cur = nxt;
      nxt = t;
    }
    memcpy(Zq + (size_t)k * nq, cur, sizeof(double) * nq);
  }
}

/* Transposed contractions, summed over components: yc = sum_k (mats^T) Zq[k] */
static void quad2dof(int dim, int n, int q, const double* B, const double* Dm,
                     const double* Zq, double* yc, double* w1, double* w2) {
  const int nl = dim == 3 ? n * n * n : n * n;
  const int nq = dim == 3 ? q * q * q : q * q;
  for (int i = 0; i < nl; ++i) yc[i] = 0.0;
  for (int k = 0; k < dim; ++k) {
    memcpy(w1, Zq + (size_t)k * nq, sizeof(double) * nq);
    int dims[3] = {q, q, dim == 3 ? q : 1};
    double* cur = w1;
    double* nxt = w2;
    for (int d = dim - 1; d >= 0; --d) {
      const double* M = (d == k) ? Dm : B;
      contract(cur, nxt, dims, d, M, q, n, 1);
      dims[d] = n;
      double* t = cur;
      cur = nxt;
      nxt = t;
    }
    for (int i = 0; i < nl; ++i) yc[i] += cur[i];
  }
}

static int max3(int a, int b) { return a > b ? a : b; }

/* y = A(kappa) x, y overwritten.  Returns 0, or 1 on a singular Jacobian. */
int oracle_apply(int dim, int n, int q, int64_t ncell, int64_t nfree,
                 const double* B, const double* Dm, const double* pts,
                 const double* w1d, const int32_t* cells,
                 const double* verts, const double* kappa, const double* x,
                 double* y, int nthreads) {
  const int nl = dim == 3 ? n * n * n : n * n;
  const int nq = dim == 3 ? q * q * q : q * q;
  const int nv = 1 << dim;
  const int wsz = max3(nl, nq) * 4;
  int err = 0;
  if (nthreads < 1) nthreads = 1;
  double* ypriv = (double*)calloc((size_t)nthreads * (size_t)(nfree > 0 ? nfree : 1), sizeof(double));
  if (!ypriv) return 2;
#pragma omp parallel num_threads(nthreads) reduction(| : err)
  {
    int tid = 0;
#ifdef _OPENMP
    tid = omp_get_thread_num();
#endif
    double* yl = ypriv + (size_t)tid * (size_t)nfree;
    double* xc = (double*)malloc(sizeof(double) * nl);
    double* yc = (double*)malloc(sizeof(double) * nl);
    double* Zq = (double*)malloc(sizeof(double) * dim * nq);
    double* M = (double*)malloc(sizeof(double) * nq * dim * dim);
    double* w1 = (double*)malloc(sizeof(double) * wsz);
    double* w2 = (double*)malloc(sizeof(double) * wsz);
#pragma omp for schedule(static)
    for (int64_t c = 0; c < ncell; ++c) {
      const int32_t* ids = cells + c * nl;
      for (int i = 0; i < nl; ++i) xc[i] = ids[i] >= 0 ? x[ids[i]] : 0.0;
      if (cell_metric(dim, q, pts, w1d, verts + c * nv * dim,
                      kappa ? kappa + c * nq : NULL, M)) {
        err |= 1;
        continue;
      }
      dof2quad(dim, n, q, B, Dm, xc, Zq, w1, w2);
      for (int t = 0; t < nq; ++t) {
        double z[3], cr[3];
        for (int k = 0; k < dim; ++k) z[k] = Zq[(size_t)k * nq + t];
        for (int r = 0; r < dim; ++r) {
          double s = 0.0;
          for (int k = 0; k < dim; ++k) s += M[(t * dim + r) * dim + k] * z[k];
          cr[r] = s;
        }
        for (int r = 0; r < dim; ++r) Zq[(size_t)r * nq + t] = cr[r];
      }
      quad2dof(dim, n, q, B, Dm, Zq, yc, w1, w2);
      for (int i = 0; i < nl; ++i)
        if (ids[i] >= 0) yl[ids[i]] += yc[i];
    }
    free(xc); free(yc); free(Zq); free(M); free(w1); free(w2);
  }
  for (int64_t i = 0; i < nfree; ++i) {
    double s = 0.0;
    for (int t = 0; t < nthreads; ++t) s += ypriv[(size_t)t * nfree + i];
    y[i] = s;
  }
  free(ypriv);
  return err;
}

/* Geometric factor in the library's layout [cell][entry][qp], entries
 * (2D) 00 01 11, (3D) 00 01 02 11 12 22, from the reference's product
 * wq * djq * sum_l jq[l][r] jq[l][k]. */
int oracle_geometry(int dim, int q, int64_t ncell, const double* pts,
                    const double* w1d, const double* verts, double* G) {
  const int nq = dim == 3 ? q * q * q : q * q;
  const int nv = 1 << dim;
  const int ng = dim * (dim + 1) / 2;
  const int pr[6] = {0, 0, 0, 1, 1, 2}, pc[6] = {0, 1, 2, 1, 2, 2};
  const int pr2[3] = {0, 0, 1}, pc2[3] = {0, 1, 1};
  double* M = (double*)malloc(sizeof(double) * nq * dim * dim);
  int err = 0;
  for (int64_t c = 0; c < ncell && !err; ++c) {
    if (cell_metric(dim, q, pts, w1d, verts + c * nv * dim, NULL, M)) {
      err = 1;
      break;
    }
    for (int e = 0; e < ng; ++e) {
      const int r = dim == 3 ? pr[e] : pr2[e];
      const int k = dim == 3 ? pc[e] : pc2[e];
      for (int t = 0; t < nq; ++t)
        G[(c * ng + e) * nq + t] = M[(t * dim + r) * dim + k];
    }
  }
  free(M);
  return err;
}

/* dJ/dkappa[cell][qp] = grad(lam) . (wq djq jq^T jq) grad(u) at each
 * quadrature point (kappa-free factor): the derivative of <lam, A(kappa) u>. */
int oracle_kappa_grad(int dim, int n, int q, int64_t ncell, const double* B,
                      const double* Dm, const double* pts, const double* w1d,
                      const int32_t* cells, const double* verts,
                      const double* u, const double* lam, double* dJdk,
                      int nthreads) {
  const int nl = dim == 3 ? n * n * n : n * n;
  const int nq = dim == 3 ? q * q * q : q * q;
  const int nv = 1 << dim;
  const int wsz = max3(nl, nq) * 4;
  int err = 0;
  if (nthreads < 1) nthreads = 1;
#pragma omp parallel num_threads(nthreads) reduction(| : err)
  {
    double* uc = (double*)malloc(sizeof(double) * nl);
    double* lc = (double*)malloc(sizeof(double) * nl);
    double* Zu = (double*)malloc(sizeof(double) * dim * nq);
    double* Zl = (double*)malloc(sizeof(double) * dim * nq);
    double* M = (double*)malloc(sizeof(double) * nq * dim * dim);
    double* w1 = (double*)malloc(sizeof(double) * wsz);
    double* w2 = (double*)malloc(sizeof(double) * wsz);
#pragma omp for schedule(static)
    for (int64_t c = 0; c < ncell; ++c) {
      const int32_t* ids = cells + c * nl;
      for (int i = 0; i < nl; ++i) {
        uc[i] = ids[i] >= 0 ? u[ids[i]] : 0.0;
        lc[i] = ids[i] >= 0 ? lam[ids[i]] : 0.0;
      }
      if (cell_metric(dim, q, pts, w1d, verts + c * nv * dim, NULL, M)) {
        err |= 1;
        continue;
      }
      dof2quad(dim, n, q, B, Dm, uc, Zu, w1, w2);
      dof2quad(dim, n, q, B, Dm, lc, Zl, w1, w2);
      for (int t = 0; t < nq; ++t) {
        double s = 0.0;
        for (int r = 0; r < dim; ++r)
          for (int k = 0; k < dim; ++k)
            s += Zl[(size_t)r * nq + t] * M[(t * dim + r) * dim + k] *
                 Zu[(size_t)k * nq + t];
        dJdk[c * nq + t] = s;
      }
    }
    free(uc); free(lc); free(Zu); free(Zl); free(M); free(w1); free(w2);
  }
  return err;
}
