// sf_api.hip -- the C ABI (include/sfhip.h): plans, tables, set-up, apply,
// adjoint and the device-resident solvers.  Host code only; kernels live in
// the other translation units.
#include <math.h>
#include <stdarg.h>
#include <stdlib.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <climits>
#include <deque>
#include <functional>
#include <map>
#include <mutex>
#include <new>
#include <string>
#include <unordered_map>
#include <vector>

#include <rocprofiler-sdk-roctx/roctx.h>

#include "sf_apply.hpp"
#include "sf_common.hpp"
#include "sf_kernels.h"
#include "sfhip.h"

using namespace sf;

namespace {
constexpr size_t kScratchBytes = SF_NOALLOC_SCRATCH_BYTES;  // per GPU
constexpr int kMaxDevices = 64;
}  // namespace

// ---------------------------------------------------------------------------
// errors

namespace {

thread_local std::string g_err;

int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

#define SF_HIP(call)                                                        \
  do {                                                                      \
    hipError_t e_ = (call);                                                 \
    if (e_ != hipSuccess)                                                   \
      return fail(SF_ERR_HIP, "%s failed: %s (%s:%d)", #call,               \
                  hipGetErrorString(e_), __FILE__, __LINE__);               \
  } while (0)

#define SF_TRY try {
#define SF_CATCH                                                       \
  }                                                                    \
  catch (const std::bad_alloc&) {                                      \
    return fail(SF_ERR_ARG, "host allocation failed");                 \
  }                                                                    \
  catch (...) {                                                        \
    return fail(SF_ERR_ARG, "unexpected C++ exception");               \
  }

// A named range on the profiler's marker timeline (rocprofv3
// --marker-trace): the apply, the solves, their chunks of CG iterations,
// the preconditioner set-up.  Costs a call into the marker library when no
// profiler listens.
struct MarkerRange {
  explicit MarkerRange(const char* name) { roctxRangePushA(name); }
  ~MarkerRange() { roctxRangePop(); }
};

// ---------------------------------------------------------------------------
// 1D tables: Gauss-Legendre on [0, 1] and Lagrange polynomials on the
// reference's node order.

// Legendre P_q(z) and P_{q-1}(z) by the three-term recurrence.
void legendre_pair(int q, double z, double* pq, double* pqm1) {
  double pn = z, pm = 1.0;
  for (int k = 2; k <= q; ++k) {
    const double pk = ((2.0 * k - 1.0) * z * pn - (k - 1.0) * pm) / k;
    pm = pn;
    pn = pk;
  }
  *pq = pn;
  *pqm1 = pm;
}

// Gauss-Legendre points / weights on [-1, 1] by Newton iteration on P_q,
// ascending; w = 2 / ((1 - z^2) P_q'(z)^2).
void gauss_legendre(int q, std::vector<double>& x, std::vector<double>& w) {
  x.assign(q, 0.0);
  w.assign(q, 0.0);
  const double pi = 3.14159265358979323846;
  for (int i = 0; i < (q + 1) / 2; ++i) {
    double z = cos(pi * (i + 0.75) / (q + 0.5));  // i-th root from the top
    double pn, pm;
    for (int it = 0; it < 100; ++it) {
      legendre_pair(q, z, &pn, &pm);
      const double dp = q * (z * pn - pm) / (z * z - 1.0);
      const double dz = pn / dp;
      z -= dz;
      if (fabs(dz) <= 1e-15) break;
    }
    legendre_pair(q, z, &pn, &pm);
    const double dp = q * (z * pn - pm) / (z * z - 1.0);
    const double wi = 2.0 / ((1.0 - z * z) * dp * dp);
    x[q - 1 - i] = z;
    x[i] = -z;
    w[i] = wi;
    w[q - 1 - i] = wi;
  }
  if (q % 2 == 1) x[q / 2] = 0.0;
}

// nodes in the reference's 1D order: [0, 1, 1/p, ..., (p-1)/p]
// (Gridap's SEGMENT node order: vertices first, then interior nodes).
std::vector<double> lagrange_nodes(int p) {
  std::vector<double> t;
  if (p == 0) {
    t.push_back(0.5);
    return t;
  }
  t.push_back(0.0);
  t.push_back(1.0);
  for (int k = 1; k < p; ++k) t.push_back((double)k / p);
  return t;
}

int build_tables(int p, int q, double* B, double* Dm, double* pts,
                 double* wts) {
  if (p < 1 || p > kMaxN - 1) return fail(SF_ERR_ARG, "order %d outside 1..%d", p, kMaxN - 1);
  if (q < 1 || q > 64) return fail(SF_ERR_ARG, "nquad %d outside 1..64", q);
  const int n = p + 1;
  std::vector<double> gx, gw;
  gauss_legendre(q, gx, gw);
  const std::vector<double> t = lagrange_nodes(p);
  for (int i = 0; i < q; ++i) {
    const double xi = 0.5 * (gx[i] + 1.0);
    if (pts) pts[i] = xi;
    if (wts) wts[i] = 0.5 * gw[i];
    for (int j = 0; j < n; ++j) {
      // phi_j(x) = prod_{m != j} (x - t_m) / (t_j - t_m)
      double den = 1.0, val = 1.0;
      for (int m = 0; m < n; ++m)
        if (m != j) {
          den *= (t[j] - t[m]);
          val *= (xi - t[m]);
        }
      // phi_j'(x) = sum_k prod_{m != j, k} (x - t_m) / den
      double der = 0.0;
      for (int k = 0; k < n; ++k) {
        if (k == j) continue;
        double pr = 1.0;
        for (int m = 0; m < n; ++m)
          if (m != j && m != k) pr *= (xi - t[m]);
        der += pr;
      }
      if (B) B[i * n + j] = val / den;
      if (Dm) Dm[i * n + j] = der / den;
    }
  }
  return SF_OK;
}

// Gauss-Lobatto-Legendre points of degree p on [0, 1] in the reference's
// node order [0, 1, interior ascending]: the roots of (1 - x^2) P_p'(x) by
// the Newton iteration x <- x - (x P_p - P_{p-1}) / ((p + 1) P_p) from the
// Chebyshev-Lobatto points cos(pi k / p).
std::vector<double> lobatto_nodes(int p) {
  std::vector<double> x(p + 1);
  const double pi = 3.14159265358979323846;
  for (int k = 0; k <= p; ++k) {
    double z = cos(pi * k / p);
    for (int it = 0; it < 100; ++it) {
      double pn, pm;
      legendre_pair(p, z, &pn, &pm);
      const double dz = (z * pn - pm) / ((p + 1) * pn);
      z -= dz;
      if (fabs(dz) <= 1e-16) break;
    }
    x[k] = 0.5 * (z + 1.0);
  }
  std::sort(x.begin(), x.end());
  std::vector<double> t;
  t.push_back(0.0);
  t.push_back(1.0);
  for (int k = 1; k < p; ++k) t.push_back(x[k]);
  return t;
}

// T1[a][b] = l_b(t_a): the GLL Lagrange polynomials (nodes s) at the
// equispaced nodes t, both in the reference's order.  Rows of the two
// vertex nodes are exact unit rows.
std::vector<double> gll_to_equispaced(int p) {
  const int n = p + 1;
  const std::vector<double> s = lobatto_nodes(p), t = lagrange_nodes(p);
  std::vector<double> T(n * n);
  for (int a = 0; a < n; ++a)
    for (int b = 0; b < n; ++b) {
      double v = 1.0;
      for (int m = 0; m < n; ++m)
        if (m != b) v *= (t[a] - s[m]) / (s[b] - s[m]);
      T[a * n + b] = v;
    }
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < n; ++b) T[a * n + b] = (a == b) ? 1.0 : 0.0;
  return T;
}

// Inverse of a small dense matrix (row-major n x n) by Gauss-Jordan with
// partial pivoting.
std::vector<double> invert_small(std::vector<double> A, int n) {
  std::vector<double> I(n * n, 0.0);
  for (int i = 0; i < n; ++i) I[i * n + i] = 1.0;
  for (int c = 0; c < n; ++c) {
    int piv = c;
    for (int r = c + 1; r < n; ++r)
      if (fabs(A[r * n + c]) > fabs(A[piv * n + c])) piv = r;
    for (int k = 0; k < n; ++k) {
      std::swap(A[c * n + k], A[piv * n + k]);
      std::swap(I[c * n + k], I[piv * n + k]);
    }
    const double d = 1.0 / A[c * n + c];
    for (int k = 0; k < n; ++k) {
      A[c * n + k] *= d;
      I[c * n + k] *= d;
    }
    for (int r = 0; r < n; ++r) {
      if (r == c) continue;
      const double f = A[r * n + c];
      if (f == 0.0) continue;
      for (int k = 0; k < n; ++k) {
        A[r * n + k] -= f * A[c * n + k];
        I[r * n + k] -= f * I[c * n + k];
      }
    }
  }
  return I;
}

// ---------------------------------------------------------------------------
// device buffer helper

template <typename T>
struct DevBuf {
  T* p = nullptr;
  int64_t n = 0;
  hipError_t alloc(int64_t count) {
    if (p && n == count) return hipSuccess;
    release();
    if (count <= 0) return hipSuccess;
    hipError_t e = hipMalloc((void**)&p, sizeof(T) * (size_t)count);
    if (e != hipSuccess) {
      p = nullptr;
      return e;
    }
    n = count;
    return hipSuccess;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  int64_t bytes() const { return n * (int64_t)sizeof(T); }
};

struct DeviceGuard {
  int prev = -1;
  bool changed = false;
  explicit DeviceGuard(int dev) {
    if (hipGetDevice(&prev) == hipSuccess && prev != dev) {
      changed = (hipSetDevice(dev) == hipSuccess);
    }
  }
  ~DeviceGuard() {
    if (changed) (void)hipSetDevice(prev);
  }
};

}  // namespace

// ---------------------------------------------------------------------------
// plan

struct sf_plan {
  int dim = 0, order = 0, n = 0, q = 0, ng = 0;
  int64_t ncell = 0, nfree = 0, nloc = 0, nqc = 0;
  int device = 0;
  std::vector<double> hB, hD, hW, hPts;
  DevBuf<double> dB, dD, dW, dPts;
  DevBuf<float> dB32, dD32;
  DevBuf<int32_t> cells;
  DevBuf<double> G, Gk, kappa;
  bool gk_valid = false;  // Gk = kappa * G is current (built on first use)
  DevBuf<float> kappa32;  // FP32 copy for the FP32 fused-geometry apply
  bool k32_valid = false;
  DevBuf<double> X;   // cell vertices (G is recomputed from them for new weights)
  DevBuf<double> Xc;  // each cell's map coefficients (kGeoCoef3 / kGeoCoef2
                      // values): what the fused-geometry kernels read
  DevBuf<float> X32, Pts32, W32;  // FP32 copies (of Xc, points, weights) for
                                  // the FP32 fused apply
  bool fused_geom = false;
  DevBuf<float> G32;
  DevBuf<bf16s> Gbf;
  DevBuf<uint32_t> Gm;  // bf16 pairs in the layout of the MFMA apply (3D)
  DevBuf<float> Gm32;   // FP32 G in the layout of the FP32-operand MFMA apply
  bool gm32_valid = false;
  float gmax = 0.f;     // max |G| of Gm (the FP16 GEMM apply's flux scale)
  // max absolute row sum (forward) and column sum (transposed) of the 1D
  // tables in use, the Gauss-Lobatto ones included once built: the bounds
  // behind the FP16 GEMM apply's scales
  double tab_rowB = 0.0, tab_rowD = 0.0, tab_colB = 0.0, tab_colD = 0.0;
  bool has_cells = false, has_geom = false, has_kappa = false;
  bool g32_valid = false, gbf_valid = false, gm_valid = false;
  DevBuf<int> flag;
  // block structure of the (block-assembled) apply
  int eb = 0, nblk = 0;
  int64_t npart = 0;  // shares of p.Ap the 3D apply writes: one per wave
  DevBuf<uint16_t> pidx;
  DevBuf<int32_t> bdofs, bexcl, cell_block;
  DevBuf<int64_t> boff, bcell;
  BlockArgs kargs{};
  // staging of the host-pointer apply
  DevBuf<double> xs, ys;
  // solver workspace
  DevBuf<double> w_r, w_p, w_q, w_t;
  DevBuf<float> w_r32, w_s32, w_d32, w_p32, w_q32;
  DevBuf<double> red_part;
  DevBuf<double> w_part;  // 3D: the apply's per-wave shares of p.Ap
  DevBuf<unsigned> red_counter;
  DevBuf<CGScalars> scal;
  DevBuf<double> dscal;  // dot results
  // preconditioner of the solvers (sf_plan_set_preconditioner): the kind,
  // and the FP32 inverse diagonal, rebuilt when the operator changes
  int pc_kind = SF_PC_NONE;
  bool pc_valid = false;
  DevBuf<float> dinv;  // Jacobi: 1 / diag(A); GLL: 1 / diag(T^T A T)
  // GLL preconditioner: the n x n table and its transpose, 1 / (cells per
  // dof), and the z / scratch vectors of the solvers
  DevBuf<double> dT1, dT1t, dT1i;  // T1, its transpose, its inverse (n x n)
  DevBuf<double> dBg, dDg;         // B T1, D T1 (+ transposes): the GLL basis
  DevBuf<float> dBg32, dDg32;      //   at the quadrature points, FP64 / FP32
  DevBuf<float> minv;
  DevBuf<double> w_w;              // scratch of sf_precond_apply
  DevBuf<double> g_b, g_y;         // a solve in the GLL basis: T^T b and y
  int max_nu = 0;  // most distinct dofs of one block
  // 3D cell order: the plan's cell i is the caller's cell perm[i] (empty:
  // the caller's order); corig is its device copy (for outputs per cell)
  std::vector<int32_t> perm;
  DevBuf<int32_t> corig;
  // apply-kernel timing (sf_plan_set_timing): event pairs recorded around
  // each apply kernel, the first `nrec` in use
  bool timing = false;
  std::vector<hipEvent_t> ev;
  size_t nrec = 0;
  hipStream_t own_stream = nullptr;

  // the stored geometric factor the stored-G forms read (after ensure_gk)
  const double* active_G() const { return has_kappa ? Gk.p : G.p; }
  const BlockArgs* blocks() const { return &kargs; }

  void release_all() {
    dB.release(); dD.release(); dW.release(); dPts.release();
    dB32.release(); dD32.release();
    cells.release(); G.release(); Gk.release(); kappa.release(); kappa32.release();
    X.release(); Xc.release();
    X32.release(); Pts32.release(); W32.release();
    G32.release(); Gbf.release(); Gm.release(); Gm32.release(); flag.release();
    pidx.release(); bdofs.release(); bexcl.release(); cell_block.release();
    boff.release(); bcell.release();
    xs.release(); ys.release();
    w_r.release(); w_p.release(); w_q.release(); w_t.release();
    w_r32.release(); w_s32.release(); w_d32.release(); w_p32.release(); w_q32.release();
    red_part.release(); w_part.release(); red_counter.release(); scal.release();
    dscal.release(); dinv.release();
    dT1.release(); dT1t.release(); dT1i.release(); minv.release(); corig.release();
    dBg.release(); dDg.release(); dBg32.release(); dDg32.release();
    w_w.release(); g_b.release(); g_y.release();
    if (own_stream) (void)hipStreamDestroy(own_stream);
    own_stream = nullptr;
    for (hipEvent_t e : ev) (void)hipEventDestroy(e);
    ev.clear();
    nrec = 0;
  }

  int64_t device_bytes() const {
    return dB.bytes() + dD.bytes() + dW.bytes() + dPts.bytes() + dB32.bytes() +
           dD32.bytes() + cells.bytes() + G.bytes() + Gk.bytes() + X.bytes() + Xc.bytes() +
           X32.bytes() + Pts32.bytes() + W32.bytes() +
           kappa.bytes() + kappa32.bytes() + G32.bytes() + Gbf.bytes() + Gm.bytes() + Gm32.bytes() + flag.bytes() +
           pidx.bytes() + bdofs.bytes() + bexcl.bytes() + cell_block.bytes() +
           boff.bytes() + bcell.bytes() +
           xs.bytes() + ys.bytes() + w_r.bytes() + w_p.bytes() + w_q.bytes() +
           w_t.bytes() + w_r32.bytes() + w_s32.bytes() + w_d32.bytes() + w_p32.bytes() +
           w_q32.bytes() + red_part.bytes() + w_part.bytes() + red_counter.bytes() +
           scal.bytes() + dscal.bytes() + dinv.bytes() + dT1.bytes() + dT1t.bytes() +
           dT1i.bytes() + dBg.bytes() + dDg.bytes() + dBg32.bytes() + dDg32.bytes() +
           minv.bytes() + w_w.bytes() + g_b.bytes() + g_y.bytes() + corig.bytes();
  }
};

namespace {

// Raise the plan's table bounds to cover the q x n tables B, D.
void table_bounds(sf_plan* P, const double* B, const double* D) {
  const int q = P->q, n = P->n;
  auto bound = [&](const double* T, double* row, double* col) {
    for (int i = 0; i < q; ++i) {
      double s = 0.0;
      for (int a = 0; a < n; ++a) s += fabs(T[i * n + a]);
      *row = s > *row ? s : *row;
    }
    for (int a = 0; a < n; ++a) {
      double s = 0.0;
      for (int i = 0; i < q; ++i) s += fabs(T[i * n + a]);
      *col = s > *col ? s : *col;
    }
  };
  bound(B, &P->tab_rowB, &P->tab_colB);
  bound(D, &P->tab_rowD, &P->tab_colD);
}

int upload_tables(sf_plan* P) {
  // device tables: the Q x N table followed by its N x Q transpose (the
  // transposed contractions of the 3D apply read rows of the latter)
  const int q = P->q, n = P->n;
  const int64_t tn = (int64_t)q * n;
  std::vector<double> b64(2 * tn), d64(2 * tn);
  std::vector<float> b32(2 * tn), d32(2 * tn);
  for (int i = 0; i < q; ++i)
    for (int j = 0; j < n; ++j) {
      const double bv = P->hB[(int64_t)i * n + j], dv = P->hD[(int64_t)i * n + j];
      b64[(int64_t)i * n + j] = b64[tn + (int64_t)j * q + i] = bv;
      d64[(int64_t)i * n + j] = d64[tn + (int64_t)j * q + i] = dv;
    }
  for (int64_t i = 0; i < 2 * tn; ++i) {
    b32[i] = (float)b64[i];
    d32[i] = (float)d64[i];
  }
  P->tab_rowB = P->tab_rowD = P->tab_colB = P->tab_colD = 0.0;
  table_bounds(P, P->hB.data(), P->hD.data());
  SF_HIP(P->dB.alloc(2 * tn));
  SF_HIP(P->dD.alloc(2 * tn));
  SF_HIP(P->dW.alloc(q));
  SF_HIP(P->dPts.alloc(q));
  SF_HIP(P->dB32.alloc(2 * tn));
  SF_HIP(P->dD32.alloc(2 * tn));
  SF_HIP(hipMemcpy(P->dB.p, b64.data(), 2 * tn * 8, hipMemcpyHostToDevice));
  SF_HIP(hipMemcpy(P->dD.p, d64.data(), 2 * tn * 8, hipMemcpyHostToDevice));
  SF_HIP(hipMemcpy(P->dW.p, P->hW.data(), q * 8, hipMemcpyHostToDevice));
  SF_HIP(hipMemcpy(P->dPts.p, P->hPts.data(), q * 8, hipMemcpyHostToDevice));
  SF_HIP(hipMemcpy(P->dB32.p, b32.data(), 2 * tn * 4, hipMemcpyHostToDevice));
  SF_HIP(hipMemcpy(P->dD32.p, d32.data(), 2 * tn * 4, hipMemcpyHostToDevice));
  return SF_OK;
}

int check_plan(const sf_plan* P) {
  if (!P) return fail(SF_ERR_ARG, "null plan");
  return SF_OK;
}

int check_ready(const sf_plan* P) {
  if (!P) return fail(SF_ERR_ARG, "null plan");
  if (!P->has_cells) return fail(SF_ERR_NOT_READY, "plan has no cell dofs (sf_plan_set_cells)");
  if (!P->has_geom)
    return fail(SF_ERR_NOT_READY, "plan has no geometry (sf_plan_set_vertices / sf_plan_set_geometry)");
  return SF_OK;
}

// A system without free dofs (every dof Dirichlet) is solved as it stands:
// the solvers return at once, converged after no iteration.
bool empty_solve(const sf_plan* P, sf_solve_stats* stats) {
  if (P->nfree > 0) return false;
  if (stats) *stats = sf_solve_stats{0, 0, 1, 0.0, 0.0};
  return true;
}

int read_flag(sf_plan* P, int* out) {
  int h = 0;
  SF_HIP(hipMemcpy(&h, P->flag.p, sizeof(int), hipMemcpyDeviceToHost));
  *out = h;
  return SF_OK;
}

int clear_flag(sf_plan* P) {
  SF_HIP(P->flag.alloc(1));
  SF_HIP(hipMemset(P->flag.p, 0, sizeof(int)));
  return SF_OK;
}

// The 3D fused-geometry applies (vertices kept, FP32 copies made): the
// forms the fused CG step and the block-assembled adjoint build on.
bool fused_f64(const sf_plan* P) { return P->fused_geom && P->dim == 3 && P->Xc.p; }
bool fused_f32(const sf_plan* P) {
  return P->fused_geom && P->dim == 3 && P->X32.p && (!P->has_kappa || P->k32_valid);
}
// Whether the apply recomputes G from the vertices: the 3D forms above, and
// in 2D without kappa (a 2D kappa scales the stored G).
bool geo_f64(const sf_plan* P) {
  return P->dim == 3 ? fused_f64(P) : (P->fused_geom && P->Xc.p && !P->has_kappa);
}
bool geo_f32(const sf_plan* P) {
  return P->dim == 3 ? fused_f32(P) : (P->fused_geom && P->X32.p && !P->has_kappa);
}

// Events around one apply kernel when the plan records timings.
struct ApplyTimer {
  sf_plan* P;
  hipStream_t s;
  hipEvent_t stop = nullptr;
  ApplyTimer(sf_plan* P_, hipStream_t s_) : P(P_), s(s_) {
    if (!P->timing) return;
    if (P->nrec * 2 + 2 > P->ev.size()) {
      hipEvent_t a = nullptr, b = nullptr;
      if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) return;
      P->ev.push_back(a);
      P->ev.push_back(b);
    }
    if (hipEventRecord(P->ev[2 * P->nrec], s) != hipSuccess) return;
    stop = P->ev[2 * P->nrec + 1];
  }
  ~ApplyTimer() {
    if (stop && hipEventRecord(stop, s) == hipSuccess) ++P->nrec;
  }
};

// kappa * G for the forms that read a stored geometric factor (the stored-G
// applies, the reduced-precision G copies, the diagonal), built on first use
// after kappa or the geometry changed: the fused-geometry kernels read kappa
// per point, so setting kappa on such a plan (the rrule's forward call, a
// finite-difference loop) costs no pass over the stored G.
int ensure_gk(sf_plan* P) {
  if (!P->has_kappa || P->gk_valid || !P->has_geom) return SF_OK;
  SF_HIP(P->Gk.alloc(P->ncell * P->ng * P->nqc));
  if (P->dim == 3)
    SF_HIP(launch_scale_kappa3(P->G.p, P->kappa.p, P->q, P->ncell, P->cell_block.p,
                               P->bcell.p, P->Gk.p, nullptr));
  else
    SF_HIP(launch_scale_kappa(P->G.p, P->kappa.p, P->ng, P->nqc, P->ncell, P->Gk.p, nullptr));
  SF_HIP(hipDeviceSynchronize());
  P->gk_valid = true;
  return SF_OK;
}

// y = A x (FP64) on stream s.  accumulate: y += A x.  zeroed: the caller
// guarantees y == 0 already (the CG direction update clears it), so the
// fill is skipped.
int do_apply_f64(sf_plan* P, double* y, const double* x, hipStream_t s,
                 int accumulate = 0, bool zeroed = false, double* dotp = nullptr) {
  if (P->nfree == 0) return SF_OK;  // nothing to compute (and x may be empty)
  if (!accumulate && !zeroed) SF_HIP(launch_fill<double>(y, P->nfree, 0.0, s));
  // the fused-geometry kernel when the plan has vertices (3D: kappa, if
  // set, read per quadrature point)
  const bool geo = geo_f64(P);
  if (!geo) {
    const int rc = ensure_gk(P);
    if (rc) return rc;
  }
  ApplyTimer timer(P, s);
  hipError_t e = apply_f64(P->dim, P->n, P->q, x, y, P->cells.p, P->active_G(),
                           P->dB.p, P->dD.p, P->ncell, P->blocks(), accumulate, s,
                           geo ? P->Xc.p : nullptr, P->dPts.p, P->dW.p, dotp,
                           geo && P->dim == 3 && P->has_kappa ? P->kappa.p : nullptr);
  if (e == hipErrorNotSupported)
    return fail(SF_ERR_UNSUPPORTED, "no kernel for dim=%d n=%d q=%d", P->dim, P->n, P->q);
  SF_HIP(e);
  return SF_OK;
}

bool is_reduced(int precision) {
  return precision == SF_PREC_F32 || precision == SF_PREC_F32_GBF16 ||
         precision == SF_PREC_MFMA_BF16 || precision == SF_PREC_MFMA_BF16X2 ||
         precision == SF_PREC_MFMA_BF16X3 || precision == SF_PREC_MFMA_F32 ||
         precision == SF_PREC_MFMA_F16;
}

// 2^-e with v 2^-e in [0.5, 1) (1 for v = 0): the scale that brings a
// bound v below 1.
float pow2_below(double v) {
  if (!(v > 0.0) || !isfinite(v)) return 1.f;
  int e = 0;
  (void)frexp(v, &e);
  return (float)ldexp(1.0, -e);
}

// The FP16 GEMM apply's conversion scales from the bounds of each table
// (row sums forward, column sums transposed) and max |G|.  Scales along the
// way: U 1 (x scaled per block), gradient paths x: pD pB, y: pB pD,
// z: pB pB, flux s3, W_x, W_y: s3 tB, W_z: s3 tD, S_D: s3 tB^2,
// S_B: s3 tB tD / 2.
GemmScales gemm_scales(const sf_plan* P) {
  GemmScales g;
  const double rB = P->tab_rowB, rD = P->tab_rowD, cB = P->tab_colB, cD = P->tab_colD;
  const double pB = pow2_below(rB), pD = pow2_below(rD), tB = pow2_below(cB),
               tD = pow2_below(cD);
  // |raw gradient| / its path scale: x, y: rB / (pD pB), z: rD / pB^2
  const double gb = fmax(rB / (pD * pB), rD / (pB * pB));
  const double s3 = pow2_below(3.0 * (double)P->gmax * gb);
  g.pB = (float)pB;
  g.pD = (float)pD;
  g.tB = (float)tB;
  g.tD = (float)tD;
  g.cx = (float)(s3 / (pD * pB));
  g.cy = (float)(s3 / (pB * pD));
  g.cz = (float)(s3 / (pB * pB));
  g.hD = (float)(0.5 * tD);  // the y term (scale s3 tB) times tD / 2
  g.hB = (float)(0.5 * tB);  // the z term (scale s3 tD) times tB / 2
  g.bx = (float)(2.0 * tB / tD);  // S_B (scale s3 tB tD / 2) to S_D's s3 tB^2
  g.vD = (float)(1.0 / (s3 * tB * tB));
  return g;
}

int prepare(sf_plan* P, int precision) {
  const int64_t ng = P->ncell * P->ng * P->nqc;
  if (precision == SF_PREC_F64) return SF_OK;
  if (precision == SF_PREC_MFMA_F32) {
    if (P->dim != 3)
      return fail(SF_ERR_UNSUPPORTED, "the MFMA apply is 3D only (dim=%d)", P->dim);
    if (!P->gm32_valid) {
      const int rc = ensure_gk(P);
      if (rc) return rc;
      SF_HIP(P->Gm32.alloc(P->ncell * P->nqc * 6));
      SF_HIP(launch_g3_to_mfma32(P->active_G(), P->Gm32.p, P->q, P->ncell, P->cell_block.p,
                                 P->bcell.p, nullptr));
      SF_HIP(hipDeviceSynchronize());
      P->gm32_valid = true;
    }
    return SF_OK;
  }
  if (precision == SF_PREC_MFMA_BF16 || precision == SF_PREC_MFMA_BF16X2 ||
      precision == SF_PREC_MFMA_BF16X3 || precision == SF_PREC_MFMA_F16) {
    if (P->dim != 3)
      return fail(SF_ERR_UNSUPPORTED, "the MFMA apply is 3D only (dim=%d)", P->dim);
    if (precision == SF_PREC_MFMA_F16 && !gemm_supported(P->n, P->q))
      return fail(SF_ERR_UNSUPPORTED,
                  "the FP16 GEMM apply is built for n = 8, q = 8..10 (got n=%d q=%d)", P->n,
                  P->q);
    if (!P->gm_valid) {
      const int rc = ensure_gk(P);
      if (rc) return rc;
      SF_HIP(P->Gm.alloc(P->ncell * P->nqc * 3));
      SF_HIP(launch_g3_to_mfma(P->active_G(), P->Gm.p, P->q, P->ncell,
                               P->cell_block.p, P->bcell.p, nullptr));
      SF_HIP(P->flag.alloc(1));
      SF_HIP(launch_absmax_bf16pairs(P->Gm.p, P->ncell * P->nqc * 3, (unsigned*)P->flag.p,
                                     nullptr));
      unsigned bits = 0;
      SF_HIP(hipMemcpy(&bits, P->flag.p, sizeof(unsigned), hipMemcpyDeviceToHost));
      memcpy(&P->gmax, &bits, sizeof(float));
      P->gm_valid = true;
    }
    return SF_OK;
  }
  if (precision == SF_PREC_F32) {
    if (P->fused_geom && P->Xc.p && !P->X32.p) {
      // FP32 map coefficients and 1D quadrature for the FP32 fused-geometry
      // apply
      const int64_t nv = P->ncell * (P->dim == 3 ? kGeoCoef3 : kGeoCoef2);
      SF_HIP(P->X32.alloc(nv));
      SF_HIP(P->Pts32.alloc(P->q));
      SF_HIP(P->W32.alloc(P->q));
      SF_HIP(launch_to_f32(P->Xc.p, nv, P->X32.p, nullptr));
      std::vector<float> pw(2 * P->q);
      for (int i = 0; i < P->q; ++i) {
        pw[i] = (float)P->hPts[i];
        pw[P->q + i] = (float)P->hW[i];
      }
      SF_HIP(hipMemcpy(P->Pts32.p, pw.data(), P->q * 4, hipMemcpyHostToDevice));
      SF_HIP(hipMemcpy(P->W32.p, pw.data() + P->q, P->q * 4, hipMemcpyHostToDevice));
      SF_HIP(hipDeviceSynchronize());
    }
    if (P->fused_geom && P->dim == 3 && P->has_kappa && !P->k32_valid) {
      SF_HIP(P->kappa32.alloc(P->ncell * P->nqc));
      SF_HIP(launch_to_f32(P->kappa.p, P->ncell * P->nqc, P->kappa32.p, nullptr));
      SF_HIP(hipDeviceSynchronize());
      P->k32_valid = true;
    }
    // the stored FP32 G only where the FP32 apply reads it (no fused geometry)
    if (!P->g32_valid && !geo_f32(P)) {
      const int rc = ensure_gk(P);
      if (rc) return rc;
      SF_HIP(P->G32.alloc(ng));
      if (P->dim == 3)
        SF_HIP(launch_g3_to_f32_pairs(P->active_G(), P->G32.p, P->q, P->ncell,
                                      P->cell_block.p, P->bcell.p, nullptr));
      else
        SF_HIP(launch_to_f32(P->active_G(), ng, P->G32.p, nullptr));
      SF_HIP(hipDeviceSynchronize());
      P->g32_valid = true;
    }
    return SF_OK;
  }
  if (precision == SF_PREC_F32_GBF16) {
    if (!P->gbf_valid) {
      const int rc = ensure_gk(P);
      if (rc) return rc;
      SF_HIP(P->Gbf.alloc(ng));
      if (P->dim == 3)
        SF_HIP(launch_g3_to_bf16_pairs(P->active_G(), P->Gbf.p, P->q, P->ncell,
                                       P->cell_block.p, P->bcell.p, nullptr));
      else
        SF_HIP(launch_to_bf16(P->active_G(), ng, P->Gbf.p, nullptr));
      SF_HIP(hipDeviceSynchronize());
      P->gbf_valid = true;
    }
    return SF_OK;
  }
  return fail(SF_ERR_UNSUPPORTED, "unknown precision %d", precision);
}

// y = A x in reduced precision (FP32 vectors), y overwritten (zeroed: y
// is already zero).
int do_apply_f32(sf_plan* P, float* y, const float* x, int precision,
                 hipStream_t s, bool zeroed = false, double* dotp = nullptr) {
  if (P->nfree == 0) return SF_OK;
  if (!zeroed) SF_HIP(launch_fill<float>(y, P->nfree, 0.0f, s));
  ApplyTimer timer(P, s);
  hipError_t e;
  if (precision == SF_PREC_F32) {
    const bool geo = geo_f32(P);
    e = apply_f32(P->dim, P->n, P->q, x, y, P->cells.p, P->G32.p, P->dB32.p,
                  P->dD32.p, P->ncell, P->blocks(), 0, s, geo ? P->X32.p : nullptr,
                  P->Pts32.p, P->W32.p, dotp,
                  geo && P->dim == 3 && P->has_kappa ? P->kappa32.p : nullptr);
  } else if (precision == SF_PREC_F32_GBF16) {
    e = apply_f32_gbf16(P->dim, P->n, P->q, x, y, P->cells.p, P->Gbf.p,
                        P->dB32.p, P->dD32.p, P->ncell, P->blocks(), 0, s);
  } else if (precision == SF_PREC_MFMA_F32) {
    e = apply_mfma(P->n, P->q, 0, x, y, P->Gm32.p, P->dB32.p, P->dD32.p, P->blocks(), s);
  } else if (precision == SF_PREC_MFMA_F16) {
    e = apply_gemm(P->n, P->q, x, y, P->Gm.p, P->dB32.p, P->dD32.p, P->blocks(),
                   gemm_scales(P), s);
  } else if (precision == SF_PREC_MFMA_BF16 || precision == SF_PREC_MFMA_BF16X2 ||
             precision == SF_PREC_MFMA_BF16X3) {
    const int split = precision == SF_PREC_MFMA_BF16 ? 1 : (precision == SF_PREC_MFMA_BF16X2 ? 2 : 3);
    e = apply_mfma(P->n, P->q, split, x, y,
                   P->Gm.p, P->dB32.p, P->dD32.p, P->blocks(), s);
  } else {
    return fail(SF_ERR_UNSUPPORTED, "precision %d has no FP32 apply", precision);
  }
  if (e == hipErrorNotSupported)
    return fail(SF_ERR_UNSUPPORTED, "no kernel for dim=%d n=%d q=%d", P->dim, P->n, P->q);
  SF_HIP(e);
  return SF_OK;
}

// Block structure of the apply from the tensor-ordered, 0-based cell
// dofs h (negative = Dirichlet):
//  * blocks of eb consecutive cells (eb = the kernel's cells per workgroup;
//    the last block may hold fewer).  A block may span the end of a mesh
//    row: the kernel computes all eb cell slots of a workgroup whether or
//    not they hold a cell, so cutting at row ends (e.g. a 64-cell row at
//    eb = 7 as 10 blocks of 6-7) left up to 9% of the slots empty, while a
//    block over two rows only lists a few more distinct dofs;
//  * per block, its distinct free dofs, sorted, the ones no other block
//    touches first (the kernel stores those and adds the shared ones with
//    atomics);
//  * per cell dof, its index in the block's list (uint16, 0xffff =
//    Dirichlet); per cell, its block.
std::vector<int64_t> block_starts(int64_t E, int eb);

int build_blocks(sf_plan* P, const std::vector<int32_t>& h) {
  int nt = 0, eb = 0;
  apply_shape_f64(P->dim, P->n, P->q, &nt, &eb);
  if (eb <= 0) return fail(SF_ERR_UNSUPPORTED, "no kernel for n=%d q=%d", P->n, P->q);
  const int64_t E = P->ncell, nl = P->nloc, nf = P->nfree;
  const std::vector<int64_t> bcell = block_starts(E, eb);
  const int64_t nb = (int64_t)bcell.size() - 1;
  if (nb > 0x7fffffffLL) return fail(SF_ERR_ARG, "too many cell blocks");
  // distinct sorted dofs per block
  std::vector<int64_t> off(nb + 1, 0);
  std::vector<int32_t> dofs;
  dofs.reserve((size_t)(E * nl));
  std::vector<int32_t> tmp;
  for (int64_t b = 0; b < nb; ++b) {
    tmp.clear();
    const int64_t c0 = bcell[b], c1 = bcell[b + 1];
    for (int64_t i = c0 * nl; i < c1 * nl; ++i)
      if (h[i] >= 0) tmp.push_back(h[i]);
    std::sort(tmp.begin(), tmp.end());
    tmp.erase(std::unique(tmp.begin(), tmp.end()), tmp.end());
    dofs.insert(dofs.end(), tmp.begin(), tmp.end());
    off[b + 1] = (int64_t)dofs.size();
  }
  // how many blocks touch each dof
  std::vector<int32_t> nblk_of(nf, 0);
  for (int32_t d : dofs) ++nblk_of[d];
  // exclusive dofs first, local indices
  std::vector<int32_t> excl(nb, 0);
  std::vector<uint16_t> li((size_t)(E * nl), kNoDof);
  std::vector<int32_t> cblk((size_t)E);
  std::vector<int32_t> part;
  std::vector<std::pair<int32_t, int32_t>> pos;
  for (int64_t b = 0; b < nb; ++b) {
    part.clear();
    int nex = 0;
    for (int pass = 0; pass < 2; ++pass)
      for (int64_t i = off[b]; i < off[b + 1]; ++i) {
        const bool ex = nblk_of[dofs[i]] == 1;
        if (ex == (pass == 0)) {
          part.push_back(dofs[i]);
          if (pass == 0) ++nex;
        }
      }
    std::copy(part.begin(), part.end(), dofs.begin() + off[b]);
    excl[b] = nex;
    if (off[b + 1] - off[b] > 0xfffe) return fail(SF_ERR_ARG, "block with too many dofs");
    pos.clear();
    for (size_t i = 0; i < part.size(); ++i) pos.push_back({part[i], (int32_t)i});
    std::sort(pos.begin(), pos.end());
    const int64_t c0 = bcell[b], c1 = bcell[b + 1];
    for (int64_t c = c0; c < c1; ++c) cblk[c] = (int32_t)b;
    for (int64_t i = c0 * nl; i < c1 * nl; ++i) {
      if (h[i] < 0) continue;
      auto it = std::lower_bound(pos.begin(), pos.end(), std::make_pair(h[i], (int32_t)-1));
      li[i] = (uint16_t)it->second;
    }
  }
  // upload
  SF_HIP(P->pidx.alloc(std::max<int64_t>((int64_t)li.size(), 1)));
  SF_HIP(P->bdofs.alloc((int64_t)std::max<size_t>(dofs.size(), 1)));
  SF_HIP(P->boff.alloc(nb + 1));
  SF_HIP(P->bcell.alloc(nb + 1));
  SF_HIP(P->bexcl.alloc(std::max<int64_t>(nb, 1)));
  SF_HIP(P->cell_block.alloc(std::max<int64_t>(E, 1)));
  if (!li.empty()) SF_HIP(hipMemcpy(P->pidx.p, li.data(), li.size() * 2, hipMemcpyHostToDevice));
  if (!dofs.empty()) SF_HIP(hipMemcpy(P->bdofs.p, dofs.data(), dofs.size() * 4, hipMemcpyHostToDevice));
  SF_HIP(hipMemcpy(P->boff.p, off.data(), (nb + 1) * 8, hipMemcpyHostToDevice));
  SF_HIP(hipMemcpy(P->bcell.p, bcell.data(), (nb + 1) * 8, hipMemcpyHostToDevice));
  if (nb > 0) SF_HIP(hipMemcpy(P->bexcl.p, excl.data(), nb * 4, hipMemcpyHostToDevice));
  if (E > 0) SF_HIP(hipMemcpy(P->cell_block.p, cblk.data(), E * 4, hipMemcpyHostToDevice));
  P->eb = eb;
  P->nblk = (int)nb;
  P->max_nu = 0;
  for (int64_t b = 0; b < nb; ++b)
    P->max_nu = std::max<int>(P->max_nu, (int)(off[b + 1] - off[b]));
  P->npart = P->dim == 3 ? nb * (nt / 64) : 0;  // p.Ap shares: 3D fused step
  P->kargs = BlockArgs{P->pidx.p, P->bdofs.p, P->boff.p, P->bexcl.p, P->bcell.p, (int)nb};
  return SF_OK;
}

// Rows of `width` values per cell in the caller's cell order, in the
// plan's order (the caller's array itself when the orders agree).
template <typename T>
const T* to_plan_order(const sf_plan* P, const T* src, int64_t width, std::vector<T>& tmp) {
  if (P->perm.empty() || !src) return src;
  tmp.resize((size_t)(P->ncell * width));
  for (int64_t i = 0; i < P->ncell; ++i) {
    const T* row = src + (int64_t)P->perm[i] * width;
    std::copy(row, row + width, tmp.data() + i * width);
  }
  return tmp.data();
}

// Starts of the apply's cell blocks: eb consecutive cells each, the last
// entry E.
std::vector<int64_t> block_starts(int64_t E, int eb) {
  std::vector<int64_t> bcell;
  for (int64_t c = 0; c < E; c += eb) bcell.push_back(c);
  bcell.push_back(E);
  return bcell;
}

// How well a cell order keeps neighbours together: the number of blocks of
// at most eb cells when the sequence (tensor-ordered 0-based dofs h) is cut
// wherever two consecutive cells share no free dof and every run is split
// into near-equal blocks.  A row-wise order is within a few per cent of
// ceil(E / eb); an order whose consecutive cells are not neighbours gives
// about one block per cell.
int64_t run_block_count(const std::vector<int32_t>& h, int64_t E, int64_t nl, int eb) {
  int64_t count = 0;
  std::vector<int32_t> prev, cur;
  int64_t run0 = 0;
  auto close_run = [&](int64_t r0, int64_t r1) { count += (r1 - r0 + eb - 1) / eb; };
  for (int64_t c = 0; c < E; ++c) {
    cur.clear();
    for (int64_t i = c * nl; i < (c + 1) * nl; ++i)
      if (h[i] >= 0) cur.push_back(h[i]);
    std::sort(cur.begin(), cur.end());
    if (c > 0) {
      bool shared = false;
      size_t a = 0, bb = 0;
      while (a < prev.size() && bb < cur.size()) {
        if (prev[a] == cur[bb]) { shared = true; break; }
        if (prev[a] < cur[bb]) ++a; else ++bb;
      }
      if (!shared) {
        close_run(run0, c);
        run0 = c;
      }
    }
    prev.swap(cur);
  }
  if (E > 0) close_run(run0, E);
  return count;
}

// A cell order built from the connectivity alone (3D hexahedra, dofs
// tensor-ordered, 1D node order [0, 1, interior]): face f = 2 d + s of a
// cell is its layer of local index s (t = 0 or 1) along direction d; two
// cells are neighbours across a face when they share its dofs.  Rows are
// chains of face neighbours through opposite faces (a hex mesh's chords),
// each from its low end; the next row is the one across the +y face of
// the row's first cell while there is one, then across the +z faces.  On
// a structured mesh numbered in any order this is the lexicographic
// (z, y, x) order; on an unstructured hex mesh it is a breadth-first walk
// over rows.  Returns order[new position] = caller's cell index.
std::vector<int32_t> connectivity_order(const std::vector<int32_t>& h, int64_t E, int n) {
  const int64_t nl = (int64_t)n * n * n;
  // face keys: the smallest free dof among the face's interior nodes
  // (local indices >= 2 in both in-face directions, p >= 2) -- held by the
  // face's two cells only; at p = 1 the face's four vertices (at least
  // three free: three vertices fix a quadrilateral face)
  std::vector<int64_t> nb(E * 6, -1);
  std::vector<int8_t> nbf(E * 6, -1);
  auto local = [&](int d, int s, int i, int j) -> int64_t {
    int idx[3];
    idx[d] = s;
    idx[d == 0 ? 1 : 0] = i;
    idx[d == 2 ? 1 : 2] = j;
    return ((int64_t)idx[2] * n + idx[1]) * n + idx[0];
  };
  if (n >= 3) {
    std::unordered_map<int32_t, int64_t> first;  // key -> c * 6 + f
    first.reserve((size_t)(E * 3));
    for (int64_t c = 0; c < E; ++c)
      for (int f = 0; f < 6; ++f) {
        const int d = f / 2, sd = f % 2;
        int32_t key = INT32_MAX;
        for (int i = 2; i < n; ++i)
          for (int j = 2; j < n; ++j) {
            const int32_t id = h[c * nl + local(d, sd, i, j)];
            if (id >= 0 && id < key) key = id;
          }
        if (key == INT32_MAX) continue;
        auto it = first.find(key);
        if (it == first.end()) {
          first.emplace(key, c * 6 + f);
        } else {
          const int64_t o = it->second;
          nb[c * 6 + f] = o / 6;
          nbf[c * 6 + f] = (int8_t)(o % 6);
          nb[o] = c;
          nbf[o] = (int8_t)f;
        }
      }
  } else {
    std::map<std::array<int32_t, 4>, int64_t> first;
    for (int64_t c = 0; c < E; ++c)
      for (int f = 0; f < 6; ++f) {
        const int d = f / 2, sd = f % 2;
        std::array<int32_t, 4> key;
        int nfree = 0;
        for (int i = 0; i < 2; ++i)
          for (int j = 0; j < 2; ++j) {
            key[2 * i + j] = h[c * nl + local(d, sd, i, j)];
            nfree += key[2 * i + j] >= 0;
          }
        if (nfree < 3) continue;
        std::sort(key.begin(), key.end());
        auto it = first.find(key);
        if (it == first.end()) {
          first.emplace(key, c * 6 + f);
        } else {
          const int64_t o = it->second;
          nb[c * 6 + f] = o / 6;
          nbf[c * 6 + f] = (int8_t)(o % 6);
          nb[o] = c;
          nbf[o] = (int8_t)f;
        }
      }
  }
  std::vector<char> seen(E, 0);
  std::vector<int32_t> order;
  order.reserve(E);
  std::deque<int64_t> todo;
  std::vector<int64_t> left, row;
  // walk from c across face f, then through the face opposite each entry
  auto walk = [&](int64_t c, int f, std::vector<int64_t>& out) {
    out.clear();
    for (;;) {
      const int64_t nx = nb[c * 6 + f];
      if (nx < 0 || seen[nx]) return;
      const int fo = nbf[c * 6 + f];
      seen[nx] = 1;
      out.push_back(nx);
      c = nx;
      f = fo ^ 1;
    }
  };
  // seeds: cells with no -x, -y, -z neighbour first (mesh corners)
  std::vector<int64_t> seeds;
  for (int64_t c = 0; c < E; ++c)
    if (nb[c * 6 + 0] < 0 && nb[c * 6 + 2] < 0 && nb[c * 6 + 4] < 0) seeds.push_back(c);
  for (int64_t c = 0; c < E; ++c) seeds.push_back(c);
  for (int64_t seed : seeds) {
    if (seen[seed]) continue;
    todo.push_back(seed);
    while (!todo.empty()) {
      const int64_t c = todo.front();
      todo.pop_front();
      if (seen[c]) continue;
      seen[c] = 1;
      walk(c, 0, left);
      row.assign(left.rbegin(), left.rend());
      row.push_back(c);
      walk(c, 1, left);
      row.insert(row.end(), left.begin(), left.end());
      for (int64_t rc : row) order.push_back((int32_t)rc);
      for (int64_t rc : row)
        for (int f = 2; f < 6; ++f) {
          const int64_t nx = nb[rc * 6 + f];
          if (nx >= 0 && !seen[nx]) todo.push_back(nx);
        }
      // the +y neighbour of the row's first cell continues the sheet
      const int64_t ny = nb[row.front() * 6 + 3];
      if (ny >= 0 && !seen[ny]) todo.push_front(ny);
    }
  }
  return order;
}

int ensure_solver_ws(sf_plan* P, bool mixed) {
  if (P->dim == 3 && P->npart > 0) SF_HIP(P->w_part.alloc(P->npart));
  SF_HIP(P->w_r.alloc(P->nfree));
  SF_HIP(P->w_p.alloc(P->nfree));
  SF_HIP(P->w_q.alloc(P->nfree));
  SF_HIP(P->w_t.alloc(P->nfree));
  if (P->pc_kind == SF_PC_GLL) {
    SF_HIP(P->g_b.alloc(P->nfree));
    SF_HIP(P->g_y.alloc(P->nfree));
  }
  if (mixed) {
    SF_HIP(P->w_r32.alloc(P->nfree));
    SF_HIP(P->w_s32.alloc(P->nfree));
    SF_HIP(P->w_d32.alloc(P->nfree));
    SF_HIP(P->w_p32.alloc(P->nfree));
    SF_HIP(P->w_q32.alloc(P->nfree));
  }
  if (!P->red_part.p) {
    SF_HIP(P->red_part.alloc(2 * kPartSlots));
    SF_HIP(P->red_counter.alloc(1));
    SF_HIP(hipMemset(P->red_counter.p, 0, sizeof(unsigned)));
    SF_HIP(P->scal.alloc(1));
    SF_HIP(P->dscal.alloc(4));
  }
  return SF_OK;
}

// diag = diag(A) on the free dofs (FP64, overwritten) on stream s.
int compute_diagonal(sf_plan* P, double* diag, hipStream_t s) {
  if (P->nfree == 0) return SF_OK;
  const int rc = ensure_gk(P);
  if (rc) return rc;
  SF_HIP(launch_fill<double>(diag, P->nfree, 0.0, s));
  hipError_t e = launch_diagonal(P->dim, P->n, P->q, P->active_G(), P->dB.p, P->dD.p,
                                 P->cells.p, P->cell_block.p, P->bcell.p, P->ncell, diag, s);
  if (e == hipErrorNotSupported)
    return fail(SF_ERR_UNSUPPORTED, "no diagonal kernel for n=%d q=%d", P->n, P->q);
  SF_HIP(e);
  return SF_OK;
}

// The solvers' preconditioner, rebuilt after the operator changed.  Jacobi:
// *dinv = the FP32 inverse diagonal of A; GLL: *gll = true and *dinv = that
// of T^T A T (the solve runs in the Gauss-Lobatto basis, solve_gll); none:
// both unset.
int ensure_pc(sf_plan* P, hipStream_t s, const float** dinv, bool* gll = nullptr) {
  *dinv = nullptr;
  if (gll) *gll = false;
  if (P->pc_kind == SF_PC_NONE || P->nfree == 0) return SF_OK;
  if (P->pc_kind == SF_PC_GLL && !gll)
    return fail(SF_ERR_UNSUPPORTED, "the Gauss-Lobatto preconditioner is not available here");
  if (!P->pc_valid) {
    MarkerRange mark("sfhip preconditioner set-up");
    const int rk = ensure_gk(P);  // the diagonal kernels read kappa * G
    if (rk) return rk;
    DevBuf<double> d, tb, td;
    SF_HIP(d.alloc(P->nfree));
    SF_HIP(P->dinv.alloc(P->nfree));
    int rc = SF_OK;
    if (P->pc_kind == SF_PC_GLL) {
      // T1, and the GLL basis at the quadrature points: B T1, D T1 (q x n)
      const int n = P->n, q = P->q;
      const std::vector<double> T = gll_to_equispaced(P->order);
      std::vector<double> Tt(n * n), bg(2 * q * n), dg(2 * q * n);
      for (int a = 0; a < n; ++a)
        for (int b = 0; b < n; ++b) Tt[b * n + a] = T[a * n + b];
      for (int i = 0; i < q; ++i)
        for (int b = 0; b < n; ++b) {
          double vb = 0.0, vd = 0.0;
          for (int a = 0; a < n; ++a) {
            vb += P->hB[i * n + a] * T[a * n + b];
            vd += P->hD[i * n + a] * T[a * n + b];
          }
          bg[i * n + b] = vb;
          dg[i * n + b] = vd;
        }
      // T1^{-1} has unit vertex rows as well (the forward pass relies on it)
      std::vector<double> Ti = invert_small(T, n);
      for (int a = 0; a < 2; ++a)
        for (int b = 0; b < n; ++b) Ti[a * n + b] = (a == b) ? 1.0 : 0.0;
      SF_HIP(P->dT1.alloc(n * n));
      SF_HIP(P->dT1t.alloc(n * n));
      SF_HIP(P->dT1i.alloc(n * n));
      SF_HIP(tb.alloc(q * n));
      SF_HIP(td.alloc(q * n));
      SF_HIP(P->minv.alloc(P->nfree));
      SF_HIP(hipMemcpy(P->dT1.p, T.data(), n * n * 8, hipMemcpyHostToDevice));
      SF_HIP(hipMemcpy(P->dT1t.p, Tt.data(), n * n * 8, hipMemcpyHostToDevice));
      SF_HIP(hipMemcpy(P->dT1i.p, Ti.data(), n * n * 8, hipMemcpyHostToDevice));
      SF_HIP(hipMemcpy(tb.p, bg.data(), q * n * 8, hipMemcpyHostToDevice));
      SF_HIP(hipMemcpy(td.p, dg.data(), q * n * 8, hipMemcpyHostToDevice));
      table_bounds(P, bg.data(), dg.data());
      // the GLL tables in the apply's layout (q x n table, then its transpose)
      {
        const int64_t tn = (int64_t)q * n;
        std::vector<double> b64(2 * tn), d64(2 * tn);
        std::vector<float> b32(2 * tn), d32(2 * tn);
        for (int i = 0; i < q; ++i)
          for (int j = 0; j < n; ++j) {
            b64[i * n + j] = b64[tn + j * q + i] = bg[i * n + j];
            d64[i * n + j] = d64[tn + j * q + i] = dg[i * n + j];
          }
        for (int64_t i = 0; i < 2 * tn; ++i) {
          b32[i] = (float)b64[i];
          d32[i] = (float)d64[i];
        }
        SF_HIP(P->dBg.alloc(2 * tn));
        SF_HIP(P->dDg.alloc(2 * tn));
        SF_HIP(P->dBg32.alloc(2 * tn));
        SF_HIP(P->dDg32.alloc(2 * tn));
        SF_HIP(hipMemcpy(P->dBg.p, b64.data(), 2 * tn * 8, hipMemcpyHostToDevice));
        SF_HIP(hipMemcpy(P->dDg.p, d64.data(), 2 * tn * 8, hipMemcpyHostToDevice));
        SF_HIP(hipMemcpy(P->dBg32.p, b32.data(), 2 * tn * 4, hipMemcpyHostToDevice));
        SF_HIP(hipMemcpy(P->dDg32.p, d32.data(), 2 * tn * 4, hipMemcpyHostToDevice));
      }
      SF_HIP(hipMemsetAsync(P->minv.p, 0, P->nfree * sizeof(float), s));
      SF_HIP(launch_mult_inverse(P->cells.p, P->ncell * P->nloc, P->minv.p, P->nfree, s));
      SF_HIP(launch_fill<double>(d.p, P->nfree, 0.0, s));
      hipError_t e = launch_diagonal(P->dim, n, q, P->active_G(), tb.p, td.p, P->cells.p,
                                     P->cell_block.p, P->bcell.p, P->ncell, d.p, s);
      if (e == hipErrorNotSupported)
        rc = fail(SF_ERR_UNSUPPORTED, "no diagonal kernel for n=%d q=%d", n, q);
      else if (e != hipSuccess)
        rc = fail(SF_ERR_HIP, "diagonal kernel: %s", hipGetErrorString(e));
    } else {
      rc = compute_diagonal(P, d.p, s);
    }
    if (rc) return rc;
    hipError_t e = hipMemsetAsync(P->flag.p, 0, sizeof(int), s);
    if (e == hipSuccess) e = launch_inv_diag(d.p, P->dinv.p, P->nfree, P->flag.p, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    SF_HIP(e);
    int bad = 0;
    rc = read_flag(P, &bad);
    if (rc) return rc;
    if (bad) return fail(SF_ERR_NONFINITE, "preconditioner diagonal not positive and finite");
    P->pc_valid = true;
  }
  if (gll) *gll = P->pc_kind == SF_PC_GLL;
  *dinv = P->dinv.p;
  return SF_OK;
}

// z = M^{-1} r by the GLL passes (w: scratch), skipped once sc->done.
template <typename V>
int pc_gll(sf_plan* P, const V* r, V* z, V* w, const CGScalars* sc, hipStream_t s) {
  hipError_t e = launch_pc_gll<V, V>(P->n, r, z, w, P->nfree, P->minv.p, P->dinv.p, P->dT1.p,
                                     P->dT1t.p, P->kargs, P->eb, P->max_nu, sc, s);
  if (e == hipErrorNotSupported)
    return fail(SF_ERR_UNSUPPORTED, "no Gauss-Lobatto preconditioner kernel for n=%d "
                "(or its blocks exceed the workgroup's LDS)", P->n);
  SF_HIP(e);
  return SF_OK;
}

// One Gauss-Lobatto pass over the free-dof vectors (launch_gll_pass).
int gll_pass(sf_plan* P, bool trans, const double* in, const float* scale, double* out,
             const double* M, hipStream_t s) {
  hipError_t e = launch_gll_pass(P->n, trans, in, scale, out, M, P->nfree, P->kargs, P->eb,
                                 P->max_nu, s);
  if (e == hipErrorNotSupported)
    return fail(SF_ERR_UNSUPPORTED, "no Gauss-Lobatto pass for n=%d "
                "(or its blocks exceed the workgroup's LDS)", P->n);
  SF_HIP(e);
  return SF_OK;
}

// The apply's 1D tables exchanged with the Gauss-Lobatto basis's for one
// scope: meanwhile every apply (FP64 and reduced precision) computes
// T^T A T.
struct GllTables {
  sf_plan* P;
  explicit GllTables(sf_plan* p) : P(p) { flip(); }
  ~GllTables() { flip(); }
  void flip() {
    std::swap(P->dB, P->dBg);
    std::swap(P->dD, P->dDg);
    std::swap(P->dB32, P->dBg32);
    std::swap(P->dD32, P->dDg32);
  }
};

int read_scalars(sf_plan* P, CGScalars* h, hipStream_t s) {
  SF_HIP(hipMemcpyAsync(h, P->scal.p, sizeof(CGScalars), hipMemcpyDeviceToHost, s));
  SF_HIP(hipStreamSynchronize(s));
  return SF_OK;
}

int dot_host(sf_plan* P, const double* a, const double* b, double* out,
             hipStream_t s) {
  Reducer red{P->red_part.p, P->red_counter.p};
  SF_HIP(launch_dot<double>(a, b, P->nfree, red, P->dscal.p, s));
  SF_HIP(hipMemcpyAsync(out, P->dscal.p, sizeof(double), hipMemcpyDeviceToHost, s));
  SF_HIP(hipStreamSynchronize(s));
  return SF_OK;
}

double now_s();

// Solver progress on stderr when SFHIP_PROGRESS=<seconds> is set: one line
// per interval from the host convergence checks (iterations, residual).
struct Progress {
  double every = -1.0, last = 0.0;
  const char* what;
  explicit Progress(const char* w) : what(w) {
    const char* e = getenv("SFHIP_PROGRESS");
    if (e && *e) every = atof(e);
    last = now_s();
  }
  void tick(int iters, double rel) {
    if (every < 0.0) return;
    const double t = now_s();
    if (t - last < every) return;
    last = t;
    fprintf(stderr, "sfhip %s: %d iterations, rel. residual %.3e\n", what, iters, rel);
    fflush(stderr);
  }
};

int write_scalars(sf_plan* P, const CGScalars* h, hipStream_t s);

// Called when the recurrence has met its target (res2 = its r.r): sets
// *stop_rr to a lower target to continue the same recurrence, or to 0 to
// accept.
using CgCheck = std::function<int(double res2, double* stop_rr)>;

// Plain CG on the FP64 operator, device vectors.  Chunks of `chunk`
// iterations are queued between host checks of the device scalars; once
// converged the remaining queued updates are no-ops.  check (optional):
// see CgCheck; continuing costs one direction update (the converging step
// moved x and r but not p).
int cg_f64(sf_plan* P, const double* b, double* x, double rtol, int maxit,
           hipStream_t s, int* iters_out, int* nonfinite_out, const float* dinv,
           const CgCheck* check = nullptr) {
  Reducer red{P->red_part.p, P->red_counter.p};
  double bb = 0.0;
  int rc = dot_host(P, b, b, &bb, s);
  if (rc) return rc;
  const double stop_rr = rtol * rtol * bb;
  rc = do_apply_f64(P, P->w_t.p, x, s);
  if (rc) return rc;
  SF_HIP(launch_cg_start<double>(b, P->w_t.p, P->w_r.p, P->w_p.p, P->w_q.p,
                                 P->nfree, red, P->scal.p, stop_rr, s, dinv));
  CGScalars h{};
  rc = read_scalars(P, &h, s);
  if (rc) return rc;
  Progress prog(dinv ? "cg (preconditioned)" : "cg");
  const int chunk = 8;
  // 3D fused-geometry apply: p.Ap from the apply (alpha_parts / resid / xdir)
  const bool fused = P->npart > 0 && P->w_part.p && fused_f64(P);
  for (;;) {
    if (h.done) {
      if (!check || h.nonfinite) break;
      double stop = 0.0;
      rc = (*check)(h.res2, &stop);
      if (rc) return rc;
      if (!(stop > 0.0)) break;
      // continue: p = z + beta p and q = 0, which the converging step left
      h.done = 0;
      h.xpend = 0;
      h.stop_rr = stop;
      rc = write_scalars(P, &h, s);
      if (rc) return rc;
      SF_HIP(launch_cg_dir<double>(P->w_p.p, P->w_r.p, P->w_q.p, P->nfree, P->scal.p, s, dinv));
    }
    if (h.iters >= maxit) break;
    prog.tick(h.iters, bb > 0 ? sqrt(h.res2 / bb) : sqrt(h.res2));
    MarkerRange mark("sfhip cg chunk (8 iterations)");
    const int todo = (maxit - h.iters) < chunk ? (maxit - h.iters) : chunk;
    for (int k = 0; k < todo; ++k) {
      if (fused) {
        rc = do_apply_f64(P, P->w_q.p, P->w_p.p, s, 0, true, P->w_part.p);
        if (rc) return rc;
        SF_HIP(launch_cg_alpha_parts(P->w_part.p, P->npart, red, P->scal.p, s));
        SF_HIP(launch_cg_resid<double>(P->w_r.p, P->w_q.p, P->nfree, red, P->scal.p, s, dinv));
        SF_HIP(launch_cg_xdir<double>(x, P->w_p.p, P->w_r.p, P->w_q.p, P->nfree, P->scal.p, s,
                                      dinv));
        continue;
      }
      rc = do_apply_f64(P, P->w_q.p, P->w_p.p, s, 0, true);
      if (rc) return rc;
      SF_HIP(launch_cg_alpha<double>(P->w_p.p, P->w_q.p, P->nfree, red, P->scal.p, s));
      SF_HIP(launch_cg_update<double>(x, P->w_r.p, P->w_p.p, P->w_q.p, P->nfree, red, P->scal.p,
                                      s, dinv));
      SF_HIP(launch_cg_dir<double>(P->w_p.p, P->w_r.p, P->w_q.p, P->nfree, P->scal.p, s, dinv));
    }
    rc = read_scalars(P, &h, s);
    if (rc) return rc;
  }
  *iters_out = h.iters;
  *nonfinite_out = h.nonfinite;
  return SF_OK;
}

// Inner CG for IR: solve A32 d = b32 (FP32 vectors, FP64 scalars) with
// b32 = w_r32 of unit norm, d = 0 initially, to ||r|| <= inner_rtol or
// maxit iterations.
int cg_f32_inner(sf_plan* P, int precision, double inner_rtol, int maxit,
                 hipStream_t s, int* iters_out, int* nonfinite_out, const float* dinv) {
  Reducer red{P->red_part.p, P->red_counter.p};
  const float* b32 = P->w_r32.p;
  float* r = P->w_s32.p;
  float* d = P->w_d32.p;
  float* p = P->w_p32.p;
  float* q = P->w_q32.p;
  SF_HIP(launch_fill<float>(d, P->nfree, 0.0f, s));
  SF_HIP(launch_cg_start<float>(b32, d, r, p, q, P->nfree, red, P->scal.p,
                                inner_rtol * inner_rtol, s, dinv));
  CGScalars h{};
  int rc = read_scalars(P, &h, s);
  if (rc) return rc;
  Progress prog("refinement inner cg");
  const int chunk = 8;
  while (!h.done && h.iters < maxit) {
    prog.tick(h.iters, sqrt(h.res2));
    const int todo = (maxit - h.iters) < chunk ? (maxit - h.iters) : chunk;
    for (int k = 0; k < todo; ++k) {
      rc = do_apply_f32(P, q, p, precision, s, true);
      if (rc) return rc;
      SF_HIP(launch_cg_alpha<float>(p, q, P->nfree, red, P->scal.p, s));
      SF_HIP(launch_cg_update<float>(d, r, p, q, P->nfree, red, P->scal.p, s, dinv));
      SF_HIP(launch_cg_dir<float>(p, r, q, P->nfree, P->scal.p, s, dinv));
    }
    rc = read_scalars(P, &h, s);
    if (rc) return rc;
  }
  *iters_out = h.iters;
  *nonfinite_out = h.nonfinite;
  return SF_OK;
}

int write_scalars(sf_plan* P, const CGScalars* h, hipStream_t s) {
  SF_HIP(hipMemcpyAsync(P->scal.p, h, sizeof(CGScalars), hipMemcpyHostToDevice, s));
  SF_HIP(hipStreamSynchronize(s));
  return SF_OK;
}

int dot_host_f32(sf_plan* P, const float* a, const float* b, double* out,
                 hipStream_t s) {
  Reducer red{P->red_part.p, P->red_counter.p};
  SF_HIP(launch_dot<float>(a, b, P->nfree, red, P->dscal.p, s));
  SF_HIP(hipMemcpyAsync(out, P->dscal.p, sizeof(double), hipMemcpyDeviceToHost, s));
  SF_HIP(hipStreamSynchronize(s));
  return SF_OK;
}

// FP64 residual w_r = b - A x and its norm.
int true_residual(sf_plan* P, const double* b, const double* x, double* rn,
                  hipStream_t s) {
  int rc = do_apply_f64(P, P->w_t.p, x, s);
  if (rc) return rc;
  SF_HIP(launch_residual(b, P->w_t.p, P->w_r.p, P->nfree, s));
  double rr = 0.0;
  rc = dot_host(P, P->w_r.p, P->w_r.p, &rr, s);
  if (rc) return rc;
  *rn = sqrt(rr);
  return SF_OK;
}

// Mixed-precision CG with reliable updates (residual replacement, van der
// Vorst & Ye 2000; the mixed-precision form of Clark et al. 2010).  One CG
// recurrence runs in FP32 on the reduced-precision operator, in units of
// the last true residual norm rn (so ||r|| starts at 1).  When the
// recurred residual has dropped below max(delta, rtol ||b|| / rn), x (FP64)
// absorbs rn * xl, the residual is replaced by the FP64 b - A x, and the
// recurrence continues with the same search direction:
//   p <- r_true + beta' p,  beta' = ||r_true||^2 / ||r_{k-1}||^2,
// rescaled to the new units.  Krylov information survives the update,
// unlike a restarted refinement.
int cg_reliable(sf_plan* P, int precision, const double* b, double* x,
                double rtol, int maxit, double delta, hipStream_t s,
                int* iters_out, int* updates_out, int* nonfinite_out,
                double* rel_out, const float* dinv) {
  Reducer red{P->red_part.p, P->red_counter.p};
  float* r = P->w_s32.p;
  float* xl = P->w_d32.p;
  float* p = P->w_p32.p;
  float* q = P->w_q32.p;
  double bb = 0.0;
  int rc = dot_host(P, b, b, &bb, s);
  if (rc) return rc;
  const double bnorm = sqrt(bb), target = rtol * bnorm;
  double rn = 0.0;
  rc = true_residual(P, b, x, &rn, s);
  if (rc) return rc;
  *iters_out = 0;
  *updates_out = 0;
  *nonfinite_out = !isfinite(rn);
  *rel_out = bnorm > 0 ? rn / bnorm : rn;
  if (!isfinite(rn) || rn <= target || rn == 0.0) return SF_OK;
  auto threshold = [&](double rnow) {
    const double t = target / rnow;
    return t > delta ? t : delta;
  };
  SF_HIP(launch_scale_to_f32(P->w_r.p, 1.0 / rn, P->w_r32.p, P->nfree, s));
  SF_HIP(launch_fill<float>(xl, P->nfree, 0.0f, s));
  const double t0 = threshold(rn);
  SF_HIP(launch_cg_start<float>(P->w_r32.p, xl, r, p, q, P->nfree, red, P->scal.p,
                                t0 * t0, s, dinv));
  CGScalars h{};
  rc = read_scalars(P, &h, s);
  if (rc) return rc;
  const int chunk = 8;
  Progress prog("mixed cg");
  // 3D FP32 fused-geometry operator: p.Ap from the apply, then the FP32
  // forms of the FP64 fused step (resid, xdir)
  const bool fused = P->npart > 0 && P->w_part.p && precision == SF_PREC_F32 && fused_f32(P);
  for (;;) {
    while (!h.done && h.iters < maxit) {
      prog.tick(h.iters, (bnorm > 0 ? rn / bnorm : rn) * sqrt(h.res2));
      const int todo = (maxit - h.iters) < chunk ? (maxit - h.iters) : chunk;
      for (int k = 0; k < todo; ++k) {
        if (fused) {
          rc = do_apply_f32(P, q, p, precision, s, true, P->w_part.p);
          if (rc) return rc;
          SF_HIP(launch_cg_alpha_parts(P->w_part.p, P->npart, red, P->scal.p, s));
          SF_HIP(launch_cg_resid<float>(r, q, P->nfree, red, P->scal.p, s, dinv));
          SF_HIP(launch_cg_xdir<float>(xl, p, r, q, P->nfree, P->scal.p, s, dinv));
          continue;
        }
        rc = do_apply_f32(P, q, p, precision, s, true);
        if (rc) return rc;
        SF_HIP(launch_cg_alpha<float>(p, q, P->nfree, red, P->scal.p, s));
        SF_HIP(launch_cg_update<float>(xl, r, p, q, P->nfree, red, P->scal.p, s, dinv));
        SF_HIP(launch_cg_dir<float>(p, r, q, P->nfree, P->scal.p, s, dinv));
      }
      rc = read_scalars(P, &h, s);
      if (rc) return rc;
    }
    *iters_out = h.iters;
    if (h.nonfinite) {
      *nonfinite_out = 1;
      return SF_OK;
    }
    // reliable update: x += rn xl, r <- b - A x (FP64)
    SF_HIP(launch_axpy_f32_to_f64(x, xl, rn, P->nfree, s));
    double rn_new = 0.0;
    rc = true_residual(P, b, x, &rn_new, s);
    if (rc) return rc;
    *updates_out += 1;
    *rel_out = bnorm > 0 ? rn_new / bnorm : rn_new;
    if (!isfinite(rn_new)) {
      *nonfinite_out = 1;
      return SF_OK;
    }
    if (rn_new <= target || rn_new == 0.0 || h.iters >= maxit) return SF_OK;
    // continue the recurrence in units of rn_new; the queued pieces after
    // `done` were no-ops, so p is the last direction and q is stale.  In
    // the new units p <- z + c p with c = (rn_new / rn) rz / rz_{k-1}
    // (rz = r.z of the replaced residual in the new units, 1 without a
    // preconditioner; rz_{k-1} = rr_new / beta in the old units).
    SF_HIP(launch_scale_to_f32(P->w_r.p, 1.0 / rn_new, r, P->nfree, s));
    double rr = 0.0, rz = 0.0;
    rc = dot_host_f32(P, r, r, &rr, s);
    if (rc) return rc;
    rz = rr;
    if (dinv) {
      SF_HIP(launch_wdot<float>(r, dinv, P->nfree, red, P->dscal.p, s));
      SF_HIP(hipMemcpyAsync(&rz, P->dscal.p, sizeof(double), hipMemcpyDeviceToHost, s));
      SF_HIP(hipStreamSynchronize(s));
    }
    const double c = (h.rr_new > 0.0) ? (rn_new / rn) * rz * h.beta / h.rr_new : 0.0;
    SF_HIP(launch_dir_coef(p, r, q, (float)c, P->nfree, s, dinv));
    SF_HIP(launch_fill<float>(xl, P->nfree, 0.0f, s));
    rn = rn_new;
    const double t = threshold(rn);
    h.res2 = rr;
    h.rr = rz;
    h.stop_rr = t * t;
    h.done = 0;
    h.xpend = 0;
    rc = write_scalars(P, &h, s);
    if (rc) return rc;
  }
}

// Mixed CG with the Krylov recurrence in FP64 and only the operator in
// reduced precision: q32 = A32 p32 with p32 the FP32 copy of the FP64
// direction, every vector update and dot in FP64.  The recurred residual
// drifts from b - A x by the operator's rounding, so it is replaced by the
// FP64 b - A x whenever it has dropped by delta (search direction kept,
// beta' = ||r_true||^2 / ||r_{k-1}||^2).  Fewer iterations than the FP32
// recurrence (which loses orthogonality faster), FP64 vector traffic.
// p32only (3D fused FP32 operator): the direction is held only in FP32 (x
// and r stay FP64), so the direction update moves 36 bytes per dof, not 48.
int cg_reliable_f64rec(sf_plan* P, int precision, const double* b, double* x,
                       double rtol, int maxit, double delta, hipStream_t s,
                       int* iters_out, int* updates_out, int* nonfinite_out,
                       double* rel_out, bool p32only, const float* dinv) {
  Reducer red{P->red_part.p, P->red_counter.p};
  double* r = P->w_r.p;
  double* p = P->w_p.p;
  float* p32 = P->w_p32.p;
  float* q32 = P->w_q32.p;
  double bb = 0.0;
  int rc = dot_host(P, b, b, &bb, s);
  if (rc) return rc;
  const double bnorm = sqrt(bb), target = rtol * bnorm;
  double rn = 0.0;
  rc = true_residual(P, b, x, &rn, s);  // w_t = A x, w_r = b - A x
  if (rc) return rc;
  *iters_out = 0;
  *updates_out = 0;
  *nonfinite_out = !isfinite(rn);
  *rel_out = bnorm > 0 ? rn / bnorm : rn;
  if (!isfinite(rn) || rn <= target || rn == 0.0) return SF_OK;
  auto threshold = [&](double rnow) { return (delta * rnow > target) ? delta * rnow : target; };
  const double t0 = threshold(rn);
  SF_HIP(launch_cg_start<double>(b, P->w_t.p, r, p, P->w_q.p, P->nfree, red, P->scal.p,
                                 t0 * t0, s, dinv));

  SF_HIP(launch_to_f32(p, P->nfree, p32, s));
  SF_HIP(launch_fill<float>(q32, P->nfree, 0.0f, s));
  CGScalars h{};
  rc = read_scalars(P, &h, s);
  if (rc) return rc;
  const int chunk = 8;
  Progress prog("mixed cg (fp64 recurrence)");
  // 3D FP32 fused-geometry operator: p.Ap from the apply
  const bool fused = P->npart > 0 && P->w_part.p && precision == SF_PREC_F32 &&
                     fused_f32(P);
  if (!fused) p32only = false;
  for (;;) {
    while (!h.done && h.iters < maxit) {
      prog.tick(h.iters, bnorm > 0 ? sqrt(h.res2) / bnorm : sqrt(h.res2));
      const int todo = (maxit - h.iters) < chunk ? (maxit - h.iters) : chunk;
      for (int k = 0; k < todo; ++k) {
        if (fused) {
          rc = do_apply_f32(P, q32, p32, precision, s, true, P->w_part.p);
          if (rc) return rc;
          SF_HIP(launch_cg_alpha_parts(P->w_part.p, P->npart, red, P->scal.p, s));
          SF_HIP(launch_cg_resid_mixed(r, q32, P->nfree, red, P->scal.p, s, dinv));
          if (p32only)
            SF_HIP(launch_cg_xdir_p32(x, p32, r, q32, P->nfree, P->scal.p, s, dinv));
          else
            SF_HIP(launch_cg_xdir_mixed(x, p, p32, r, q32, P->nfree, P->scal.p, s, dinv));
          continue;
        }
        rc = do_apply_f32(P, q32, p32, precision, s, true);
        if (rc) return rc;
        SF_HIP(launch_cg_alpha_mixed(p, q32, P->nfree, red, P->scal.p, s));
        SF_HIP(launch_cg_update_mixed(x, r, p, q32, P->nfree, red, P->scal.p, s, dinv));
        SF_HIP(launch_cg_dir_mixed(p, p32, r, q32, P->nfree, P->scal.p, s, dinv));
      }
      rc = read_scalars(P, &h, s);
      if (rc) return rc;
    }
    *iters_out = h.iters;
    if (h.nonfinite) {
      *nonfinite_out = 1;
      return SF_OK;
    }
    // reliable update: r <- b - A x in FP64
    double rn_new = 0.0;
    rc = true_residual(P, b, x, &rn_new, s);
    if (rc) return rc;
    *updates_out += 1;
    *rel_out = bnorm > 0 ? rn_new / bnorm : rn_new;
    if (!isfinite(rn_new)) {
      *nonfinite_out = 1;
      return SF_OK;
    }
    if (rn_new <= target || rn_new == 0.0 || h.iters >= maxit) return SF_OK;
    // p <- z + c p, c = r.z / rz_{k-1} (rz_{k-1} = rr_new / beta)
    double rz = rn_new * rn_new;
    if (dinv) {
      SF_HIP(launch_wdot<double>(r, dinv, P->nfree, red, P->dscal.p, s));
      SF_HIP(hipMemcpyAsync(&rz, P->dscal.p, sizeof(double), hipMemcpyDeviceToHost, s));
      SF_HIP(hipStreamSynchronize(s));
    }
    const double c = (h.rr_new > 0.0) ? rz * h.beta / h.rr_new : 0.0;
    if (p32only)
      SF_HIP(launch_dir_coef_p32(p32, r, q32, c, P->nfree, s, dinv));
    else
      SF_HIP(launch_dir_coef_mixed(p, p32, r, q32, c, P->nfree, s, dinv));
    const double t = threshold(rn_new);
    h.res2 = rn_new * rn_new;
    h.rr = rz;
    h.stop_rr = t * t;
    h.done = 0;
    h.xpend = 0;
    rc = write_scalars(P, &h, s);
    if (rc) return rc;
  }
}

double now_s() {
  using namespace std::chrono;
  return duration<double>(steady_clock::now().time_since_epoch()).count();
}

// A solve with the Gauss-Lobatto preconditioner.  PCG on A with
// M^{-1} = T diag(T^T A T)^{-1} T^T is, in y = T^{-1} x, Jacobi PCG on
// A_g = T^T A T with right-hand side T^T b, and A_g is the same apply on
// the GLL tables: run(bg, y, rtol_g, maxit, &its, &nonfinite) runs one of
// the solvers on it, with no pass per iteration.  y starts at T^{-1} x and
// x = T y after.  The solver's own target rtol_g (relative to ||T^T b||)
// starts at rtol ||b|| / ||T^T b|| and tightens while the true
// ||b - A x|| / ||b|| stays above rtol (*restarts_out: the runs after the
// first).  Budget and *iters_out in the units of run's count.
template <typename Run>
int solve_gll(sf_plan* P, const double* b, double* x, double rtol, int maxit, hipStream_t s,
              Run&& run, int* iters_out, double* rel_out, int* nonfinite_out,
              int* restarts_out) {
  double* bg = P->g_b.p;
  double* y = P->g_y.p;
  *iters_out = 0;
  *nonfinite_out = 0;
  *restarts_out = 0;
  double bb = 0.0, gg = 0.0;
  int rc = dot_host(P, b, b, &bb, s);
  if (rc) return rc;
  rc = gll_pass(P, true, b, P->minv.p, bg, P->dT1t.p, s);
  if (rc) return rc;
  rc = gll_pass(P, false, x, nullptr, y, P->dT1i.p, s);
  if (rc) return rc;
  rc = dot_host(P, bg, bg, &gg, s);
  if (rc) return rc;
  const double bnorm = sqrt(bb);
  double rtol_g = gg > 0 ? rtol * bnorm / sqrt(gg) : rtol;
  double rel = 0.0;
  for (;;) {
    int it = 0, nonfinite = 0;
    {
      GllTables swap(P);
      rc = run(bg, y, rtol_g, maxit - *iters_out, &it, &nonfinite);
      if (rc) return rc;
    }
    *iters_out += it;
    rc = gll_pass(P, false, y, nullptr, x, P->dT1.p, s);
    if (rc) return rc;
    double rn = 0.0;
    rc = true_residual(P, b, x, &rn, s);
    if (rc) return rc;
    rel = bnorm > 0 ? rn / bnorm : rn;
    if (nonfinite || !isfinite(rel)) {
      *nonfinite_out = 1;
      break;
    }
    if (!(rel > rtol) || *iters_out >= maxit || it == 0) break;
    const double f = 0.8 * rtol / rel;
    rtol_g *= f > 1e-3 ? f : 1e-3;
    ++*restarts_out;
  }
  *rel_out = rel;
  return SF_OK;
}

}  // namespace

static int finish_geometry(sf_plan* P);

namespace {

// FP64 CG on the plan's workspace (already in place): the body shared by
// sf_cg and sf_cg_noalloc.
int cg_solve(sf_plan* P, const double* b, double* x, double rtol, int maxit,
             sf_solve_stats* stats, hipStream_t s) {
  MarkerRange mark("sfhip cg");
  const float* dinv = nullptr;
  bool gll = false;
  int rc = ensure_pc(P, s, &dinv, &gll);
  if (rc) return rc;
  SF_HIP(hipStreamSynchronize(s));
  const double t0 = now_s();
  int iters = 0, nonfinite = 0, restarts = 0;
  double bb = 0.0, rel = 0.0;
  if (gll) {
    // CG in the Gauss-Lobatto basis (solve_gll); each time the recurrence
    // meets its target the true residual of x = T y is checked with the
    // apply's own tables, and the same recurrence continues to a lower
    // target while it is above rtol (no restart, no Krylov space lost)
    double* bg = P->g_b.p;
    double* y = P->g_y.p;
    double gg = 0.0;
    rc = dot_host(P, b, b, &bb, s);
    if (rc) return rc;
    rc = gll_pass(P, true, b, P->minv.p, bg, P->dT1t.p, s);
    if (rc) return rc;
    rc = gll_pass(P, false, x, nullptr, y, P->dT1i.p, s);
    if (rc) return rc;
    rc = dot_host(P, bg, bg, &gg, s);
    if (rc) return rc;
    const double bnorm = sqrt(bb);
    auto true_rel = [&](double* out) -> int {
      int r = gll_pass(P, false, y, nullptr, x, P->dT1.p, s);
      if (r) return r;
      r = do_apply_f64(P, P->w_t.p, x, s);
      if (r) return r;
      Reducer red{P->red_part.p, P->red_counter.p};
      SF_HIP(launch_diff_norm2(b, P->w_t.p, P->nfree, red, P->dscal.p, s));
      double rr = 0.0;
      SF_HIP(hipMemcpyAsync(&rr, P->dscal.p, sizeof(double), hipMemcpyDeviceToHost, s));
      SF_HIP(hipStreamSynchronize(s));
      *out = bnorm > 0 ? sqrt(rr) / bnorm : sqrt(rr);
      return SF_OK;
    };
    bool fresh = false;  // rel is that of the current y
    CgCheck check = [&](double res2, double* stop_rr) -> int {
      *stop_rr = 0.0;
      GllTables own(P);  // flipped back: the apply's own tables
      int r = true_rel(&rel);
      if (r) return r;
      fresh = true;
      if (rel > rtol && isfinite(rel)) {
        const double f = 0.8 * rtol / rel;
        const double g = f > 1e-3 ? f : 1e-3;
        *stop_rr = res2 * g * g;
        ++restarts;
        fresh = false;
      }
      return SF_OK;
    };
    {
      GllTables swap(P);
      rc = cg_f64(P, bg, y, gg > 0 ? rtol * bnorm / sqrt(gg) : rtol, maxit, s, &iters,
                  &nonfinite, dinv, &check);
    }
    if (rc) return rc;
    if (!fresh) {  // stopped by maxit or a non-finite value: x = T y still to form
      rc = true_rel(&rel);
      if (rc) return rc;
    }
  } else {
    rc = dot_host(P, b, b, &bb, s);
    if (rc) return rc;
  }
  // The recurred residual drifts from the true one; when the recurrence
  // stops short of rtol in the true residual, restart from x (the restart
  // recomputes r = b - A x).
  while (!gll) {
    int it = 0;
    rc = cg_f64(P, b, x, rtol, maxit - iters, s, &it, &nonfinite, dinv);
    if (rc) return rc;
    iters += it;
    double rn = 0.0;
    rc = true_residual(P, b, x, &rn, s);
    if (rc) return rc;
    rel = bb > 0 ? rn / sqrt(bb) : rn;
    if (nonfinite || !(rel > rtol) || iters >= maxit || it == 0) break;
    ++restarts;
  }
  const double t1 = now_s();
  if (stats) {
    stats->iterations = iters;
    stats->outer_iterations = restarts;
    stats->rel_residual = rel;
    stats->converged = (rel <= rtol) ? 1 : 0;
    stats->seconds = t1 - t0;
  }
  if (nonfinite || !isfinite(rel)) return fail(SF_ERR_NONFINITE, "CG met a non-finite value");
  if (rel > rtol && iters >= maxit)
    return fail(SF_ERR_NOT_CONVERGED, "CG did not reach rtol %g in %d iterations (rel. residual %g)", rtol, maxit, rel);
  return SF_OK;
}

// Iterative refinement: FP64 residuals r = b - A x, inner CG on r / ||r||
// in reduced precision (cg_f32_inner), x += ||r|| d, until ||r|| <= rtol ||b||
// or max_outer inner solves.
int refine(sf_plan* P, const double* b, double* x, double rtol, int max_outer,
           int inner_precision, double inner_rtol, int inner_maxit, hipStream_t s,
           int* inner_total, int* outer_out, int* nonfinite_out, double* rel_out,
           const float* dinv) {
  double bb = 0.0;
  int rc = dot_host(P, b, b, &bb, s);
  if (rc) return rc;
  const double bnorm = sqrt(bb);
  int outer = 0, nonfinite = 0;
  double rel = 0.0;
  MarkerRange mark("sfhip refinement");
  for (;;) {
    MarkerRange step("sfhip refinement outer step");
    double rn = 0.0;
    rc = true_residual(P, b, x, &rn, s);
    if (rc) return rc;
    rel = bnorm > 0 ? rn / bnorm : rn;
    if (!isfinite(rel)) {
      nonfinite = 1;
      break;
    }
    if (rel <= rtol || outer >= max_outer || rn == 0.0) break;
    // inner solve on the unit-norm residual in reduced precision
    SF_HIP(launch_scale_to_f32(P->w_r.p, 1.0 / rn, P->w_r32.p, P->nfree, s));
    int it = 0, nf = 0;
    rc = cg_f32_inner(P, inner_precision, inner_rtol, inner_maxit, s, &it, &nf, dinv);
    if (rc) return rc;
    *inner_total += it;
    ++outer;
    if (nf) {
      nonfinite = 1;
      break;
    }
    SF_HIP(launch_axpy_f32_to_f64(x, P->w_d32.p, rn, P->nfree, s));
  }
  *outer_out = outer;
  *nonfinite_out = nonfinite;
  *rel_out = rel;
  return SF_OK;
}

// Offsets of the workspace pieces inside the scratch array, each on a
// 256-byte boundary.
struct ScratchCarve {
  size_t off = 0;
  size_t take(size_t bytes) {
    off = (off + 255) & ~(size_t)255;
    const size_t at = off;
    off += bytes;
    return at;
  }
};

// A plan buffer pointed at scratch memory for one call, restored on exit
// (the buffer never frees what it points at meanwhile: every alloc() it
// sees asks for the size it already has).
template <typename T>
struct BufSwap {
  DevBuf<T>& b;
  T* p0;
  int64_t n0;
  BufSwap(DevBuf<T>& b_, T* p, int64_t n) : b(b_), p0(b_.p), n0(b_.n) {
    b.p = p;
    b.n = n;
  }
  ~BufSwap() {
    b.p = p0;
    b.n = n0;
  }
};

}  // namespace

// The no-allocation CG's workspace: one fixed array per GPU, held until the
// process ends.  Host code never uses this name as an address (on the host
// it is a stand-in): scratch_base() asks the runtime for each device's
// copy once and keeps it.
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

std::mutex g_scratch_mu;
char* g_scratch_base[kMaxDevices] = {};

// The device address of `scratch` on device dev (current), looked up once.
int scratch_base(int dev, char** out) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (dev < 0 || dev >= kMaxDevices) return fail(SF_ERR_ARG, "device %d out of range", dev);
  if (!g_scratch_base[dev]) {
    void* ptr = nullptr;
    hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
      return fail(SF_ERR_HIP, "hipGetSymbolAddress(scratch) on device %d: %s", dev,
                  hipGetErrorString(e));
    g_scratch_base[dev] = (char*)ptr;
  }
  *out = g_scratch_base[dev];
  return SF_OK;
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI

extern "C" {

const char* sf_last_error(void) { return g_err.c_str(); }

const char* sf_version(void) { return "sfhip 0.1 (gfx950)"; }

int sf_tables_1d(int order, int nquad, double* B, double* Dm, double* points,
                 double* weights) {
  SF_TRY
  return build_tables(order, nquad, B, Dm, points, weights);
  SF_CATCH
}

int sf_gauss_points_for_degree(int degree) {
  if (degree < 0) return 1;
  return (degree + 2) / 2;  // ceil((degree + 1) / 2)
}

int64_t sf_flops_model(int64_t n_cells, int dim, int nquad, int n1d,
                       int count_only_fma_ops, int structured_mesh_opt) {
  // restates count_manual_flops_poisson_matrix_free, src/Utils.jl:33-72
  const int D = dim;
  int64_t b2 = 0;
  for (int d = 1; d <= D; ++d) {
    int64_t t = D;
    for (int k = 1; k <= d; ++k) t *= nquad;
    for (int k = d; k <= D; ++k) t *= n1d;
    b2 += t;
  }
  if (!count_only_fma_ops) b2 *= 2;
  int64_t DD = 1;
  for (int k = 0; k < D; ++k) DD *= D;
  int64_t b3 = 0;
  if (!structured_mesh_opt) {
    b3 = (D - 1) * DD + (int64_t)D * D;
    if (!count_only_fma_ops) b3 = 2 * b3 + DD + 2 * D;
  } else {
    if (!count_only_fma_ops) b3 = D;
  }
  int64_t qd = 1;
  for (int k = 0; k < D; ++k) qd *= nquad;
  b3 *= qd;
  const int64_t b4 = b2;
  return n_cells * (b2 + b3 + b4);
}

int sf_plan_create(sf_plan** out, int dim, int order, int nquad,
                   int64_t n_cells, int64_t n_free, int flags) {
  SF_TRY
  if (!out) return fail(SF_ERR_ARG, "null output pointer");
  *out = nullptr;
  if (dim != 2 && dim != 3) return fail(SF_ERR_ARG, "dim must be 2 or 3, got %d", dim);
  if (order < 1 || order > kMaxN - 1) return fail(SF_ERR_ARG, "order %d outside 1..%d", order, kMaxN - 1);
  if (nquad < 1) return fail(SF_ERR_ARG, "nquad must be >= 1");
  if (n_cells < 0 || n_free < 0) return fail(SF_ERR_ARG, "negative sizes");
  if (n_free > 0x7fffffffLL) return fail(SF_ERR_ARG, "n_free above int32 range");
  if (flags != 0) return fail(SF_ERR_ARG, "flags must be 0");
  int nt = 0, eb = 0;
  apply_shape_f64(dim, order + 1, nquad, &nt, &eb);
  if (nt == 0)
    return fail(SF_ERR_UNSUPPORTED, "no kernel for dim=%d n=%d q=%d (q must be in n-1..n+3)", dim, order + 1, nquad);
  sf_plan* P = new sf_plan();
  P->dim = dim;
  P->order = order;
  P->n = order + 1;
  P->q = nquad;
  P->ng = dim * (dim + 1) / 2;
  P->ncell = n_cells;
  P->nfree = n_free;
  P->nloc = dim == 3 ? (int64_t)P->n * P->n * P->n : (int64_t)P->n * P->n;
  P->nqc = dim == 3 ? (int64_t)P->q * P->q * P->q : (int64_t)P->q * P->q;
  if (hipGetDevice(&P->device) != hipSuccess) {
    delete P;
    return fail(SF_ERR_HIP, "no HIP device");
  }
  P->hB.resize(P->q * P->n);
  P->hD.resize(P->q * P->n);
  P->hW.resize(P->q);
  P->hPts.resize(P->q);
  int rc = build_tables(order, nquad, P->hB.data(), P->hD.data(), P->hPts.data(), P->hW.data());
  if (rc == SF_OK) rc = upload_tables(P);
  if (rc == SF_OK) rc = clear_flag(P);
  if (rc != SF_OK) {
    P->release_all();
    delete P;
    return rc;
  }
  *out = P;
  return SF_OK;
  SF_CATCH
}

int sf_plan_destroy(sf_plan* plan) {
  SF_TRY
  if (!plan) return SF_OK;
  {
    DeviceGuard g(plan->device);
    plan->release_all();
  }
  delete plan;
  return SF_OK;
  SF_CATCH
}

int sf_plan_get_info(const sf_plan* P, sf_plan_info* info) {
  SF_TRY
  if (check_plan(P)) return SF_ERR_ARG;
  if (!info) return fail(SF_ERR_ARG, "null info");
  info->dim = P->dim;
  info->order = P->order;
  info->n1d = P->n;
  info->q1d = P->q;
  info->n_cells = P->ncell;
  info->n_free = P->nfree;
  info->n_qp = P->ncell * P->nqc;
  info->has_cells = P->has_cells;
  info->has_geometry = P->has_geom;
  info->has_kappa = P->has_kappa;
  apply_shape_f64(P->dim, P->n, P->q, &info->threads_per_block, &info->cells_per_block);
  info->device_bytes = P->device_bytes();
  info->n_blocks = P->nblk;
  info->launches_per_apply = 1;
  info->cells_reordered = P->perm.empty() ? 0 : 1;
  return SF_OK;
  SF_CATCH
}

int sf_plan_set_timing(sf_plan* P, int enable) {
  SF_TRY
  if (check_plan(P)) return SF_ERR_ARG;
  P->timing = enable != 0;
  if (P->timing) P->nrec = 0;  // a new record; stopping keeps what was recorded
  return SF_OK;
  SF_CATCH
}

int sf_plan_read_timing(sf_plan* P, double* ms, int max, int* count) {
  SF_TRY
  if (check_plan(P)) return SF_ERR_ARG;
  if (!count || (max > 0 && !ms)) return fail(SF_ERR_ARG, "null output");
  DeviceGuard g(P->device);
  const int n = (int)std::min<size_t>(P->nrec, (size_t)std::max(max, 0));
  for (int i = 0; i < n; ++i) {
    SF_HIP(hipEventSynchronize(P->ev[2 * i + 1]));
    float t = 0.f;
    SF_HIP(hipEventElapsedTime(&t, P->ev[2 * i], P->ev[2 * i + 1]));
    ms[i] = (double)t;
  }
  *count = (int)P->nrec;
  P->nrec = 0;
  return SF_OK;
  SF_CATCH
}

int sf_plan_set_tables(sf_plan* P, const double* B, const double* Dm,
                       const double* weights) {
  SF_TRY
  if (check_plan(P)) return SF_ERR_ARG;
  if (!B || !Dm) return fail(SF_ERR_ARG, "null table");
  DeviceGuard g(P->device);
  const int64_t tn = (int64_t)P->q * P->n;
  for (int64_t i = 0; i < tn; ++i)
    if (!isfinite(B[i]) || !isfinite(Dm[i])) return fail(SF_ERR_ARG, "non-finite table entry");
  if (weights) {
    for (int i = 0; i < P->q; ++i)
      if (!(weights[i] > 0.0) || !isfinite(weights[i]))
        return fail(SF_ERR_ARG, "quadrature weight %d is not positive and finite", i);
    // the geometric factor holds the weights: with the vertices at hand it
    // is recomputed below, a caller-supplied G cannot be
    if (P->has_geom && !P->X.p)
      return fail(SF_ERR_NOT_READY,
                  "new quadrature weights after a geometry the plan cannot recompute: "
                  "set the tables before the geometry");
  }
  SF_HIP(hipDeviceSynchronize());  // kernels in flight may still read the tables
  P->hB.assign(B, B + tn);
  P->hD.assign(Dm, Dm + tn);
  if (weights) P->hW.assign(weights, weights + P->q);
  P->pc_valid = false;
  int rc = upload_tables(P);
  if (rc || !weights || !P->has_geom) return rc;
  // G, kappa * G and every reduced-precision copy derive from the weights
  P->X32.release();
  P->Pts32.release();
  P->W32.release();
  rc = clear_flag(P);
  if (rc) return rc;
  SF_HIP(launch_geometry(P->dim, P->X.p, P->dPts.p, P->dW.p, P->q, P->ncell, P->cell_block.p,
                         P->bcell.p, P->G.p, P->flag.p, nullptr));
  SF_HIP(hipDeviceSynchronize());
  return finish_geometry(P);
  SF_CATCH
}

int sf_plan_set_cells(sf_plan* P, const int32_t* cell_dofs, int index_base,
                      const int32_t* local_to_tensor) {
  SF_TRY
  if (check_plan(P)) return SF_ERR_ARG;
  if (!cell_dofs && P->ncell > 0) return fail(SF_ERR_ARG, "null cell_dofs");
  if (index_base != 0 && index_base != 1) return fail(SF_ERR_ARG, "index_base must be 0 or 1");
  DeviceGuard g(P->device);
  const int64_t nl = P->nloc;
  std::vector<int32_t> perm(nl);
  if (local_to_tensor) {
    std::vector<char> seen(nl, 0);
    for (int64_t k = 0; k < nl; ++k) {
      const int32_t t = local_to_tensor[k];
      if (t < 0 || t >= nl || seen[t])
        return fail(SF_ERR_ARG, "local_to_tensor is not a permutation of 0..%lld", (long long)nl - 1);
      seen[t] = 1;
      perm[k] = t;
    }
  } else {
    for (int64_t k = 0; k < nl; ++k) perm[k] = (int32_t)k;
  }
  std::vector<int32_t> h((size_t)(P->ncell * nl));
  const int64_t nf = P->nfree;
  SF_HIP(hipDeviceSynchronize());  // kernels in flight may still read the plan
  for (int64_t c = 0; c < P->ncell; ++c) {
    const int32_t* src = cell_dofs + c * nl;
    int32_t* dst = h.data() + c * nl;
    for (int64_t k = 0; k < nl; ++k) {
      const int64_t id = (int64_t)src[k];
      int64_t v;
      if (index_base == 1) {
        v = id >= 1 ? id - 1 : -1;
      } else {
        v = id >= 0 ? id : -1;
      }
      if (v >= nf)
        return fail(SF_ERR_DOF_RANGE, "cell %lld local dof %lld: id %lld outside the %lld free dofs",
                    (long long)c, (long long)k, (long long)id, (long long)nf);
      dst[perm[k]] = (int32_t)v;
    }
  }
  P->pc_valid = false;
  // 3D: blocks are runs of consecutive cells, so a caller order whose
  // consecutive cells are not neighbours (any numbering but a row-wise one)
  // would give blocks whose cells share no dof: each would gather and add
  // eb n^3 dofs, every one through an atomic.  The plan then keeps its cells
  // in the order built from the connectivity when that order keeps more
  // neighbours together (fewer run-split blocks), and maps per-cell inputs
  // and outputs between the two orders.
  P->perm.clear();
  P->corig.release();
  if (P->dim == 3 && P->ncell > 1) {
    int nt = 0, eb = 0;
    apply_shape_f64(3, P->n, P->q, &nt, &eb);
    const int64_t E = P->ncell;
    const int64_t nb_caller = run_block_count(h, E, nl, eb);
    // (a row-wise order is within a few per cent of the minimum count: the
    // connectivity order is only worth building when it is well above)
    if (eb > 0 && nb_caller > 5 * ((E + eb - 1) / eb) / 4) {
      std::vector<int32_t> ord = connectivity_order(h, E, P->n);
      std::vector<int32_t> hp(h.size());
      for (int64_t i = 0; i < E; ++i)
        std::copy(h.begin() + (int64_t)ord[i] * nl, h.begin() + ((int64_t)ord[i] + 1) * nl,
                  hp.begin() + i * nl);
      const int64_t nb_conn = run_block_count(hp, E, nl, eb);
      if (nb_conn < nb_caller) {
        h.swap(hp);
        P->perm.swap(ord);
        SF_HIP(P->corig.alloc(E));
        SF_HIP(hipMemcpy(P->corig.p, P->perm.data(), E * sizeof(int32_t), hipMemcpyHostToDevice));
      }
    }
  }
  SF_HIP(P->cells.alloc(P->ncell * nl));
  if (!h.empty())
    SF_HIP(hipMemcpy(P->cells.p, h.data(), h.size() * sizeof(int32_t), hipMemcpyHostToDevice));
  if (P->dim == 3) {
    // new blocks: the block-interleaved geometry must be set again
    P->has_cells = false;
    P->has_geom = false;
    P->fused_geom = false;
    P->X.release();
    P->Xc.release();
    P->X32.release();
    P->g32_valid = P->gbf_valid = P->gm_valid = P->gm32_valid = false;
  }
  // 2D: G is stored per cell, independent of the blocks, and stays valid
  int rc = build_blocks(P, h);
  if (rc) return rc;
  P->has_cells = true;
  return SF_OK;
  SF_CATCH
}

static int finish_geometry(sf_plan* P) {
  P->g32_valid = P->gbf_valid = P->gm_valid = P->gm32_valid = false;
  P->pc_valid = false;
  P->gk_valid = false;
  // a plan whose FP64 apply reads the stored G gets kappa * G now (so its
  // applies, sf_cg_noalloc's included, never build it); with the fused
  // geometry it is built on first use
  return geo_f64(P) ? SF_OK : ensure_gk(P);
}

// 3D geometry lives in the block-interleaved layout of the apply, which
// depends on the cell blocks: the cells must be set first.
static int need_cells_for_geometry(const sf_plan* P) {
  if (P->dim == 3 && !P->has_cells)
    return fail(SF_ERR_NOT_READY, "set the cell dofs (sf_plan_set_cells) before the geometry");
  return SF_OK;
}

// The coefficient vectors of each cell's multilinear map from its vertices
// (vertex v = b0 + 2 b1 + 4 b2, D coordinates each), in the order the
// fused-geometry kernels read them: 3D a e p b c g d, 2D a e b (the terms
// xi, eta, zeta, xi eta, xi zeta, eta zeta, xi eta zeta of x - V0).
static std::vector<double> map_coefficients(int dim, const double* X, int64_t ncell) {
  const int nc = dim == 3 ? kGeoCoef3 : kGeoCoef2;
  std::vector<double> C((size_t)(ncell * nc));
  for (int64_t e = 0; e < ncell; ++e) {
    double* c = C.data() + e * nc;
    if (dim == 3) {
      const double* V = X + e * 24;
      for (int a = 0; a < 3; ++a) {
        const double v0 = V[a], v1 = V[3 + a], v2 = V[6 + a], v3 = V[9 + a], v4 = V[12 + a],
                     v5 = V[15 + a], v6 = V[18 + a], v7 = V[21 + a];
        c[a] = v1 - v0;                                          // a: xi
        c[3 + a] = v2 - v0;                                      // e: eta
        c[6 + a] = v4 - v0;                                      // p: zeta
        c[9 + a] = (v3 - v2) - (v1 - v0);                        // b: xi eta
        c[12 + a] = (v5 - v4) - (v1 - v0);                       // c: xi zeta
        c[15 + a] = (v6 - v4) - (v2 - v0);                       // g: eta zeta
        c[18 + a] = ((v7 - v6) - (v5 - v4)) - ((v3 - v2) - (v1 - v0));  // d: xi eta zeta
      }
    } else {
      const double* V = X + e * 8;
      for (int a = 0; a < 2; ++a) {
        const double v0 = V[a], v1 = V[2 + a], v2 = V[4 + a], v3 = V[6 + a];
        c[a] = v1 - v0;                    // a: xi
        c[2 + a] = v2 - v0;                // e: eta
        c[4 + a] = (v3 - v2) - (v1 - v0);  // b: xi eta
      }
    }
  }
  return C;
}

int sf_plan_set_vertices(sf_plan* P, const double* vertices) {
  SF_TRY
  if (check_plan(P)) return SF_ERR_ARG;
  if (!vertices && P->ncell > 0) return fail(SF_ERR_ARG, "null vertices");
  if (need_cells_for_geometry(P)) return SF_ERR_NOT_READY;
  DeviceGuard g(P->device);
  SF_HIP(hipDeviceSynchronize());  // kernels in flight may still read the plan
  const int64_t nv = P->ncell * (P->dim == 3 ? 24 : 8);
  DevBuf<double>& X = P->X;
  P->fused_geom = false;
  P->X32.release();
  SF_HIP(X.alloc(nv));
  SF_HIP(P->Xc.alloc(P->ncell * (P->dim == 3 ? kGeoCoef3 : kGeoCoef2)));
  int rc = SF_OK;
  do {
    if (nv > 0) {
      std::vector<double> tmp;
      const double* Xp = to_plan_order(P, vertices, nv / P->ncell, tmp);
      hipError_t e = hipMemcpy(X.p, Xp, nv * 8, hipMemcpyHostToDevice);
      if (e != hipSuccess) { rc = fail(SF_ERR_HIP, "vertex upload: %s", hipGetErrorString(e)); break; }
      const std::vector<double> C = map_coefficients(P->dim, Xp, P->ncell);
      e = hipMemcpy(P->Xc.p, C.data(), C.size() * 8, hipMemcpyHostToDevice);
      if (e != hipSuccess) { rc = fail(SF_ERR_HIP, "map coefficient upload: %s", hipGetErrorString(e)); break; }
    }
    hipError_t e = P->G.alloc(P->ncell * P->ng * P->nqc);
    if (e != hipSuccess) { rc = fail(SF_ERR_HIP, "G alloc: %s", hipGetErrorString(e)); break; }
    rc = clear_flag(P);
    if (rc) break;
    e = launch_geometry(P->dim, X.p, P->dPts.p, P->dW.p, P->q, P->ncell, P->cell_block.p,
                        P->bcell.p, P->G.p, P->flag.p, nullptr);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e != hipSuccess) { rc = fail(SF_ERR_HIP, "geometry kernel: %s", hipGetErrorString(e)); break; }
    int bad = 0;
    rc = read_flag(P, &bad);
    if (rc) break;
    if (bad) {
      P->has_geom = false;
      rc = fail(SF_ERR_JACOBIAN, "det J <= 0 or non-finite at some quadrature point");
      break;
    }
    P->has_geom = true;
    rc = finish_geometry(P);
  } while (0);
  // The apply recomputes G from the map coefficients (SFHIP_FUSED_GEOMETRY=0
  // reads the stored G instead; in 2D a kappa, which scales the stored G,
  // does too).  Every form keeps the vertices (24 or 8 values per cell): new
  // quadrature weights (sf_plan_set_tables) recompute G from them.
  const char* fg = getenv("SFHIP_FUSED_GEOMETRY");
  P->fused_geom = rc == SF_OK && !(fg && fg[0] == '0');
  if (rc != SF_OK) {
    X.release();
    P->Xc.release();
  }
  return rc;
  SF_CATCH
}

int sf_plan_set_geometry(sf_plan* P, const double* G) {
  SF_TRY
  if (check_plan(P)) return SF_ERR_ARG;
  if (!G && P->ncell > 0) return fail(SF_ERR_ARG, "null G");
  if (need_cells_for_geometry(P)) return SF_ERR_NOT_READY;
  DeviceGuard g(P->device);
  const int64_t n = P->ncell * P->ng * P->nqc;
  for (int64_t i = 0; i < n; ++i)
    if (!isfinite(G[i])) return fail(SF_ERR_ARG, "non-finite G entry at %lld", (long long)i);
  SF_HIP(hipDeviceSynchronize());  // kernels in flight may still read the plan
  SF_HIP(P->G.alloc(n));
  if (n > 0) {
    if (P->dim == 3) {
      DevBuf<double> tmp;
      SF_HIP(tmp.alloc(n));
      std::vector<double> hrows;
      const double* Gp = to_plan_order(P, G, P->ng * P->nqc, hrows);
      hipError_t e = hipMemcpy(tmp.p, Gp, n * 8, hipMemcpyHostToDevice);
      if (e == hipSuccess)
        e = launch_g3_relayout(tmp.p, P->G.p, P->q, P->ncell, P->cell_block.p, P->bcell.p, 1, nullptr);
      if (e == hipSuccess) e = hipDeviceSynchronize();
      tmp.release();
      SF_HIP(e);
    } else {
      SF_HIP(hipMemcpy(P->G.p, G, n * 8, hipMemcpyHostToDevice));
    }
  }
  P->has_geom = true;
  P->fused_geom = false;  // a caller-supplied G is what the apply must use
  P->X.release();
  P->Xc.release();
  P->X32.release();
  return finish_geometry(P);
  SF_CATCH
}

int sf_plan_get_geometry(const sf_plan* P, double* G) {
  SF_TRY
  if (check_plan(P)) return SF_ERR_ARG;
  if (!P->has_geom) return fail(SF_ERR_NOT_READY, "plan has no geometry");
  if (!G) return fail(SF_ERR_ARG, "null G");
  DeviceGuard g(P->device);
  const int64_t n = P->ncell * P->ng * P->nqc;
  if (n > 0) {
    if (P->dim == 3) {
      DevBuf<double> tmp;
      SF_HIP(tmp.alloc(n));
      hipError_t e = launch_g3_relayout(P->G.p, tmp.p, P->q, P->ncell, P->cell_block.p,
                                        P->bcell.p, 0, nullptr);
      std::vector<double> hrows;
      if (!P->perm.empty()) hrows.resize((size_t)n);
      double* dst = P->perm.empty() ? G : hrows.data();
      if (e == hipSuccess) e = hipMemcpy(dst, tmp.p, n * 8, hipMemcpyDeviceToHost);
      tmp.release();
      SF_HIP(e);
      if (!P->perm.empty()) {
        const int64_t w = P->ng * P->nqc;
        for (int64_t i = 0; i < P->ncell; ++i)
          std::copy(hrows.begin() + i * w, hrows.begin() + (i + 1) * w,
                    G + (int64_t)P->perm[i] * w);
      }
    } else {
      SF_HIP(hipMemcpy(G, P->G.p, n * 8, hipMemcpyDeviceToHost));
    }
  }
  return SF_OK;
  SF_CATCH
}

int sf_plan_set_kappa(sf_plan* P, const double* kappa) {
  SF_TRY
  if (check_plan(P)) return SF_ERR_ARG;
  if (!kappa && P->ncell > 0) return fail(SF_ERR_ARG, "null kappa");
  DeviceGuard g(P->device);
  const int64_t n = P->ncell * P->nqc;
  for (int64_t i = 0; i < n; ++i)
    if (!(kappa[i] > 0.0) || !isfinite(kappa[i]))
      return fail(SF_ERR_ARG, "kappa[%lld] = %g is not positive and finite", (long long)i, kappa[i]);
  // an apply, adjoint or solver still running on another stream reads
  // kappa, kappa * G and their copies: let it finish before they change
  SF_HIP(hipDeviceSynchronize());
  SF_HIP(P->kappa.alloc(n));
  std::vector<double> krows;
  const double* kp = to_plan_order(P, kappa, P->nqc, krows);
  if (n > 0) SF_HIP(hipMemcpy(P->kappa.p, kp, n * 8, hipMemcpyHostToDevice));
  P->has_kappa = true;
  P->k32_valid = false;
  P->pc_valid = false;
  if (P->has_geom) return finish_geometry(P);
  P->g32_valid = P->gbf_valid = P->gm_valid = P->gm32_valid = false;
  return SF_OK;
  SF_CATCH
}

int sf_plan_prepare(sf_plan* P, int precision) {
  SF_TRY
  int rc = check_ready(P);
  if (rc) return rc;
  DeviceGuard g(P->device);
  return prepare(P, precision);
  SF_CATCH
}

int sf_apply(sf_plan* P, double* y, const double* x, void* stream) {
  SF_TRY
  MarkerRange mark("sfhip apply");
  int rc = check_ready(P);
  if (rc) return rc;
  if ((!x || !y) && P->nfree > 0) return fail(SF_ERR_ARG, "null vector");
  DeviceGuard g(P->device);
  return do_apply_f64(P, y, x, (hipStream_t)stream);
  SF_CATCH
}

int sf_apply_add(sf_plan* P, double* y, const double* x, void* stream) {
  SF_TRY
  int rc = check_ready(P);
  if (rc) return rc;
  if ((!x || !y) && P->nfree > 0) return fail(SF_ERR_ARG, "null vector");
  if (P->nfree == 0) return SF_OK;
  DeviceGuard g(P->device);
  return do_apply_f64(P, y, x, (hipStream_t)stream, 1);
  SF_CATCH
}

int sf_apply_host(sf_plan* P, double* y, const double* x) {
  SF_TRY
  int rc = check_ready(P);
  if (rc) return rc;
  if ((!x || !y) && P->nfree > 0) return fail(SF_ERR_ARG, "null vector");
  DeviceGuard g(P->device);
  SF_HIP(P->xs.alloc(P->nfree));
  SF_HIP(P->ys.alloc(P->nfree));
  if (P->nfree > 0)
    SF_HIP(hipMemcpy(P->xs.p, x, P->nfree * 8, hipMemcpyHostToDevice));
  rc = do_apply_f64(P, P->ys.p, P->xs.p, nullptr);
  if (rc) return rc;
  if (P->nfree > 0)
    SF_HIP(hipMemcpy(y, P->ys.p, P->nfree * 8, hipMemcpyDeviceToHost));
  SF_HIP(hipDeviceSynchronize());
  return SF_OK;
  SF_CATCH
}

int sf_apply_f32(sf_plan* P, float* y, const float* x, int precision,
                 void* stream) {
  SF_TRY
  MarkerRange mark("sfhip apply (reduced precision)");
  int rc = check_ready(P);
  if (rc) return rc;
  if ((!x || !y) && P->nfree > 0) return fail(SF_ERR_ARG, "null vector");
  if (!is_reduced(precision))
    return fail(SF_ERR_ARG, "precision %d is not a reduced-precision operator", precision);
  DeviceGuard g(P->device);
  rc = prepare(P, precision);
  if (rc) return rc;
  return do_apply_f32(P, y, x, precision, (hipStream_t)stream);
  SF_CATCH
}

int sf_adjoint(sf_plan* P, const double* u, const double* lam,
               double* dJdkappa, double* dJdu, void* stream) {
  SF_TRY
  MarkerRange mark("sfhip adjoint");
  int rc = check_ready(P);
  if (rc) return rc;
  if ((!u || !lam) && P->nfree > 0) return fail(SF_ERR_ARG, "null vector");
  DeviceGuard g(P->device);
  hipStream_t s = (hipStream_t)stream;
  if (P->nfree == 0) {  // every dof Dirichlet: u = lam = 0, so dJ/dkappa = 0
    if (dJdkappa) SF_HIP(launch_fill<double>(dJdkappa, P->ncell * P->nqc, 0.0, s));
    return SF_OK;
  }
  if (fused_f64(P)) {
    // block-assembled kernel, geometry from the cell vertices
    if (!dJdkappa) return dJdu ? do_apply_f64(P, dJdu, lam, s) : SF_OK;
    if (dJdu) SF_HIP(launch_fill<double>(dJdu, P->nfree, 0.0, s));
    hipError_t e = adjoint3d_fused_f64(P->n, P->q, u, lam, dJdkappa, dJdu, &P->kargs,
                                       P->dB.p, P->dD.p, P->Xc.p, P->dPts.p, P->dW.p,
                                       P->has_kappa ? P->kappa.p : nullptr, P->corig.p, s);
    if (e == hipErrorNotSupported)
      return fail(SF_ERR_UNSUPPORTED, "no adjoint kernel for dim=%d n=%d q=%d", P->dim, P->n,
                  P->q);
    SF_HIP(e);
    return SF_OK;
  }
  if (dJdu) SF_HIP(launch_fill<double>(dJdu, P->nfree, 0.0, s));
  hipError_t e = adjoint_f64(P->dim, P->n, P->q, u, lam, dJdkappa, dJdu,
                             P->cells.p, P->G.p, P->has_kappa ? P->kappa.p : nullptr,
                             P->dB.p, P->dD.p, P->ncell, P->cell_block.p, P->bcell.p,
                             P->corig.p, s);
  if (e == hipErrorNotSupported)
    return fail(SF_ERR_UNSUPPORTED, "no adjoint kernel for dim=%d n=%d q=%d", P->dim, P->n, P->q);
  SF_HIP(e);
  return SF_OK;
  SF_CATCH
}

int sf_diagonal(sf_plan* P, double* diag, void* stream) {
  SF_TRY
  int rc = check_ready(P);
  if (rc) return rc;
  if (!diag && P->nfree > 0) return fail(SF_ERR_ARG, "null diag");
  DeviceGuard g(P->device);
  return compute_diagonal(P, diag, (hipStream_t)stream);
  SF_CATCH
}

int sf_plan_set_preconditioner(sf_plan* P, int kind) {
  SF_TRY
  if (check_plan(P)) return SF_ERR_ARG;
  if (kind != SF_PC_NONE && kind != SF_PC_JACOBI && kind != SF_PC_GLL)
    return fail(SF_ERR_ARG, "unknown preconditioner %d", kind);
  if (kind == SF_PC_GLL && P->dim != 3)
    return fail(SF_ERR_UNSUPPORTED, "the Gauss-Lobatto preconditioner is 3D only");
  if (kind != P->pc_kind) P->pc_valid = false;
  P->pc_kind = kind;
  if (kind == SF_PC_NONE) {
    DeviceGuard g(P->device);
    P->dinv.release();
  }
  return SF_OK;
  SF_CATCH
}

int sf_precond_apply(sf_plan* P, double* z, const double* r, void* stream) {
  SF_TRY
  int rc = check_ready(P);
  if (rc) return rc;
  if ((!z || !r) && P->nfree > 0) return fail(SF_ERR_ARG, "null vector");
  DeviceGuard g(P->device);
  hipStream_t s = (hipStream_t)stream;
  if (P->nfree == 0) return SF_OK;
  const float* dinv = nullptr;
  bool gll = false;
  rc = ensure_pc(P, s, &dinv, &gll);
  if (rc) return rc;
  if (gll) {
    SF_HIP(P->w_w.alloc(P->nfree));
    return pc_gll<double>(P, r, z, P->w_w.p, nullptr, s);
  }
  if (dinv) {
    SF_HIP(launch_scale_vec(r, dinv, z, P->nfree, s));
    return SF_OK;
  }
  SF_HIP(hipMemcpyAsync(z, r, P->nfree * 8, hipMemcpyDeviceToDevice, s));
  return SF_OK;
  SF_CATCH
}

int sf_cg(sf_plan* P, const double* b, double* x, double rtol, int maxit,
          sf_solve_stats* stats, void* stream) {
  SF_TRY
  int rc = check_ready(P);
  if (rc) return rc;
  if ((!b || !x) && P->nfree > 0) return fail(SF_ERR_ARG, "null vector");
  if (!(rtol >= 0.0) || maxit < 0) return fail(SF_ERR_ARG, "bad rtol / maxit");
  if (empty_solve(P, stats)) return SF_OK;
  DeviceGuard g(P->device);
  hipStream_t s = (hipStream_t)stream;
  rc = ensure_solver_ws(P, false);
  if (rc) return rc;
  return cg_solve(P, b, x, rtol, maxit, stats, s);
  SF_CATCH
}

int sf_cg_noalloc(sf_plan* P, const double* b, double* x, double rtol, int maxit,
                  sf_solve_stats* stats, void* stream) {
  SF_TRY
  int rc = check_ready(P);
  if (rc) return rc;
  if ((!b || !x) && P->nfree > 0) return fail(SF_ERR_ARG, "null vector");
  if (!(rtol >= 0.0) || maxit < 0) return fail(SF_ERR_ARG, "bad rtol / maxit");
  if (empty_solve(P, stats)) return SF_OK;
  if (P->pc_kind != SF_PC_NONE && !P->pc_valid && P->nfree > 0)
    return fail(SF_ERR_NOT_READY,
                "the preconditioner is not built yet (one sf_cg or sf_precond_apply call "
                "builds it; the no-allocation CG cannot)");
  DeviceGuard g(P->device);
  hipStream_t s = (hipStream_t)stream;
  // the workspace sf_cg allocates, carved out of this GPU's scratch array
  const bool gll = P->pc_kind == SF_PC_GLL;
  const int64_t nf = P->nfree;
  const int64_t np = (P->dim == 3 && P->npart > 0) ? P->npart : 0;
  ScratchCarve c;
  const size_t o_r = c.take(nf * 8), o_p = c.take(nf * 8), o_q = c.take(nf * 8),
               o_t = c.take(nf * 8);
  const size_t o_gb = gll ? c.take(nf * 8) : 0, o_gy = gll ? c.take(nf * 8) : 0;
  const size_t o_red = c.take(2 * (size_t)kPartSlots * 8), o_part = c.take(np * 8);
  const size_t o_cnt = c.take(sizeof(unsigned)), o_sc = c.take(sizeof(CGScalars));
  const size_t o_ds = c.take(4 * 8);
  if (c.off > kScratchBytes)
    return fail(SF_ERR_ARG,
                "the no-allocation CG needs %zu bytes of workspace for %lld free dofs; its "
                "scratch holds %zu",
                c.off, (long long)nf, (size_t)kScratchBytes);
  char* base = nullptr;
  rc = scratch_base(P->device, &base);
  if (rc) return rc;
  BufSwap<double> s_r(P->w_r, (double*)(base + o_r), nf), s_p(P->w_p, (double*)(base + o_p), nf),
      s_q(P->w_q, (double*)(base + o_q), nf), s_t(P->w_t, (double*)(base + o_t), nf),
      s_gb(P->g_b, gll ? (double*)(base + o_gb) : P->g_b.p, gll ? nf : P->g_b.n),
      s_gy(P->g_y, gll ? (double*)(base + o_gy) : P->g_y.p, gll ? nf : P->g_y.n),
      s_red(P->red_part, (double*)(base + o_red), 2 * (int64_t)kPartSlots),
      s_part(P->w_part, np ? (double*)(base + o_part) : P->w_part.p, np ? np : P->w_part.n),
      s_ds(P->dscal, (double*)(base + o_ds), 4);
  BufSwap<unsigned> s_cnt(P->red_counter, (unsigned*)(base + o_cnt), 1);
  BufSwap<CGScalars> s_sc(P->scal, (CGScalars*)(base + o_sc), 1);
  // the arrival counter of the single-launch reductions starts at zero
  SF_HIP(launch_fill<unsigned>(P->red_counter.p, 1, 0u, s));
  return cg_solve(P, b, x, rtol, maxit, stats, s);
  SF_CATCH
}

int sf_cg_mixed(sf_plan* P, const double* b, double* x, double rtol, int maxit,
                int inner_precision, double delta, int fp64_recurrence,
                sf_solve_stats* stats, void* stream) {
  SF_TRY
  int rc = check_ready(P);
  if (rc) return rc;
  if ((!b || !x) && P->nfree > 0) return fail(SF_ERR_ARG, "null vector");
  if (!(rtol >= 0.0) || maxit < 0 || !(delta > 0.0 && delta < 1.0))
    return fail(SF_ERR_ARG, "bad rtol / maxit / delta");
  if (fp64_recurrence < 0 || fp64_recurrence > 2)
    return fail(SF_ERR_ARG, "fp64_recurrence must be 0, 1 or 2 (got %d)", fp64_recurrence);
  if (!is_reduced(inner_precision))
    return fail(SF_ERR_UNSUPPORTED, "inner precision %d not available", inner_precision);
  if (empty_solve(P, stats)) return SF_OK;
  DeviceGuard g(P->device);
  hipStream_t s = (hipStream_t)stream;
  rc = prepare(P, inner_precision);
  if (rc) return rc;
  rc = ensure_solver_ws(P, true);
  if (rc) return rc;
  MarkerRange mark("sfhip mixed cg");
  const float* dinv = nullptr;
  bool gll = false;
  rc = ensure_pc(P, s, &dinv, &gll);
  if (rc) return rc;
  SF_HIP(hipStreamSynchronize(s));
  const double t0 = now_s();
  int iters = 0, updates = 0, nonfinite = 0;
  double rel = 0.0;
  auto run = [&](const double* rb, double* rx, double rt, int mi, int* it, int* nf) {
    int up = 0;
    double rl = 0.0;
    const int r = fp64_recurrence
                      ? cg_reliable_f64rec(P, inner_precision, rb, rx, rt, mi, delta, s, it,
                                           &up, nf, &rl, fp64_recurrence == 2, dinv)
                      : cg_reliable(P, inner_precision, rb, rx, rt, mi, delta, s, it, &up,
                                    nf, &rl, dinv);
    updates += up;
    if (!gll) rel = rl;
    return r;
  };
  int restarts = 0;
  if (gll)
    rc = solve_gll(P, b, x, rtol, maxit, s, run, &iters, &rel, &nonfinite, &restarts);
  else
    rc = run(b, x, rtol, maxit, &iters, &nonfinite);
  if (rc) return rc;
  SF_HIP(hipStreamSynchronize(s));
  const double t1 = now_s();
  if (stats) {
    stats->iterations = iters;
    stats->outer_iterations = updates;
    stats->rel_residual = rel;
    stats->converged = (rel <= rtol) ? 1 : 0;
    stats->seconds = t1 - t0;
  }
  if (nonfinite) return fail(SF_ERR_NONFINITE, "mixed CG met a non-finite value");
  if (rel > rtol)
    return fail(SF_ERR_NOT_CONVERGED, "mixed CG did not reach rtol %g in %d iterations (rel. residual %g)", rtol, maxit, rel);
  return SF_OK;
  SF_CATCH
}

int sf_cg_ir(sf_plan* P, const double* b, double* x, double rtol,
             int max_outer, int inner_precision, double inner_rtol,
             int inner_maxit, sf_solve_stats* stats, void* stream) {
  SF_TRY
  int rc = check_ready(P);
  if (rc) return rc;
  if ((!b || !x) && P->nfree > 0) return fail(SF_ERR_ARG, "null vector");
  if (!(rtol >= 0.0) || max_outer < 0 || !(inner_rtol > 0.0) || inner_maxit < 1)
    return fail(SF_ERR_ARG, "bad tolerances / iteration limits");
  if (!is_reduced(inner_precision))
    return fail(SF_ERR_UNSUPPORTED, "inner precision %d not available for refinement", inner_precision);
  if (empty_solve(P, stats)) return SF_OK;
  DeviceGuard g(P->device);
  hipStream_t s = (hipStream_t)stream;
  rc = prepare(P, inner_precision);
  if (rc) return rc;
  rc = ensure_solver_ws(P, true);
  if (rc) return rc;
  const float* dinv = nullptr;
  bool gll = false;
  rc = ensure_pc(P, s, &dinv, &gll);
  if (rc) return rc;
  SF_HIP(hipStreamSynchronize(s));
  const double t0 = now_s();
  int inner_total = 0, outer = 0, nonfinite = 0;
  double rel = 0.0;
  auto run = [&](const double* rb, double* rx, double rt, int mo, int* it, int* nf) {
    int inner = 0;
    double rl = 0.0;
    const int r = refine(P, rb, rx, rt, mo, inner_precision, inner_rtol, inner_maxit, s,
                         &inner, it, nf, &rl, dinv);
    inner_total += inner;
    if (!gll) rel = rl;
    return r;
  };
  int restarts = 0;
  if (gll)
    rc = solve_gll(P, b, x, rtol, max_outer, s, run, &outer, &rel, &nonfinite, &restarts);
  else
    rc = run(b, x, rtol, max_outer, &outer, &nonfinite);
  if (rc) return rc;
  SF_HIP(hipStreamSynchronize(s));
  const double t1 = now_s();
  if (stats) {
    stats->iterations = inner_total;
    stats->outer_iterations = outer;
    stats->rel_residual = rel;
    stats->converged = (rel <= rtol) ? 1 : 0;
    stats->seconds = t1 - t0;
  }
  if (nonfinite) return fail(SF_ERR_NONFINITE, "refinement met a non-finite value");
  if (rel > rtol)
    return fail(SF_ERR_NOT_CONVERGED, "refinement did not reach rtol %g in %d outer steps (rel. residual %g)", rtol, max_outer, rel);
  return SF_OK;
  SF_CATCH
}

}  // extern "C"
