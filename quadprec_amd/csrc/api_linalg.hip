// api_linalg.hip -- C ABI of the factorisations, triangular solves, matvec, the solver
// crate's SPD solves and the overlap build; device-side vector kernels they need.
#include <cfloat>
#include <cmath>
#include <cstring>
#include <vector>

#include "hp_chol.hpp"
#include "hp_internal.hpp"

namespace hp {

template <class A>
int chol_exact(Mat a, long n, long nb, bool unblocked, Mat S, ErrRec* err, double* inv_h, double* inv_l, hipStream_t s);
template <class A>
int fwd_exact(Mat l, double* yh, double* yl, long n, hipStream_t s);
template <class A>
int bwd_exact(Mat l, double* xh, double* xl, long n, hipStream_t s);
template <class A>
int matvec_exact(Mat a, const double* xh, const double* xl, double* yh, double* yl, long n, hipStream_t s);
int zero_upper_f64(double* l, long n, hipStream_t s);
int launch_overlap_s(const double* centers, const double* alpha, size_t n, double* shi, double* slo, hipStream_t s);

namespace {

// y = x + sign*z elementwise in dd (sign = -1: dd sub, dd.rs:159-165)
__global__ void k_dd_axpy(const double* xh, const double* xl, const double* zh, const double* zl, double* yh,
                          double* yl, long n, int negate) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  for (; i < n; i += (long)gridDim.x * blockDim.x) {
    dd x = mkdd(xh[i], xl[i]), z = mkdd(zh[i], zl[i]);
    dd y = negate ? dd_sub(x, z) : dd_add(x, z);
    yh[i] = y.hi;
    yl[i] = y.lo;
  }
}
// norm_f128 (crates/solver/src/lib.rs:243-249): sequential s = s + v*v, sqrt, to_f64
__global__ void k_norm_dd(const double* vh, const double* vl, long n, double* out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  dd s = mkdd(0.0);
  for (long i = 0; i < n; ++i) {
    dd v = mkdd(vh[i], vl[i]);
    s = dd_add(s, dd_mul(v, v));
  }
  dd r = dd_sqrt(s);
  out[0] = r.hi + r.lo;
}
__global__ void k_fill(double* p, long n, double v) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  for (; i < n; i += (long)gridDim.x * blockDim.x) p[i] = v;
}
__global__ void k_get_diag(const double* hi, const double* lo, long n, double* dh, double* dl) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    dh[i] = hi[i * n + i];
    if (lo) dl[i] = lo[i * n + i];
  }
}
unsigned grid1(long n) {
  long g = (n + 255) / 256;
  return (unsigned)std::max<long>(1, std::min<long>(g, 65536));
}

// Device workspace for the factorisations: partial-sum planes S (n x n, lower part used),
// 64 inverse pivots and the error record.
struct Work {
  DevBuf S, inv, err;
  int alloc(long n, int limbs) {
    int rc;
    if ((rc = S.alloc((size_t)n * n * limbs * sizeof(double)))) return rc;
    if ((rc = inv.alloc(2 * 64 * sizeof(double)))) return rc;
    if ((rc = err.alloc(sizeof(ErrRec)))) return rc;
    return HP_OK;
  }
};

// Factor on device; status HP_ENOTSPD with pivot / value from the device error record.
template <class A>
int chol_dev(Mat a, long n, long nb, bool unblocked, size_t* pivot, double* value, hipStream_t s) {
  Work w;
  int rc = w.alloc(std::max<long>(n, 1), A::LIMBS);
  if (rc) return rc;
  HP_HIP(hipMemsetAsync(w.err.p, 0, sizeof(ErrRec), s));
  Mat S{w.S.d(), A::LIMBS == 2 ? w.S.d() + (size_t)n * n : nullptr, n};
  rc = chol_exact<A>(a, n, nb, unblocked, S, (ErrRec*)w.err.p, w.inv.d(), w.inv.d() + 64, s);
  if (rc) return rc;
  ErrRec e;
  HP_HIP(hipMemcpyAsync(&e, w.err.p, sizeof(e), hipMemcpyDeviceToHost, s));
  HP_HIP(hipStreamSynchronize(s));
  if (e.flag) {
    if (pivot) *pivot = (size_t)e.pivot;
    if (value) *value = e.value;
    set_error("Cholesky failed at pivot %lld (value=%.4e)", e.pivot, e.value);
    return HP_ENOTSPD;
  }
  return HP_OK;
}

int norm_dev(const double* vh, const double* vl, long n, double* scratch, double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_norm_dd, dim3(1), dim3(64), 0, s, vh, vl, n, scratch);
  HP_LAUNCH_CHECK();
  HP_HIP(hipMemcpyAsync(out, scratch, sizeof(double), hipMemcpyDeviceToHost, s));
  HP_HIP(hipStreamSynchronize(s));
  return HP_OK;
}

// cond_estimate (linalg.rs:185-194) / cond_estimate_f64 (lib.rs:215-224) from the
// factor's diagonal (read back from the device).
int cond_dev(const double* hi, const double* lo, long n, double* kappa, hipStream_t s) {
  DevBuf d;
  int rc = d.alloc(2 * n * sizeof(double));
  if (rc) return rc;
  hipLaunchKernelGGL(k_get_diag, dim3(grid1(n)), dim3(256), 0, s, hi, lo, n, d.d(), d.d() + n);
  HP_LAUNCH_CHECK();
  std::vector<double> h(2 * n, 0.0);
  HP_HIP(hipMemcpyAsync(h.data(), d.p, 2 * n * sizeof(double), hipMemcpyDeviceToHost, s));
  HP_HIP(hipStreamSynchronize(s));
  double dmax = 0.0, dmin = DBL_MAX;
  for (long i = 0; i < n; ++i) {
    double v = std::fabs(lo ? h[i] + h[n + i] : h[i]);
    dmax = std::fmax(dmax, v);
    dmin = std::fmin(dmin, v);
  }
  double r = dmax / dmin;
  *kappa = r * r;
  return HP_OK;
}

int axpy_dev(const double* xh, const double* xl, const double* zh, const double* zl, double* yh, double* yl, long n,
             bool negate, hipStream_t s) {
  hipLaunchKernelGGL(k_dd_axpy, dim3(grid1(n)), dim3(256), 0, s, xh, xl, zh, zl, yh, yl, n, negate ? 1 : 0);
  HP_LAUNCH_CHECK();
  return HP_OK;
}

// dd host (AoS) staging helpers
struct DDVec {
  DevBuf b;
  long n = 0;
  int alloc(long count) { n = count; return b.alloc(2 * std::max<long>(count, 1) * sizeof(double)); }
  double* hi() { return b.d(); }
  double* lo() { return b.d() + n; }
};

// residual_norm_f128_from_f128 (lib.rs:237-241): ||b - A x|| with the exact matvec
int residual_norm(Mat a, DDVec& x, DDVec& b, long n, double* out, hipStream_t s) {
  DDVec ax, r;
  DevBuf sc;
  int rc;
  if ((rc = ax.alloc(n)) || (rc = r.alloc(n)) || (rc = sc.alloc(8))) return rc;
  if ((rc = matvec_exact<ArDD>(a, x.hi(), x.lo(), ax.hi(), ax.lo(), n, s))) return rc;
  if ((rc = axpy_dev(b.hi(), b.lo(), ax.hi(), ax.lo(), r.hi(), r.lo(), n, true, s))) return rc;
  return norm_dev(r.hi(), r.lo(), n, sc.d(), out, s);
}

int solve_cholesky_dd(Mat l, const double* bh, const double* bl, double* xh, double* xl, long n, hipStream_t s) {
  if (xh != bh) {
    HP_HIP(hipMemcpyAsync(xh, bh, n * sizeof(double), hipMemcpyDeviceToDevice, s));
    HP_HIP(hipMemcpyAsync(xl, bl, n * sizeof(double), hipMemcpyDeviceToDevice, s));
  }
  int rc = fwd_exact<ArDD>(l, xh, xl, n, s);
  if (rc) return rc;
  return bwd_exact<ArDD>(l, xh, xl, n, s);
}

// solve_pure_f128 (lib.rs:127-144) on device-resident dd A (planes) and b
int solve_pure_dd(Mat a, DDVec& b, long n, hp_solution* sol, double* x_dd, double* x_f64, hipStream_t s) {
  DevBuf lb;
  int rc;
  if ((rc = lb.alloc(2 * (size_t)n * n * sizeof(double)))) return rc;
  HP_HIP(hipMemcpyAsync(lb.p, a.hi, (size_t)n * n * sizeof(double), hipMemcpyDeviceToDevice, s));
  HP_HIP(hipMemcpyAsync(lb.d() + (size_t)n * n, a.lo, (size_t)n * n * sizeof(double), hipMemcpyDeviceToDevice, s));
  Mat l{lb.d(), lb.d() + (size_t)n * n, n};
  size_t pv = 0;
  double vv = 0;
  rc = chol_dev<ArDD>(l, n, 0, true, &pv, &vv, s);
  if (rc == HP_ENOTSPD) { sol->pivot = pv; return rc; }
  if (rc) return rc;
  if ((rc = cond_dev(l.hi, l.lo, n, &sol->kappa, s))) return rc;
  DDVec x;
  if ((rc = x.alloc(n))) return rc;
  if ((rc = solve_cholesky_dd(l, b.hi(), b.lo(), x.hi(), x.lo(), n, s))) return rc;
  if ((rc = residual_norm(a, x, b, n, &sol->residual_norm, s))) return rc;
  if ((rc = download_planes(x.hi(), n, 2, x_dd, s))) return rc;
  for (long i = 0; i < n; ++i) x_f64[i] = x_dd[2 * i] + x_dd[2 * i + 1];
  sol->strategy = HP_STRATEGY_FLOAT128;
  sol->refinement_iters = 0;
  return HP_OK;
}

// iterative_refinement (lib.rs:75-125)
int iterative_refinement_dd(Mat a, DDVec& b, const double* x64_dev, long n, hp_solution* sol, double* x_dd,
                            double* x_f64, hipStream_t s) {
  DevBuf lb, sc;
  int rc;
  if ((rc = lb.alloc(2 * (size_t)n * n * sizeof(double))) || (rc = sc.alloc(8))) return rc;
  HP_HIP(hipMemcpyAsync(lb.p, a.hi, (size_t)n * n * sizeof(double), hipMemcpyDeviceToDevice, s));
  HP_HIP(hipMemcpyAsync(lb.d() + (size_t)n * n, a.lo, (size_t)n * n * sizeof(double), hipMemcpyDeviceToDevice, s));
  Mat l{lb.d(), lb.d() + (size_t)n * n, n};
  size_t pv = 0;
  double vv = 0;
  rc = chol_dev<ArDD>(l, n, 0, true, &pv, &vv, s);
  if (rc == HP_ENOTSPD) { sol->pivot = pv; return rc; }
  if (rc) return rc;
  if ((rc = cond_dev(l.hi, l.lo, n, &sol->kappa, s))) return rc;
  DDVec x, ax, r, dx;
  if ((rc = x.alloc(n)) || (rc = ax.alloc(n)) || (rc = r.alloc(n)) || (rc = dx.alloc(n))) return rc;
  HP_HIP(hipMemcpyAsync(x.hi(), x64_dev, n * sizeof(double), hipMemcpyDeviceToDevice, s));
  hipLaunchKernelGGL(k_fill, dim3(grid1(n)), dim3(256), 0, s, x.lo(), n, 0.0);
  HP_LAUNCH_CHECK();
  double bnorm = 0.0;
  if ((rc = norm_dev(b.hi(), b.lo(), n, sc.d(), &bnorm, s))) return rc;
  double best = DBL_MAX;
  int iters = 0;
  for (int it = 0; it < 30; ++it) {
    if ((rc = matvec_exact<ArDD>(a, x.hi(), x.lo(), ax.hi(), ax.lo(), n, s))) return rc;
    if ((rc = axpy_dev(b.hi(), b.lo(), ax.hi(), ax.lo(), r.hi(), r.lo(), n, true, s))) return rc;
    double rnorm = 0.0;
    if ((rc = norm_dev(r.hi(), r.lo(), n, sc.d(), &rnorm, s))) return rc;
    iters += 1;
    if (rnorm < 1e-28 * bnorm) break;
    if (iters > 3 && rnorm > 0.5 * best) break;
    if (rnorm < best) best = rnorm;
    if ((rc = solve_cholesky_dd(l, r.hi(), r.lo(), dx.hi(), dx.lo(), n, s))) return rc;
    if ((rc = axpy_dev(x.hi(), x.lo(), dx.hi(), dx.lo(), x.hi(), x.lo(), n, false, s))) return rc;
  }
  if ((rc = residual_norm(a, x, b, n, &sol->residual_norm, s))) return rc;
  if ((rc = download_planes(x.hi(), n, 2, x_dd, s))) return rc;
  for (long i = 0; i < n; ++i) x_f64[i] = x_dd[2 * i] + x_dd[2 * i + 1];
  sol->strategy = HP_STRATEGY_MIXED;
  sol->refinement_iters = iters;
  return HP_OK;
}

int check_ptrs(const void* a, const void* b, const void* c) {
  if (!a || !b || !c) {
    set_error("null pointer argument");
    return HP_EINVAL;
  }
  return HP_OK;
}

}  // namespace
}  // namespace hp

using namespace hp;

extern "C" {

int hp_cholesky_blocked(double* a, size_t n, size_t nb, size_t* pivot, double* value) {
  if (n == 0) return HP_OK;
  if (!a) { set_error("null pointer argument"); return HP_EINVAL; }
  hipStream_t s = hipStreamPerThread;
  DevBuf da;
  int rc;
  if ((rc = da.alloc(2 * n * n * sizeof(double)))) return rc;
  if ((rc = upload_planes(a, n * n, 2, da.d(), s))) return rc;
  Mat m{da.d(), da.d() + n * n, (long)n};
  rc = chol_dev<ArDD>(m, (long)n, (long)nb, false, pivot, value, s);
  if (rc) return rc;
  return download_planes(da.d(), n * n, 2, a, s);
}

int hp_cholesky(double* a, size_t n, size_t* pivot, double* value) {
  if (n == 0) return HP_OK;
  if (!a) { set_error("null pointer argument"); return HP_EINVAL; }
  hipStream_t s = hipStreamPerThread;
  DevBuf da;
  int rc;
  if ((rc = da.alloc(2 * n * n * sizeof(double)))) return rc;
  if ((rc = upload_planes(a, n * n, 2, da.d(), s))) return rc;
  Mat m{da.d(), da.d() + n * n, (long)n};
  rc = chol_dev<ArDD>(m, (long)n, 0, true, pivot, value, s);
  if (rc) return rc;
  return download_planes(da.d(), n * n, 2, a, s);
}

int hp_dev_cholesky(double* a_hi, double* a_lo, size_t n, size_t nb, int unblocked, size_t* pivot, double* value,
                    void* stream) {
  if (n == 0) return HP_OK;
  Mat m{a_hi, a_lo, (long)n};
  return chol_dev<ArDD>(m, (long)n, (long)nb, unblocked != 0, pivot, value, (hipStream_t)stream);
}

static int vec_op(int which, const double* l, const double* b, double* out, size_t n) {
  if (n == 0) return HP_OK;
  int rc = check_ptrs(l, b, out);
  if (rc) return rc;
  hipStream_t s = hipStreamPerThread;
  DevBuf dl;
  DDVec x, y;
  if ((rc = dl.alloc(2 * n * n * sizeof(double))) || (rc = x.alloc((long)n)) || (rc = y.alloc((long)n))) return rc;
  if ((rc = upload_planes(l, n * n, 2, dl.d(), s)) || (rc = upload_planes(b, n, 2, x.hi(), s))) return rc;
  Mat m{dl.d(), dl.d() + n * n, (long)n};
  switch (which) {
    case 0: rc = fwd_exact<ArDD>(m, x.hi(), x.lo(), (long)n, s); break;
    case 1: rc = bwd_exact<ArDD>(m, x.hi(), x.lo(), (long)n, s); break;
    case 2: rc = solve_cholesky_dd(m, x.hi(), x.lo(), x.hi(), x.lo(), (long)n, s); break;
    case 3:
      rc = matvec_exact<ArDD>(m, x.hi(), x.lo(), y.hi(), y.lo(), (long)n, s);
      if (!rc) return download_planes(y.hi(), n, 2, out, s);
      break;
  }
  if (rc) return rc;
  return download_planes(x.hi(), n, 2, out, s);
}

int hp_forward_solve(const double* l, const double* b, double* y, size_t n) { return vec_op(0, l, b, y, n); }
int hp_backward_solve(const double* l, const double* y, double* x, size_t n) { return vec_op(1, l, y, x, n); }
int hp_solve_cholesky(const double* l, const double* b, double* x, size_t n) { return vec_op(2, l, b, x, n); }
int hp_matvec(const double* a, const double* x, double* y, size_t n) { return vec_op(3, a, x, y, n); }

double hp_cond_estimate(const double* l, size_t n) {
  double dmax = 0.0, dmin = DBL_MAX;
  for (size_t i = 0; i < n; ++i) {
    double d = std::fabs(l[2 * (i * n + i)] + l[2 * (i * n + i) + 1]);
    dmax = std::fmax(dmax, d);
    dmin = std::fmin(dmin, d);
  }
  double r = dmax / dmin;
  return r * r;
}

int hp_solve_spd_f128(const double* a_dd, const double* b_dd, size_t n, hp_solution* sol, double* x_dd,
                      double* x_f64) {
  if (!sol) { set_error("null solution"); return HP_EINVAL; }
  std::memset(sol, 0, sizeof(*sol));
  if (n == 0) return HP_OK;
  int rc = check_ptrs(a_dd, b_dd, x_dd);
  if (rc) return rc;
  hipStream_t s = hipStreamPerThread;
  DevBuf da;
  DDVec b;
  if ((rc = da.alloc(2 * n * n * sizeof(double))) || (rc = b.alloc((long)n))) return rc;
  if ((rc = upload_planes(a_dd, n * n, 2, da.d(), s)) || (rc = upload_planes(b_dd, n, 2, b.hi(), s))) return rc;
  Mat a{da.d(), da.d() + n * n, (long)n};
  return solve_pure_dd(a, b, (long)n, sol, x_dd, x_f64, s);
}

int hp_solve_spd(const double* a, const double* b, size_t n, hp_solution* sol, double* x_dd, double* x_f64) {
  if (!sol) { set_error("null solution"); return HP_EINVAL; }
  std::memset(sol, 0, sizeof(*sol));
  if (n == 0) return HP_OK;
  int rc = check_ptrs(a, b, x_dd);
  if (rc) return rc;
  hipStream_t s = hipStreamPerThread;
  const long N = (long)n;
  // promote_matrix (lib.rs:228-230): dd planes with lo = 0
  DevBuf a128, l64, x64;
  DDVec b128;
  if ((rc = a128.alloc(2 * n * n * sizeof(double))) || (rc = l64.alloc(n * n * sizeof(double))) ||
      (rc = x64.alloc(n * sizeof(double))) || (rc = b128.alloc(N)))
    return rc;
  HP_HIP(hipMemcpyAsync(a128.p, a, n * n * sizeof(double), hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_fill, dim3(grid1(N * N)), dim3(256), 0, s, a128.d() + n * n, N * N, 0.0);
  HP_LAUNCH_CHECK();
  HP_HIP(hipMemcpyAsync(b128.hi(), b, n * sizeof(double), hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_fill, dim3(grid1(N)), dim3(256), 0, s, b128.lo(), N, 0.0);
  HP_LAUNCH_CHECK();
  Mat A128{a128.d(), a128.d() + n * n, N};
  // factor_and_solve_f64 (lib.rs:149-166) with the Linux cholesky_f64 / solve_cholesky_f64
  HP_HIP(hipMemcpyAsync(l64.p, a128.p, n * n * sizeof(double), hipMemcpyDeviceToDevice, s));
  Mat L64{l64.d(), nullptr, N};
  size_t pv = 0;
  double vv = 0;
  rc = chol_dev<ArF64>(L64, N, 0, true, &pv, &vv, s);
  if (rc == HP_ENOTSPD) {
    set_error("");
    rc = solve_pure_dd(A128, b128, N, sol, x_dd, x_f64, s);  // lib.rs:66-67
    if (rc == HP_ENOTSPD) set_error("Matrix is not positive definite (failed at pivot %zu)", sol->pivot);
    return rc;
  }
  if (rc) return rc;
  if ((rc = zero_upper_f64(l64.d(), N, s))) return rc;
  HP_HIP(hipMemcpyAsync(x64.p, b, n * sizeof(double), hipMemcpyHostToDevice, s));
  if ((rc = fwd_exact<ArF64>(L64, x64.d(), nullptr, N, s)) || (rc = bwd_exact<ArF64>(L64, x64.d(), nullptr, N, s)))
    return rc;
  double kappa = 0.0;
  if ((rc = cond_dev(l64.d(), nullptr, N, &kappa, s))) return rc;
  if (kappa < 1e10) {  // lib.rs:49-60
    DDVec x;
    if ((rc = x.alloc(N))) return rc;
    HP_HIP(hipMemcpyAsync(x.hi(), x64.p, n * sizeof(double), hipMemcpyDeviceToDevice, s));
    hipLaunchKernelGGL(k_fill, dim3(grid1(N)), dim3(256), 0, s, x.lo(), N, 0.0);
    HP_LAUNCH_CHECK();
    if ((rc = residual_norm(A128, x, b128, N, &sol->residual_norm, s))) return rc;
    HP_HIP(hipMemcpyAsync(x_f64, x64.p, n * sizeof(double), hipMemcpyDeviceToHost, s));
    HP_HIP(hipStreamSynchronize(s));
    for (size_t i = 0; i < n; ++i) { x_dd[2 * i] = x_f64[i]; x_dd[2 * i + 1] = 0.0; }
    sol->kappa = kappa;
    sol->strategy = HP_STRATEGY_FLOAT64;
    sol->refinement_iters = 0;
    return HP_OK;
  }
  rc = iterative_refinement_dd(A128, b128, x64.d(), N, sol, x_dd, x_f64, s);
  if (rc == HP_ENOTSPD) set_error("Matrix is not positive definite (failed at pivot %zu)", sol->pivot);
  return rc;
}

int hp_set_split_factor_panel(int nb) {
  if (nb != 64 && nb != 128 && nb != 256 && nb != 512) {
    set_error("split factor panel width must be 64, 128, 256 or 512");
    return HP_EINVAL;
  }
  g_split_panel.store(nb);
  return HP_OK;
}

int hp_dev_solve_spd_refined(int factor_kind, const double* s_hi, const double* s_lo, const double* b_hi,
                             const double* b_lo, size_t n, double tol, int max_iter, double* x_hi, double* x_lo,
                             hp_refine_stats* stats, void* stream) {
  hp_refine_stats local;
  if (!stats) stats = &local;
  return solve_refined(factor_kind, s_hi, s_lo, b_hi, b_lo, (long)n, tol, max_iter, x_hi, x_lo, stats,
                       (hipStream_t)stream);
}

int hp_solve_spd_refined(int factor_kind, const double* s_dd, const double* b_dd, size_t n, double tol, int max_iter,
                         double* x_dd, hp_refine_stats* stats) {
  hp_refine_stats local;
  if (!stats) stats = &local;
  std::memset(stats, 0, sizeof(*stats));
  if (n == 0) return HP_OK;
  int rc = check_ptrs(s_dd, b_dd, x_dd);
  if (rc) return rc;
  hipStream_t s = hipStreamPerThread;
  DevBuf ds;
  DDVec b, x;
  if ((rc = ds.alloc(2 * n * n * sizeof(double))) || (rc = b.alloc((long)n)) || (rc = x.alloc((long)n))) return rc;
  if ((rc = upload_planes(s_dd, n * n, 2, ds.d(), s)) || (rc = upload_planes(b_dd, n, 2, b.hi(), s))) return rc;
  rc = solve_refined(factor_kind, ds.d(), ds.d() + n * n, b.hi(), b.lo(), (long)n, tol, max_iter, x.hi(), x.lo(),
                     stats, s);
  if (rc) return rc;
  return download_planes(x.hi(), n, 2, x_dd, s);
}

int hp_overlap_s(const double* centers, const double* alpha, size_t n, double* s_dd) {
  if (n == 0) return HP_OK;
  int rc = check_ptrs(centers, alpha, s_dd);
  if (rc) return rc;
  hipStream_t s = hipStreamPerThread;
  DevBuf dc, da, ds;
  if ((rc = dc.alloc(3 * n * sizeof(double))) || (rc = da.alloc(n * sizeof(double))) ||
      (rc = ds.alloc(2 * n * n * sizeof(double))))
    return rc;
  HP_HIP(hipMemcpyAsync(dc.p, centers, 3 * n * sizeof(double), hipMemcpyHostToDevice, s));
  HP_HIP(hipMemcpyAsync(da.p, alpha, n * sizeof(double), hipMemcpyHostToDevice, s));
  if ((rc = launch_overlap_s(dc.d(), da.d(), n, ds.d(), ds.d() + n * n, s))) return rc;
  return download_planes(ds.d(), n * n, 2, s_dd, s);
}

int hp_dev_overlap_s(const double* centers, const double* alpha, size_t n, double* s_hi, double* s_lo, void* stream) {
  return launch_overlap_s(centers, alpha, n, s_hi, s_lo, (hipStream_t)stream);
}

}  // extern "C"
