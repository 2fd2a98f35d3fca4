// k_solve.hip -- mixed-precision SPD solves with quad-double residual refinement.
//
// The solver crate's iterative_refinement (crates/solver/src/lib.rs:75-125) factors once and
// corrects x with solves against residuals; its limiting accuracy is kappa times the
// residual's precision.  To reach relative errors of 1e-28..1e-30 at kappa up to ~1e22 the
// residual b - S x is formed here in quad-double (dd x dd products are exact as 8 doubles,
// accumulated error-free in a cascaded 4-limb accumulator), x is kept in dd, and the
// factor is chosen by what the condition number needs:
//   HP_FACTOR_F64       binary64 Cholesky of S.hi (kappa * 2^-53 < 1), f64 correction solves
//   HP_FACTOR_DD        dd cholesky_blocked (reference order, VALU), dd correction solves
//   HP_FACTOR_DD_SPLIT  dd blocked Cholesky whose trailing updates P P^T run on the INT8
//                       matrix cores (split GEMM), dd correction solves
// Each refinement step contracts the error by ~kappa * u_factor.
#include <cfloat>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include "hp_chol.hpp"
#include "hp_internal.hpp"

namespace hp {

template <class A>
int chol_exact(Mat a, long n, long nb, bool unblocked, Mat S, ErrRec* err, double* inv_h, double* inv_l, hipStream_t s);
template <class A>
int chol_panel_exact(Mat a, Mat S, long c0, long jb, long n, ErrRec* err, double* inv_h, double* inv_l, hipStream_t s);
template <class A>
int fwd_exact(Mat l, double* yh, double* yl, long n, hipStream_t s);

namespace {

constexpr int BW = 64;

__device__ __forceinline__ double rl(double v, int lane) {
  long long b = __builtin_bit_cast(long long, v);
  int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), lane);
  int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
  return __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned int)lo);
}
__device__ __forceinline__ dd rl(dd v, int lane) { return mkdd(rl(v.hi, lane), rl(v.lo, lane)); }

// ---------------------------------------------------------------- backward solve (fast)
// L^T x = y, right-looking in 64-row blocks from the bottom: the block's diagonal solve in
// one wavefront, then every row above takes the block's contributions (reads of row j of
// L are coalesced across the rows i).  Not the reference's order; used for corrections.
template <class A>
__global__ __launch_bounds__(64) void k_bwd_diag(Mat l, double* yh, double* yl, long b0, int bw) {
  using T = typename A::T;
  __shared__ double Lh[BW][BW + 1], Ll[A::LIMBS == 2 ? BW : 1][BW + 1];
  const int lane = threadIdx.x;
  for (int j = 0; j < bw; ++j)
    if (lane <= j) {
      T v = A::ld(l, b0 + j, b0 + lane);
      if constexpr (A::LIMBS == 2) { Lh[j][lane] = v.hi; Ll[j][lane] = v.lo; }
      else Lh[j][lane] = v;
    }
  __syncthreads();
  T y = lane < bw ? A::ldv(yh, yl, b0 + lane) : A::zero();
  for (int j = bw - 1; j >= 0; --j) {
    T ljj;
    if constexpr (A::LIMBS == 2) ljj = mkdd(Lh[j][j], Ll[j][j]);
    else ljj = Lh[j][j];
    T xj = A::div(y, ljj);  // meaningful on lane j
    if (lane == j) y = xj;
    T b = rl(xj, j);
    if (lane < j) {
      T lv;
      if constexpr (A::LIMBS == 2) lv = mkdd(Lh[j][lane], Ll[j][lane]);
      else lv = Lh[j][lane];
      y = A::sub(y, A::mul(lv, b));
    }
  }
  if (lane < bw) A::stv(yh, yl, b0 + lane, y);
}
template <class A>
__global__ __launch_bounds__(256) void k_bwd_update(Mat l, double* yh, double* yl, long b0, int bw) {
  using T = typename A::T;
  __shared__ double Xh[BW], Xl[BW];
  if (threadIdx.x < bw) {
    Xh[threadIdx.x] = yh[b0 + threadIdx.x];
    if constexpr (A::LIMBS == 2) Xl[threadIdx.x] = yl[b0 + threadIdx.x];
  }
  __syncthreads();
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= b0) return;
  T y = A::ldv(yh, yl, i);
  for (int j = bw - 1; j >= 0; --j) {
    T xv;
    if constexpr (A::LIMBS == 2) xv = mkdd(Xh[j], Xl[j]);
    else xv = Xh[j];
    y = A::sub(y, A::mul(A::ld(l, b0 + j, i), xv));
  }
  A::stv(yh, yl, i, y);
}

// ---------------------------------------------------------------- qd residual
// r_i = b_i - sum_j S_ij x_j with S, x, b in dd.  Each product is split exactly into 8
// doubles (TwoProd of the four limb pairs) and added into a 4-limb cascaded accumulator
// (TwoSum carries the error down, only the last limb rounds), then the 64 lanes' partial
// sums are combined the same way.  One wavefront per row.
__device__ __forceinline__ void acc_add(double (&a)[4], double t) {
  double s, e;
  two_sum(a[0], t, s, e); a[0] = s;
  two_sum(a[1], e, s, e); a[1] = s;
  two_sum(a[2], e, s, e); a[2] = s;
  a[3] = a[3] + e;
}
__global__ __launch_bounds__(256) void k_resid_qd(const double* __restrict__ sh, const double* __restrict__ sl,
                                                  const double* __restrict__ xh, const double* __restrict__ xl,
                                                  const double* __restrict__ bh, const double* __restrict__ bl,
                                                  double* __restrict__ rh, double* __restrict__ rl_, long n) {
  const long row = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (row >= n) return;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  const double* rsh = sh + row * n;
  const double* rsl = sl + row * n;
  for (long j = lane; j < n; j += 64) {
    double a0 = rsh[j], a1 = rsl[j], b0 = xh[j], b1 = xl[j];
    double p, e;
    two_prod(a0, b0, p, e); acc_add(acc, p); acc_add(acc, e);
    two_prod(a0, b1, p, e); acc_add(acc, p); acc_add(acc, e);
    two_prod(a1, b0, p, e); acc_add(acc, p); acc_add(acc, e);
    two_prod(a1, b1, p, e); acc_add(acc, p); acc_add(acc, e);
  }
  for (int off = 32; off > 0; off >>= 1) {
    double o[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) o[q] = __shfl_down(acc[q], off, 64);
    if (lane < off) {
#pragma unroll
      for (int q = 0; q < 4; ++q) acc_add(acc, o[q]);
    }
  }
  if (lane == 0) {
    double r[4] = {-acc[0], -acc[1], -acc[2], -acc[3]};
    acc_add(r, bh[row]);
    acc_add(r, bl[row]);
    mf_renormalize(r);
    rh[row] = r[0];
    rl_[row] = r[1];
  }
}

// ---------------------------------------------------------------- vector helpers
__global__ void k_dd_to_f64(const double* h, const double* l, double* o, long n) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) o[i] = h[i] + l[i];
}
__global__ void k_dd_add_f64(double* xh, double* xl, const double* d, long n) {  // x += d (dd += f64)
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    dd r = dd_add_f(mkdd(xh[i], xl[i]), d[i]);
    xh[i] = r.hi;
    xl[i] = r.lo;
  }
}
__global__ void k_dd_add(double* xh, double* xl, const double* dh, const double* dl, long n) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    dd r = dd_add(mkdd(xh[i], xl[i]), mkdd(dh[i], dl[i]));
    xh[i] = r.hi;
    xl[i] = r.lo;
  }
}
// max |v| over n doubles into out (as ordered bit patterns of non-negative doubles)
__global__ void k_absmax(const double* v, long n, unsigned long long* out) {
  __shared__ double red[256];
  double m = 0.0;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    m = fmax(m, fabs(v[i]));
  red[threadIdx.x] = m;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicMax(out, __builtin_bit_cast(unsigned long long, red[0]));
}
// a[r0+i][r0+j] -= c[i][j] (j <= i) -- the trailing update of the split factorisation
__global__ void k_sub_lower(double* ah, double* al, long lda, long r0, const double* ch, const double* cl, long m) {
  long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  for (; idx < m * m; idx += (long)gridDim.x * blockDim.x) {
    long i = idx / m, j = idx % m;
    if (j > i) continue;
    long o = (r0 + i) * lda + r0 + j;
    dd r = dd_sub(mkdd(ah[o], al[o]), mkdd(ch[idx], cl[idx]));
    ah[o] = r.hi;
    al[o] = r.lo;
  }
}
unsigned g1(long n) { return (unsigned)std::max<long>(1, std::min<long>((n + 255) / 256, 65536)); }

template <class A>
int bwd_fast(Mat l, double* yh, double* yl, long n, hipStream_t s) {
  long nblk = (n + BW - 1) / BW;
  for (long b = nblk - 1; b >= 0; --b) {
    long b0 = b * BW;
    int bw = (int)std::min<long>(BW, n - b0);
    hipLaunchKernelGGL(k_bwd_diag<A>, dim3(1), dim3(64), 0, s, l, yh, yl, b0, bw);
    HP_LAUNCH_CHECK();
    if (b0 > 0) {
      hipLaunchKernelGGL(k_bwd_update<A>, dim3((unsigned)((b0 + 255) / 256)), dim3(256), 0, s, l, yh, yl, b0, bw);
      HP_LAUNCH_CHECK();
    }
  }
  return HP_OK;
}

double absmax_dev(const double* v, long n, unsigned long long* scratch, hipStream_t s, int* rc) {
  *rc = HP_OK;
  if (hipMemsetAsync(scratch, 0, sizeof(unsigned long long), s) != hipSuccess) { *rc = HP_EDEVICE; return 0; }
  hipLaunchKernelGGL(k_absmax, dim3(std::min<unsigned>(g1(n), 1024)), dim3(256), 0, s, v, n, scratch);
  unsigned long long h = 0;
  if (hipMemcpyAsync(&h, scratch, sizeof(h), hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess) {
    *rc = HP_EDEVICE;
    set_error("absmax: HIP error");
    return 0;
  }
  return __builtin_bit_cast(double, h);
}

}  // namespace

// dd blocked Cholesky with the trailing update P P^T on the INT8 matrix cores.
int chol_split(Mat a, long n, long nb, Mat S, ErrRec* err, double* inv_h, double* inv_l, double* tmp, hipStream_t s) {
  if (nb <= 0) nb = 256;
  for (long c0 = 0; c0 < n; c0 += nb) {
    long jb = std::min(nb, n - c0);
    int rc = chol_panel_exact<ArDD>(a, S, c0, jb, n, err, inv_h, inv_l, s);
    if (rc) return rc;
    long rem = n - c0 - jb;
    if (rem <= 0) continue;
    ErrRec e;
    HP_HIP(hipMemcpyAsync(&e, err, sizeof(e), hipMemcpyDeviceToHost, s));
    HP_HIP(hipStreamSynchronize(s));
    if (e.flag) return HP_OK;  // reported by the caller
    SplitOperand P{a.hi + (c0 + jb) * a.ld + c0, a.lo + (c0 + jb) * a.ld + c0, a.ld, true};
    double* th = tmp;
    double* tl = tmp + (size_t)rem * rem;
    if ((rc = split_gemm_general(P, P, th, tl, (size_t)rem, (size_t)rem, (size_t)rem, (size_t)jb, 31, s))) return rc;
    hipLaunchKernelGGL(k_sub_lower, dim3(g1(rem * rem)), dim3(256), 0, s, a.hi, a.lo, a.ld, c0 + jb, th, tl, rem);
    HP_LAUNCH_CHECK();
  }
  return HP_OK;
}

int solve_refined(int kind, const double* sh, const double* sl, const double* bh, const double* bl, long n, double tol,
                  int max_iter, double* xh, double* xl, hp_refine_stats* st, hipStream_t s) {
  using clk = std::chrono::steady_clock;
  std::memset(st, 0, sizeof(*st));
  st->factor_kind = kind;
  if (n == 0) return HP_OK;
  const size_t nn = (size_t)n * n;
  DevBuf L, Sw, tmp, inv, err, vec, sc;
  int rc;
  const int limbs = (kind == HP_FACTOR_F64) ? 1 : 2;
  if ((rc = L.alloc(nn * limbs * sizeof(double))) || (rc = Sw.alloc(nn * limbs * sizeof(double))) ||
      (rc = inv.alloc(2 * 64 * sizeof(double))) || (rc = err.alloc(sizeof(ErrRec))) ||
      (rc = vec.alloc(6 * n * sizeof(double))) || (rc = sc.alloc(64)))
    return rc;
  if (kind == HP_FACTOR_DD_SPLIT && (rc = tmp.alloc(2 * nn * sizeof(double)))) return rc;
  HP_HIP(hipMemsetAsync(err.p, 0, sizeof(ErrRec), s));
  HP_HIP(hipStreamSynchronize(s));
  auto t0 = clk::now();
  Mat Lm{L.d(), limbs == 2 ? L.d() + nn : nullptr, n};
  Mat Sm{Sw.d(), limbs == 2 ? Sw.d() + nn : nullptr, n};
  HP_HIP(hipMemcpyAsync(Lm.hi, sh, nn * sizeof(double), hipMemcpyDeviceToDevice, s));
  if (limbs == 2) HP_HIP(hipMemcpyAsync(Lm.lo, sl, nn * sizeof(double), hipMemcpyDeviceToDevice, s));
  ErrRec* er = (ErrRec*)err.p;
  switch (kind) {
    case HP_FACTOR_F64: rc = chol_exact<ArF64>(Lm, n, 64, false, Sm, er, inv.d(), inv.d() + 64, s); break;
    case HP_FACTOR_DD: rc = chol_exact<ArDD>(Lm, n, 64, false, Sm, er, inv.d(), inv.d() + 64, s); break;
    case HP_FACTOR_DD_SPLIT: rc = chol_split(Lm, n, 256, Sm, er, inv.d(), inv.d() + 64, tmp.d(), s); break;
    default: set_error("solve_refined: unknown factor kind %d", kind); return HP_EINVAL;
  }
  if (rc) return rc;
  ErrRec e;
  HP_HIP(hipMemcpyAsync(&e, er, sizeof(e), hipMemcpyDeviceToHost, s));
  HP_HIP(hipStreamSynchronize(s));
  auto t1 = clk::now();
  st->factor_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
  if (e.flag) {
    set_error("Cholesky failed at pivot %lld (value=%.4e)", e.pivot, e.value);
    st->pivot = (size_t)e.pivot;
    return HP_ENOTSPD;
  }
  double* rh = vec.d();
  double* rlo = rh + n;
  double* dh = rh + 2 * n;
  double* dl = rh + 3 * n;
  double* d64 = rh + 4 * n;
  unsigned long long* scr = (unsigned long long*)sc.p;
  HP_HIP(hipMemsetAsync(xh, 0, n * sizeof(double), s));
  HP_HIP(hipMemsetAsync(xl, 0, n * sizeof(double), s));
  double last = 1.0;
  int it = 0;
  for (; it < max_iter; ++it) {
    if (it == 0) {
      HP_HIP(hipMemcpyAsync(rh, bh, n * sizeof(double), hipMemcpyDeviceToDevice, s));
      HP_HIP(hipMemcpyAsync(rlo, bl, n * sizeof(double), hipMemcpyDeviceToDevice, s));
    } else {
      hipLaunchKernelGGL(k_resid_qd, dim3((unsigned)((n + 3) / 4)), dim3(256), 0, s, sh, sl, xh, xl, bh, bl, rh, rlo, n);
      HP_LAUNCH_CHECK();
    }
    double dn;
    if (kind == HP_FACTOR_F64) {
      hipLaunchKernelGGL(k_dd_to_f64, dim3(g1(n)), dim3(256), 0, s, rh, rlo, d64, n);
      HP_LAUNCH_CHECK();
      if ((rc = fwd_exact<ArF64>(Lm, d64, nullptr, n, s)) || (rc = bwd_fast<ArF64>(Lm, d64, nullptr, n, s))) return rc;
      hipLaunchKernelGGL(k_dd_add_f64, dim3(g1(n)), dim3(256), 0, s, xh, xl, d64, n);
      HP_LAUNCH_CHECK();
      dn = absmax_dev(d64, n, scr, s, &rc);
    } else {
      HP_HIP(hipMemcpyAsync(dh, rh, n * sizeof(double), hipMemcpyDeviceToDevice, s));
      HP_HIP(hipMemcpyAsync(dl, rlo, n * sizeof(double), hipMemcpyDeviceToDevice, s));
      if ((rc = fwd_exact<ArDD>(Lm, dh, dl, n, s)) || (rc = bwd_fast<ArDD>(Lm, dh, dl, n, s))) return rc;
      hipLaunchKernelGGL(k_dd_add, dim3(g1(n)), dim3(256), 0, s, xh, xl, dh, dl, n);
      HP_LAUNCH_CHECK();
      dn = absmax_dev(dh, n, scr, s, &rc);
    }
    if (rc) return rc;
    double xn = absmax_dev(xh, n, scr, s, &rc);
    if (rc) return rc;
    last = xn > 0 ? dn / xn : dn;
    if (last <= tol) { ++it; break; }
  }
  auto t2 = clk::now();
  st->iterations = it;
  st->last_correction = last;
  st->refine_ms = std::chrono::duration<double, std::milli>(t2 - t1).count();
  return HP_OK;
}

}  // namespace hp
