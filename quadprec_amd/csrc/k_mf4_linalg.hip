// k_mf4_linalg.hip -- MultiFloat<4> (f256) reciprocal, division and square root
// (crates/quad/src/multifloat/div.rs:14-77, math.rs:22-40) and the MF linear algebra of
// multifloat/linalg.rs: mf_cholesky (89-141), mf_cholesky_solve (145-172), mf_gemv (17-36).
// Every element follows the reference's operation order, so results are bit-identical:
//   * mf_cholesky, right-looking: S_ij accumulates L_ik L_jk for k = 0, 1, ... from ZERO,
//     the same sums the reference forms row by row (`s = s + row[k] * rj[k]`,
//     `sum = sum + ljk * ljk`); column k then takes diag = a_kk - S_kk, l_kk = sqrt(diag),
//     l_ik = (a_ik - S_ik) * recip(l_kk).  Per column: one single-block launch for the
//     column, one launch for the trailing update (one MF multiply-add per element, every
//     element in parallel).  The upper triangle keeps A.
//   * forward solve: y_j / l_jj, then every y_i (i > j) takes - l_ij y_j: the per-element
//     order of the reference (j ascending), one block.
//   * backward solve: x_i - l_ji x_j for j = i+1 .. n-1 ascending starts from x_{i+1}, the
//     last value formed, so the chain is serial: the products are formed in parallel into
//     LDS and one thread runs the subtractions (as the dd backward solve does).
//   * the solves' divisions by l_ii are div.rs's special cases and a multiply by
//     recip(l_ii); the reciprocals, a function of L alone, are formed first for every row in
//     parallel, which takes 2n reciprocal chains off the serial path.
//   * gemv: one thread per row, s = s + a_ij x_j in j order, 64 x 16 tiles through LDS.
// Matrices are limb planes on the device (plane l at offset l * rows * cols).
#include <algorithm>
#include <vector>

#include "hp_arith.hpp"
#include "hp_internal.hpp"

namespace hp {
namespace {

struct MfMat {
  double* p;
  long rows, cols;
  __device__ __forceinline__ mf4 ld(long i, long j) const {
    const long o = i * cols + j, pl = rows * cols;
    return mkmf(p[o], p[o + pl], p[o + 2 * pl], p[o + 3 * pl]);
  }
  __device__ __forceinline__ void st(long i, long j, mf4 v) const {
    const long o = i * cols + j, pl = rows * cols;
    p[o] = v.l0;
    p[o + pl] = v.l1;
    p[o + 2 * pl] = v.l2;
    p[o + 3 * pl] = v.l3;
  }
};
__device__ __forceinline__ mf4 ldv(const double* v, long n, long i) {
  return mkmf(v[i], v[i + n], v[i + 2 * n], v[i + 3 * n]);
}
__device__ __forceinline__ void stv(double* v, long n, long i, mf4 x) {
  v[i] = x.l0;
  v[i + n] = x.l1;
  v[i + 2 * n] = x.l2;
  v[i + 3 * n] = x.l3;
}

struct CholErr {
  int flag;
  int pad;
  long pivot;
};

// column k: the pivot on thread 0 (shared through LDS), then l_ik for i > k
__global__ __launch_bounds__(256) void k_mfc_col(MfMat a, MfMat S, long n, long k, CholErr* err) {
  __shared__ double inv[4];
  __shared__ int bad;
  if (err->flag) return;
  if (threadIdx.x == 0) {
    const mf4 diag = mf_sub(a.ld(k, k), S.ld(k, k));
    bad = diag.l0 <= 0.0;  // multifloat/linalg.rs:103 (a NaN pivot passes, as in the reference)
    if (bad) {
      err->flag = 1;
      err->pivot = k;
    } else {
      const mf4 ljj = mf_sqrt(diag);
      a.st(k, k, ljj);
      const mf4 r = mf_recip(ljj);
      inv[0] = r.l0;
      inv[1] = r.l1;
      inv[2] = r.l2;
      inv[3] = r.l3;
    }
  }
  __syncthreads();
  if (bad) return;
  const mf4 r = mkmf(inv[0], inv[1], inv[2], inv[3]);
  for (long i = k + 1 + threadIdx.x; i < n; i += blockDim.x) a.st(i, k, mf_mul(mf_sub(a.ld(i, k), S.ld(i, k)), r));
}

// S_ij += l_ik l_jk for k < j <= i < n (16 x 16 tiles over the trailing square)
__global__ __launch_bounds__(256) void k_mfc_update(MfMat a, MfMat S, long n, long k, const CholErr* err) {
  if (err->flag) return;
  const long i = k + 1 + (long)blockIdx.y * 16 + (threadIdx.x >> 4);
  const long j = k + 1 + (long)blockIdx.x * 16 + (threadIdx.x & 15);
  if (i >= n || j > i) return;
  S.st(i, j, mf_add(S.ld(i, j), mf_mul(a.ld(i, k), a.ld(j, k))));
}

// mf_div(a, b) given r = mf_recip(b): div.rs:44-77's special cases, then a * r -- the same
// operations, with the reciprocal (a function of b alone) formed ahead
__device__ __forceinline__ mf4 mf_div_by(mf4 a, mf4 b, mf4 r) {
  if (b.l0 == 0.0) {
    const double sign = (!signbit(a.l0) == !signbit(b.l0)) ? 1.0 : -1.0;
    return mkmf(a.l0 == 0.0 ? (double)NAN : sign * (double)INFINITY);
  }
  if (a.l0 == 0.0) return mkmf(0.0);
  if (!isfinite(a.l0) || !isfinite(b.l0)) return mkmf(a.l0 / b.l0);
  return mf_mul(a, r);
}
// the reciprocals of L's diagonal, every row in parallel (the solves' divisors)
__global__ void k_mf_diag_recip(MfMat l, double* r, long n) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) stv(r, n, i, mf_recip(l.ld(i, i)));
}

// forward solve L y = b in place (y holds b), one block
__global__ __launch_bounds__(1024) void k_mf_fwd(MfMat l, double* y, const double* rd, long n) {
  for (long j = 0; j < n; ++j) {
    if (threadIdx.x == 0) stv(y, n, j, mf_div_by(ldv(y, n, j), l.ld(j, j), ldv(rd, n, j)));
    __syncthreads();
    const mf4 yj = ldv(y, n, j);
    for (long i = j + 1 + threadIdx.x; i < n; i += blockDim.x) stv(y, n, i, mf_sub_chain(ldv(y, n, i), mf_mul(l.ld(i, j), yj)));
    __syncthreads();
  }
}

// backward solve L^T x = y in place (x holds y), one block
__global__ __launch_bounds__(256) void k_mf_bwd(MfMat l, double* x, const double* rd, long n) {
  __shared__ double P[4][256];
  for (long i = n - 1; i >= 0; --i) {
    mf4 s = mkmf(0.0);
    if (threadIdx.x == 0) s = ldv(x, n, i);
    for (long jc = i + 1; jc < n; jc += 256) {
      const long j = jc + threadIdx.x;
      if (j < n) {
        const mf4 p = mf_mul(l.ld(j, i), ldv(x, n, j));
        P[0][threadIdx.x] = p.l0;
        P[1][threadIdx.x] = p.l1;
        P[2][threadIdx.x] = p.l2;
        P[3][threadIdx.x] = p.l3;
      }
      __syncthreads();
      if (threadIdx.x == 0) {
        const int cnt = (int)std::min(256L, n - jc);
        for (int q = 0; q < cnt; ++q) s = mf_sub_chain(s, mkmf(P[0][q], P[1][q], P[2][q], P[3][q]));
      }
      __syncthreads();
    }
    if (threadIdx.x == 0) stv(x, n, i, mf_div_by(s, l.ld(i, i), ldv(rd, n, i)));
    __syncthreads();
  }
}

// y = A x, A n x m: one thread per row, 64 x 16 tiles of A (and x) staged through LDS
__global__ __launch_bounds__(64) void k_mf_gemv(MfMat a, const double* x, double* y) {
  __shared__ double T[4][64][17];
  __shared__ double X[4][16];
  const int t = threadIdx.x;
  const long n = a.rows, m = a.cols, r0 = (long)blockIdx.x * 64;
  mf4 s = mkmf(0.0);
  for (long c0 = 0; c0 < m; c0 += 16) {
    const int w = (int)std::min(16L, m - c0);
    for (int e = t; e < 64 * 16; e += 64) {
      const int r = e >> 4, c = e & 15;
      const bool ok = r0 + r < n && c < w;
      const mf4 v = ok ? a.ld(r0 + r, c0 + c) : mkmf(0.0);
      T[0][r][c] = v.l0;
      T[1][r][c] = v.l1;
      T[2][r][c] = v.l2;
      T[3][r][c] = v.l3;
    }
    if (t < 16) {
      const mf4 v = t < w ? ldv(x, m, c0 + t) : mkmf(0.0);
      X[0][t] = v.l0;
      X[1][t] = v.l1;
      X[2][t] = v.l2;
      X[3][t] = v.l3;
    }
    __syncthreads();
    for (int c = 0; c < w; ++c)
      s = mf_add(s, mf_mul(mkmf(T[0][t][c], T[1][t][c], T[2][t][c], T[3][t][c]), mkmf(X[0][c], X[1][c], X[2][c], X[3][c])));
    __syncthreads();
  }
  if (r0 + t < n) stv(y, n, r0 + t, s);
}

// elementwise: 0 recip(a), 1 sqrt(a), 2 a / b
__global__ void k_mf_elem(int op, const double* a, const double* b, double* out, long n) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const mf4 x = ldv(a, n, i);
  mf4 r;
  if (op == 0) r = mf_recip(x);
  else if (op == 1) r = mf_sqrt(x);
  else r = mf_div(x, ldv(b, n, i));
  stv(out, n, i, r);
}

// elementwise with the generic MultiFloat<N> op codes: 0 add, 1 sub, 2 mul, 3 div,
// 4 recip(a), 5 sqrt(a), 6 renormalize(a)
__global__ void k_mf4_elem_n(int op, const double* a, const double* b, double* out, long n) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const mf4 x = ldv(a, n, i);
  mf4 r = x;
  switch (op) {
    case 0: r = mf_add(x, ldv(b, n, i)); break;
    case 1: r = mf_sub(x, ldv(b, n, i)); break;
    case 2: r = mf_mul(x, ldv(b, n, i)); break;
    case 3: r = mf_div(x, ldv(b, n, i)); break;
    case 4: r = mf_recip(x); break;
    case 5: r = mf_sqrt(x); break;
    default: {
      double l[4] = {x.l0, x.l1, x.l2, x.l3};
      mf_renormalize(l);
      r = mkmf(l[0], l[1], l[2], l[3]);
    }
  }
  stv(out, n, i, r);
}

// MultiFloat<4>::from_f128 / to_f128 (multifloat/conv.rs:120-131) over arrays: limbs
// (hi, lo, 0, 0), and (limb 0, limb 1) with limbs 2 and 3 dropped; AoS in and out
__global__ void k_mf4_from_dd(const double2* __restrict__ in, double4* __restrict__ out, long n) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = make_double4(in[i].x, in[i].y, 0.0, 0.0);
}
__global__ void k_dd_from_mf4(const double4* __restrict__ in, double2* __restrict__ out, long n) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = make_double2(in[i].x, in[i].y);
}

unsigned blocks_of(long n, int t) { return (unsigned)((n + t - 1) / t); }

}  // namespace

int launch_mfn_elem4(int op, const double* a, const double* b, double* out, size_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_mf4_elem_n, dim3(blocks_of((long)n, 256)), dim3(256), 0, s, op, a, b, out, (long)n);
  HP_LAUNCH_CHECK();
  return HP_OK;
}

}  // namespace hp

using namespace hp;

namespace {
int null_arg() {
  set_error("null pointer argument");
  return HP_EINVAL;
}
}  // namespace

extern "C" {

int hp_mf_cholesky(double* a, size_t n, size_t* pivot) {
  if (n == 0) return HP_OK;
  if (!a) return null_arg();
  hipStream_t s = hipStreamPerThread;
  const size_t nn = n * n;
  DevBuf da, dS, derr;
  int rc;
  if ((rc = da.alloc(4 * nn * sizeof(double))) || (rc = dS.alloc(4 * nn * sizeof(double))) ||
      (rc = derr.alloc(sizeof(CholErr))))
    return rc;
  if ((rc = upload_planes(a, nn, 4, da.d(), s))) return rc;
  HP_HIP(hipMemsetAsync(dS.p, 0, 4 * nn * sizeof(double), s));
  HP_HIP(hipMemsetAsync(derr.p, 0, sizeof(CholErr), s));
  const MfMat A{da.d(), (long)n, (long)n}, S{dS.d(), (long)n, (long)n};
  auto* err = static_cast<CholErr*>(derr.p);
  for (long k = 0; k < (long)n; ++k) {
    hipLaunchKernelGGL(k_mfc_col, dim3(1), dim3(256), 0, s, A, S, (long)n, k, err);
    const long t = (long)n - k - 1;
    if (t > 0) hipLaunchKernelGGL(k_mfc_update, dim3(blocks_of(t, 16), blocks_of(t, 16)), dim3(256), 0, s, A, S,
                                  (long)n, k, static_cast<const CholErr*>(err));
  }
  HP_LAUNCH_CHECK();
  CholErr he;
  HP_HIP(hipMemcpyAsync(&he, err, sizeof(he), hipMemcpyDeviceToHost, s));
  HP_HIP(hipStreamSynchronize(s));
  if ((rc = download_planes(da.d(), nn, 4, a, s))) return rc;  // on failure: the columns before the pivot, as
  if (he.flag) {                                                // the reference leaves them
    if (pivot) *pivot = (size_t)he.pivot;
    set_error("Cholesky failed: matrix is not positive definite (pivot %ld)", he.pivot);
    return HP_ENOTSPD;
  }
  return HP_OK;
}

int hp_mf_cholesky_solve(const double* l, const double* b, double* x, size_t n) {
  if (n == 0) return HP_OK;
  if (!l || !b || !x) return null_arg();
  hipStream_t s = hipStreamPerThread;
  const size_t nn = n * n;
  DevBuf dl, dx, dr;
  int rc;
  if ((rc = dl.alloc(4 * nn * sizeof(double))) || (rc = dx.alloc(4 * n * sizeof(double))) ||
      (rc = dr.alloc(4 * n * sizeof(double))))
    return rc;
  if ((rc = upload_planes(l, nn, 4, dl.d(), s)) || (rc = upload_planes(b, n, 4, dx.d(), s))) return rc;
  const MfMat L{dl.d(), (long)n, (long)n};
  hipLaunchKernelGGL(k_mf_diag_recip, dim3(blocks_of((long)n, 64)), dim3(64), 0, s, L, dr.d(), (long)n);
  hipLaunchKernelGGL(k_mf_fwd, dim3(1), dim3(1024), 0, s, L, dx.d(), static_cast<const double*>(dr.d()), (long)n);
  hipLaunchKernelGGL(k_mf_bwd, dim3(1), dim3(256), 0, s, L, dx.d(), static_cast<const double*>(dr.d()), (long)n);
  HP_LAUNCH_CHECK();
  return download_planes(dx.d(), n, 4, x, s);
}

int hp_mf_gemv(const double* a, const double* x, double* y, size_t n, size_t m) {
  if (n == 0) return HP_OK;
  if (!a || !x || !y) return null_arg();
  hipStream_t s = hipStreamPerThread;
  DevBuf da, dx, dy;
  int rc;
  if ((rc = da.alloc(4 * n * m * sizeof(double))) || (rc = dx.alloc(4 * m * sizeof(double))) ||
      (rc = dy.alloc(4 * n * sizeof(double))))
    return rc;
  if ((m && (rc = upload_planes(a, n * m, 4, da.d(), s))) || (m && (rc = upload_planes(x, m, 4, dx.d(), s))))
    return rc;
  hipLaunchKernelGGL(k_mf_gemv, dim3(blocks_of((long)n, 64)), dim3(64), 0, s, MfMat{da.d(), (long)n, (long)m}, dx.d(),
                     dy.d());
  HP_LAUNCH_CHECK();
  return download_planes(dy.d(), n, 4, y, s);
}

// dd <-> MF4 conversion of n values (from_f128 / to_f128), device or host AoS arrays
// through the device
int hp_mf4_from_dd(const double* dd, double* mf4, size_t n) {
  if (n == 0) return HP_OK;
  if (!dd || !mf4) return null_arg();
  hipStream_t s = hipStreamPerThread;
  DevBuf di, dout;
  int rc;
  if ((rc = di.alloc(2 * n * sizeof(double))) || (rc = dout.alloc(4 * n * sizeof(double)))) return rc;
  HP_HIP(hipMemcpyAsync(di.p, dd, 2 * n * sizeof(double), hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_mf4_from_dd, dim3(blocks_of((long)n, 256)), dim3(256), 0, s,
                     static_cast<const double2*>(di.p), static_cast<double4*>(dout.p), (long)n);
  HP_LAUNCH_CHECK();
  HP_HIP(hipMemcpyAsync(mf4, dout.p, 4 * n * sizeof(double), hipMemcpyDeviceToHost, s));
  HP_HIP(hipStreamSynchronize(s));
  return HP_OK;
}
int hp_dd_from_mf4(const double* mf4, double* dd, size_t n) {
  if (n == 0) return HP_OK;
  if (!dd || !mf4) return null_arg();
  hipStream_t s = hipStreamPerThread;
  DevBuf di, dout;
  int rc;
  if ((rc = di.alloc(4 * n * sizeof(double))) || (rc = dout.alloc(2 * n * sizeof(double)))) return rc;
  HP_HIP(hipMemcpyAsync(di.p, mf4, 4 * n * sizeof(double), hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_dd_from_mf4, dim3(blocks_of((long)n, 256)), dim3(256), 0, s,
                     static_cast<const double4*>(di.p), static_cast<double2*>(dout.p), (long)n);
  HP_LAUNCH_CHECK();
  HP_HIP(hipMemcpyAsync(dd, dout.p, 2 * n * sizeof(double), hipMemcpyDeviceToHost, s));
  HP_HIP(hipStreamSynchronize(s));
  return HP_OK;
}

// elementwise MF4 ops on n values (AoS, 4 doubles each): op 0 recip(a), 1 sqrt(a), 2 a / b
int hp_mf4_elementwise(int op, const double* a, const double* b, double* out, size_t n) {
  if (op < 0 || op > 2) {
    set_error("unknown MF4 op %d", op);
    return HP_EINVAL;
  }
  if (n == 0) return HP_OK;
  if (!a || !out || (op == 2 && !b)) return null_arg();
  hipStream_t s = hipStreamPerThread;
  DevBuf da, db, dout;
  int rc;
  if ((rc = da.alloc(4 * n * sizeof(double))) || (rc = db.alloc(4 * n * sizeof(double))) ||
      (rc = dout.alloc(4 * n * sizeof(double))))
    return rc;
  if ((rc = upload_planes(a, n, 4, da.d(), s)) || (op == 2 && (rc = upload_planes(b, n, 4, db.d(), s)))) return rc;
  hipLaunchKernelGGL(k_mf_elem, dim3(blocks_of((long)n, 256)), dim3(256), 0, s, op, da.d(), db.d(), dout.d(), (long)n);
  HP_LAUNCH_CHECK();
  return download_planes(dout.d(), n, 4, out, s);
}

}  // extern "C"
