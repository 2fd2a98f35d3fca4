// k_mfn.hip -- MultiFloat<N> for N = 1..8 (f512 = MultiFloat<8> among them): the generic
// element operations and linear algebra of crates/quad/src/multifloat, bit-identical to
// the reference's per-element order.
//   * mf_gemm / mf_gemm_atb (multifloat/linalg.rs:39-85): one element per thread, 16 x 16
//     tiles, 16-deep k slabs of the N limb planes through LDS (as the MultiFloat<4> kernel).
//   * mf_gemv (17-36) and mf_dot (7-14, the one-row gemv): one thread per row, 64 x 16
//     tiles through LDS.
//   * mf_cholesky (89-141), right-looking with the reference's per-element sums (S_ij
//     accumulates l_ik l_jk for k ascending from ZERO): per column one single-block launch
//     (diag = a_kk - S_kk, sqrt, recip, the column) and one trailing-update launch.
//   * mf_cholesky_solve (145-172): the forward solve on one block (y_j / l_jj, then every
//     later y_i takes - l_ij y_j), the backward solve with the products formed in parallel
//     and the serial subtraction chain on one thread (x_i's terms run j ascending from
//     x_{i+1}); divisions are div.rs's special cases and a multiply by the reciprocal, the
//     reciprocals of L's diagonal formed first in parallel.
// MultiFloat<4> requests go to the tuned f256 kernels (k_mf4_gemm.hip, k_mf4_linalg.hip),
// which give the same bits.  Device matrices are limb planes: plane l of an r x c matrix
// at offset l * r * c.
#include <algorithm>

#include "hp_internal.hpp"
#include "hp_mfn.hpp"

namespace hp {
namespace {

constexpr int TB = 16, KB = 16;

template <int N>
__device__ __forceinline__ mfn<N> ldp(const double* p, long plane, long o) {
  mfn<N> r;
#pragma unroll
  for (int l = 0; l < N; ++l) r.l[l] = p[o + l * plane];
  return r;
}
template <int N>
__device__ __forceinline__ void stp(double* p, long plane, long o, const mfn<N>& v) {
#pragma unroll
  for (int l = 0; l < N; ++l) p[o + l * plane] = v.l[l];
}

template <int N, bool TRANS_A>
__global__ __launch_bounds__(256) void k_mfn_gemm(const double* __restrict__ a, const double* __restrict__ b,
                                                  double* __restrict__ c, int m, int n, int k) {
  __shared__ double As[N][KB][TB + 1];
  __shared__ double Bs[N][KB][TB + 1];
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
  const int i = blockIdx.y * TB + ty, j = blockIdx.x * TB + tx;
  const size_t pa = (size_t)m * k, pb = (size_t)k * n, pc = (size_t)m * n;
  mfn<N> s = mfn_from<N>(0.0);
  for (int p0 = 0; p0 < k; p0 += KB) {
    {
      const int gi = blockIdx.y * TB + ty, gp = p0 + tx;
      const bool ok = gi < m && gp < k;
      const size_t off = TRANS_A ? (size_t)gp * m + gi : (size_t)gi * k + gp;
#pragma unroll
      for (int l = 0; l < N; ++l) As[l][tx][ty] = ok ? a[l * pa + off] : 0.0;
      const int gp2 = p0 + ty;
      const bool ok2 = j < n && gp2 < k;
      const size_t off2 = (size_t)gp2 * n + j;
#pragma unroll
      for (int l = 0; l < N; ++l) Bs[l][ty][tx] = ok2 ? b[l * pb + off2] : 0.0;
    }
    __syncthreads();
    const int kmax = min(KB, k - p0);
    for (int kk = 0; kk < kmax; ++kk) {
      mfn<N> av, bv;
#pragma unroll
      for (int l = 0; l < N; ++l) {
        av.l[l] = As[l][kk][ty];
        bv.l[l] = Bs[l][kk][tx];
      }
      s = mfn_add<N>(s, mfn_mul<N>(av, bv));
    }
    __syncthreads();
  }
  if (i < m && j < n) stp<N>(c, (long)pc, (long)i * n + j, s);
}

// y = A x, A n x m (one thread per row, 64 x 16 tiles of A and x through LDS)
template <int N>
__global__ __launch_bounds__(64) void k_mfn_gemv(const double* __restrict__ a, const double* __restrict__ x,
                                                 double* __restrict__ y, long n, long m) {
  __shared__ double T[N][64][17];
  __shared__ double X[N][16];
  const int t = threadIdx.x;
  const long r0 = (long)blockIdx.x * 64, pa = n * m;
  mfn<N> s = mfn_from<N>(0.0);
  for (long c0 = 0; c0 < m; c0 += 16) {
    const int w = (int)std::min(16L, m - c0);
    for (int e = t; e < 64 * 16; e += 64) {
      const int r = e >> 4, c = e & 15;
      const bool ok = r0 + r < n && c < w;
#pragma unroll
      for (int l = 0; l < N; ++l) T[l][r][c] = ok ? a[(r0 + r) * m + c0 + c + l * pa] : 0.0;
    }
    if (t < 16) {
#pragma unroll
      for (int l = 0; l < N; ++l) X[l][t] = t < w ? x[c0 + t + l * m] : 0.0;
    }
    __syncthreads();
    for (int c = 0; c < w; ++c) {
      mfn<N> av, xv;
#pragma unroll
      for (int l = 0; l < N; ++l) {
        av.l[l] = T[l][t][c];
        xv.l[l] = X[l][c];
      }
      s = mfn_add<N>(s, mfn_mul<N>(av, xv));
    }
    __syncthreads();
  }
  if (r0 + t < n) stp<N>(y, n, r0 + t, s);
}

// elementwise, op: 0 add, 1 sub, 2 mul, 3 div, 4 recip(a), 5 sqrt(a), 6 renormalize(a)
template <int N>
__global__ __launch_bounds__(128) void k_mfn_elem(int op, const double* __restrict__ a, const double* __restrict__ b,
                           double* __restrict__ out, long n) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const mfn<N> x = ldp<N>(a, n, i);
  mfn<N> r = x;
  switch (op) {
    case 0: r = mfn_add<N>(x, ldp<N>(b, n, i)); break;
    case 1: r = mfn_sub<N>(x, ldp<N>(b, n, i)); break;
    case 2: r = mfn_mul<N>(x, ldp<N>(b, n, i)); break;
    case 3: r = mfn_div<N>(x, ldp<N>(b, n, i)); break;
    case 4: r = mfn_recip<N>(x); break;
    case 5: r = mfn_sqrt<N>(x); break;
    default: mfn_renormalize<N>(r.l); break;
  }
  stp<N>(out, n, i, r);
}

struct MfnErr {
  int flag;
  int pad;
  long pivot;
};

// column k of the factor: the pivot on thread 0, then l_ik = (a_ik - S_ik) * recip(l_kk)
template <int N>
__global__ __launch_bounds__(256) void k_mfnc_col(double* a, const double* S, long n, long k, MfnErr* err) {
  __shared__ double inv[N];
  __shared__ int bad;
  if (err->flag) return;
  const long pl = n * n;
  if (threadIdx.x == 0) {
    const mfn<N> diag = mfn_sub<N>(ldp<N>(a, pl, k * n + k), ldp<N>(S, pl, k * n + k));
    bad = diag.l[0] <= 0.0;  // multifloat/linalg.rs:103 (a NaN pivot passes, as in the reference)
    if (bad) {
      err->flag = 1;
      err->pivot = k;
    } else {
      const mfn<N> ljj = mfn_sqrt<N>(diag);
      stp<N>(a, pl, k * n + k, ljj);
      const mfn<N> r = mfn_recip<N>(ljj);
#pragma unroll
      for (int l = 0; l < N; ++l) inv[l] = r.l[l];
    }
  }
  __syncthreads();
  if (bad) return;
  mfn<N> r;
#pragma unroll
  for (int l = 0; l < N; ++l) r.l[l] = inv[l];
  for (long i = k + 1 + threadIdx.x; i < n; i += blockDim.x)
    stp<N>(a, pl, i * n + k, mfn_mul<N>(mfn_sub<N>(ldp<N>(a, pl, i * n + k), ldp<N>(S, pl, i * n + k)), r));
}
// S_ij += l_ik l_jk for k < j <= i < n (16 x 16 tiles over the trailing square)
template <int N>
__global__ __launch_bounds__(256) void k_mfnc_update(const double* a, double* S, long n, long k, const MfnErr* err) {
  if (err->flag) return;
  const long i = k + 1 + (long)blockIdx.y * 16 + (threadIdx.x >> 4);
  const long j = k + 1 + (long)blockIdx.x * 16 + (threadIdx.x & 15);
  if (i >= n || j > i) return;
  const long pl = n * n;
  stp<N>(S, pl, i * n + j,
         mfn_add<N>(ldp<N>(S, pl, i * n + j), mfn_mul<N>(ldp<N>(a, pl, i * n + k), ldp<N>(a, pl, j * n + k))));
}

// a / b with r = recip(b) formed ahead: div.rs's special cases, then a * r
template <int N>
__device__ __forceinline__ mfn<N> mfn_div_by(const mfn<N>& a, const mfn<N>& b, const mfn<N>& r) {
  mfn<N> q;
  if (mfn_div_special<N>(a, b, q)) return q;
  return mfn_mul<N>(a, r);
}
template <int N>
__global__ __launch_bounds__(64) void k_mfn_diag_recip(const double* l, double* r, long n) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) stp<N>(r, n, i, mfn_recip<N>(ldp<N>(l, n * n, i * n + i)));
}
// forward solve L y = b in place (y holds b), one block of 256 threads
template <int N>
__global__ __launch_bounds__(256) void k_mfn_fwd(const double* l, double* y, const double* rd, long n) {
  const long pl = n * n;
  for (long j = 0; j < n; ++j) {
    if (threadIdx.x == 0) stp<N>(y, n, j, mfn_div_by<N>(ldp<N>(y, n, j), ldp<N>(l, pl, j * n + j), ldp<N>(rd, n, j)));
    __syncthreads();
    const mfn<N> yj = ldp<N>(y, n, j);
    for (long i = j + 1 + threadIdx.x; i < n; i += blockDim.x)
      stp<N>(y, n, i, mfn_sub<N>(ldp<N>(y, n, i), mfn_mul<N>(ldp<N>(l, pl, i * n + j), yj)));
    __syncthreads();
  }
}
// backward solve L^T x = y in place (x holds y), one block
template <int N>
__global__ __launch_bounds__(256) void k_mfn_bwd(const double* l, double* x, const double* rd, long n) {
  __shared__ double P[N][256];
  const long pl = n * n;
  for (long i = n - 1; i >= 0; --i) {
    mfn<N> s = mfn_from<N>(0.0);
    if (threadIdx.x == 0) s = ldp<N>(x, n, i);
    for (long jc = i + 1; jc < n; jc += 256) {
      const long j = jc + threadIdx.x;
      if (j < n) {
        const mfn<N> p = mfn_mul<N>(ldp<N>(l, pl, j * n + i), ldp<N>(x, n, j));
#pragma unroll
        for (int q = 0; q < N; ++q) P[q][threadIdx.x] = p.l[q];
      }
      __syncthreads();
      if (threadIdx.x == 0) {
        const int cnt = (int)std::min(256L, n - jc);
        for (int q = 0; q < cnt; ++q) {
          mfn<N> v;
#pragma unroll
          for (int t = 0; t < N; ++t) v.l[t] = P[t][q];
          s = mfn_sub<N>(s, v);
        }
      }
      __syncthreads();
    }
    if (threadIdx.x == 0) stp<N>(x, n, i, mfn_div_by<N>(s, ldp<N>(l, pl, i * n + i), ldp<N>(rd, n, i)));
    __syncthreads();
  }
}

// limbs 0 .. min(from, to) - 1 copied, the rest zero (multifloat/conv.rs:119-157's
// from_f128 / to_f128 / from_f256 / to_f256); AoS in and out
__global__ void k_mfn_convert(const double* __restrict__ in, double* __restrict__ out, int from, int to, long n) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  for (int l = 0; l < to; ++l) out[i * to + l] = l < from ? in[i * from + l] : 0.0;
}

unsigned blocks_of(long n, int t) { return (unsigned)((n + t - 1) / t); }

template <int N>
int mfn_gemm_dev(int trans_a, const double* a, const double* b, double* c, size_t m, size_t n, size_t k,
                 hipStream_t s) {
  dim3 grid((unsigned)((n + TB - 1) / TB), (unsigned)((m + TB - 1) / TB));
  if (trans_a)
    hipLaunchKernelGGL((k_mfn_gemm<N, true>), grid, dim3(256), 0, s, a, b, c, (int)m, (int)n, (int)k);
  else
    hipLaunchKernelGGL((k_mfn_gemm<N, false>), grid, dim3(256), 0, s, a, b, c, (int)m, (int)n, (int)k);
  HP_LAUNCH_CHECK();
  return HP_OK;
}
template <int N>
int mfn_gemv_dev(const double* a, const double* x, double* y, size_t n, size_t m, hipStream_t s) {
  hipLaunchKernelGGL(k_mfn_gemv<N>, dim3(blocks_of((long)n, 64)), dim3(64), 0, s, a, x, y, (long)n, (long)m);
  HP_LAUNCH_CHECK();
  return HP_OK;
}
template <int N>
int mfn_elem_dev(int op, const double* a, const double* b, double* out, size_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_mfn_elem<N>, dim3(blocks_of((long)n, 128)), dim3(128), 0, s, op, a, b, out, (long)n);
  HP_LAUNCH_CHECK();
  return HP_OK;
}
template <int N>
int mfn_chol_dev(double* a, double* S, MfnErr* err, size_t n, hipStream_t s) {
  for (long k = 0; k < (long)n; ++k) {
    hipLaunchKernelGGL(k_mfnc_col<N>, dim3(1), dim3(256), 0, s, a, static_cast<const double*>(S), (long)n, k, err);
    const long t = (long)n - k - 1;
    if (t > 0)
      hipLaunchKernelGGL(k_mfnc_update<N>, dim3(blocks_of(t, 16), blocks_of(t, 16)), dim3(256), 0, s,
                         static_cast<const double*>(a), S, (long)n, k, static_cast<const MfnErr*>(err));
  }
  HP_LAUNCH_CHECK();
  return HP_OK;
}
template <int N>
int mfn_solve_dev(const double* l, double* x, double* rd, size_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_mfn_diag_recip<N>, dim3(blocks_of((long)n, 64)), dim3(64), 0, s, l, rd, (long)n);
  hipLaunchKernelGGL(k_mfn_fwd<N>, dim3(1), dim3(256), 0, s, l, x, static_cast<const double*>(rd), (long)n);
  hipLaunchKernelGGL(k_mfn_bwd<N>, dim3(1), dim3(256), 0, s, l, x, static_cast<const double*>(rd), (long)n);
  HP_LAUNCH_CHECK();
  return HP_OK;
}

// N = 1..8 except 4 (the f256 kernels)
#define HP_MFN_SWITCH(NL, CALL)                  \
  switch (NL) {                                  \
    case 1: return CALL(1);                      \
    case 2: return CALL(2);                      \
    case 3: return CALL(3);                      \
    case 5: return CALL(5);                      \
    case 6: return CALL(6);                      \
    case 7: return CALL(7);                      \
    case 8: return CALL(8);                      \
    default: break;                              \
  }

int bad_limbs(int nl) {
  set_error("MultiFloat<N>: N = %d outside 1..8", nl);
  return HP_EINVAL;
}

}  // namespace

int launch_mfn_gemm(int nl, int trans_a, const double* a, const double* b, double* c, size_t m, size_t n, size_t k,
                    hipStream_t s) {
  if (nl == 4) return launch_mf4_gemm(trans_a, a, b, c, m, n, k, s);
  if (nl < 1 || nl > 8) return bad_limbs(nl);
  if (m == 0 || n == 0) return HP_OK;
  if ((m + TB - 1) / TB > 65535 || k > (1u << 30) || n > (1u << 30)) {
    set_error("mf_gemm: dimension too large");
    return HP_ESHAPE;
  }
#define C_(K) mfn_gemm_dev<K>(trans_a, a, b, c, m, n, k, s)
  HP_MFN_SWITCH(nl, C_)
#undef C_
  return bad_limbs(nl);
}

}  // namespace hp

using namespace hp;

namespace {
int null_arg() {
  set_error("null pointer argument");
  return HP_EINVAL;
}
int check_limbs(int nl) { return (nl < 1 || nl > 8) ? bad_limbs(nl) : HP_OK; }
}  // namespace

extern "C" {

int hp_dev_mfn_gemm(int nl, int trans_a, const double* a, const double* b, double* c, size_t m, size_t n, size_t k,
                    void* stream) {
  return launch_mfn_gemm(nl, trans_a, a, b, c, m, n, k, (hipStream_t)stream);
}

int hp_mfn_gemv(int nl, const double* a, const double* x, double* y, size_t n, size_t m) {
  if (int rc = check_limbs(nl)) return rc;
  if (nl == 4) return hp_mf_gemv(a, x, y, n, m);
  if (n == 0) return HP_OK;
  if (!a || !x || !y) return null_arg();
  hipStream_t s = hipStreamPerThread;
  DevBuf da, dx, dy;
  int rc;
  if ((rc = da.alloc(nl * n * m * sizeof(double))) || (rc = dx.alloc(nl * m * sizeof(double))) ||
      (rc = dy.alloc(nl * n * sizeof(double))))
    return rc;
  if ((m && (rc = upload_planes(a, n * m, nl, da.d(), s))) || (m && (rc = upload_planes(x, m, nl, dx.d(), s))))
    return rc;
#define C_(K) mfn_gemv_dev<K>(da.d(), dx.d(), dy.d(), n, m, s)
  rc = [&]() -> int {
    HP_MFN_SWITCH(nl, C_)
    return bad_limbs(nl);
  }();
#undef C_
  if (rc) return rc;
  return download_planes(dy.d(), n, nl, y, s);
}

int hp_mfn_dot(int nl, const double* a, const double* b, size_t n, double* out) {
  if (!out) return null_arg();
  return hp_mfn_gemv(nl, a, b, out, 1, n);  // mf_dot is the one-row gemv: s = s + a_i b_i from ZERO
}

int hp_mfn_elementwise(int nl, int op, const double* a, const double* b, double* out, size_t n) {
  if (int rc = check_limbs(nl)) return rc;
  if (op < 0 || op > 6) {
    set_error("unknown MultiFloat op %d", op);
    return HP_EINVAL;
  }
  if (n == 0) return HP_OK;
  const bool binary = op <= 3;
  if (!a || !out || (binary && !b)) return null_arg();
  hipStream_t s = hipStreamPerThread;
  DevBuf da, db, dout;
  int rc;
  if ((rc = da.alloc(nl * n * sizeof(double))) || (rc = db.alloc(nl * n * sizeof(double))) ||
      (rc = dout.alloc(nl * n * sizeof(double))))
    return rc;
  if ((rc = upload_planes(a, n, nl, da.d(), s)) || (binary && (rc = upload_planes(b, n, nl, db.d(), s)))) return rc;
  if (nl == 4) {
    rc = launch_mfn_elem4(op, da.d(), db.d(), dout.d(), n, s);
  } else {
#define C_(K) mfn_elem_dev<K>(op, da.d(), db.d(), dout.d(), n, s)
    rc = [&]() -> int {
      HP_MFN_SWITCH(nl, C_)
      return bad_limbs(nl);
    }();
#undef C_
  }
  if (rc) return rc;
  return download_planes(dout.d(), n, nl, out, s);
}

int hp_mfn_cholesky(int nl, double* a, size_t n, size_t* pivot) {
  if (int rc = check_limbs(nl)) return rc;
  if (nl == 4) return hp_mf_cholesky(a, n, pivot);
  if (n == 0) return HP_OK;
  if (!a) return null_arg();
  hipStream_t s = hipStreamPerThread;
  const size_t nn = n * n;
  DevBuf da, dS, derr;
  int rc;
  if ((rc = da.alloc(nl * nn * sizeof(double))) || (rc = dS.alloc(nl * nn * sizeof(double))) ||
      (rc = derr.alloc(sizeof(MfnErr))))
    return rc;
  if ((rc = upload_planes(a, nn, nl, da.d(), s))) return rc;
  HP_HIP(hipMemsetAsync(dS.p, 0, nl * nn * sizeof(double), s));
  HP_HIP(hipMemsetAsync(derr.p, 0, sizeof(MfnErr), s));
  auto* err = static_cast<MfnErr*>(derr.p);
#define C_(K) mfn_chol_dev<K>(da.d(), dS.d(), err, n, s)
  rc = [&]() -> int {
    HP_MFN_SWITCH(nl, C_)
    return bad_limbs(nl);
  }();
#undef C_
  if (rc) return rc;
  MfnErr he;
  HP_HIP(hipMemcpyAsync(&he, err, sizeof(he), hipMemcpyDeviceToHost, s));
  HP_HIP(hipStreamSynchronize(s));
  if ((rc = download_planes(da.d(), nn, nl, a, s))) return rc;  // on failure: the columns before the pivot
  if (he.flag) {
    if (pivot) *pivot = (size_t)he.pivot;
    set_error("Cholesky failed: matrix is not positive definite (pivot %ld)", he.pivot);
    return HP_ENOTSPD;
  }
  return HP_OK;
}

int hp_mfn_cholesky_solve(int nl, const double* l, const double* b, double* x, size_t n) {
  if (int rc = check_limbs(nl)) return rc;
  if (nl == 4) return hp_mf_cholesky_solve(l, b, x, n);
  if (n == 0) return HP_OK;
  if (!l || !b || !x) return null_arg();
  hipStream_t s = hipStreamPerThread;
  const size_t nn = n * n;
  DevBuf dl, dx, dr;
  int rc;
  if ((rc = dl.alloc(nl * nn * sizeof(double))) || (rc = dx.alloc(nl * n * sizeof(double))) ||
      (rc = dr.alloc(nl * n * sizeof(double))))
    return rc;
  if ((rc = upload_planes(l, nn, nl, dl.d(), s)) || (rc = upload_planes(b, n, nl, dx.d(), s))) return rc;
#define C_(K) mfn_solve_dev<K>(dl.d(), dx.d(), dr.d(), n, s)
  rc = [&]() -> int {
    HP_MFN_SWITCH(nl, C_)
    return bad_limbs(nl);
  }();
#undef C_
  if (rc) return rc;
  return download_planes(dx.d(), n, nl, x, s);
}

int hp_mfn_convert(int nl_from, int nl_to, const double* in, double* out, size_t n) {
  if (int rc = check_limbs(nl_from)) return rc;
  if (int rc = check_limbs(nl_to)) return rc;
  if (n == 0) return HP_OK;
  if (!in || !out) return null_arg();
  hipStream_t s = hipStreamPerThread;
  DevBuf di, dout;
  int rc;
  if ((rc = di.alloc(nl_from * n * sizeof(double))) || (rc = dout.alloc(nl_to * n * sizeof(double)))) return rc;
  HP_HIP(hipMemcpyAsync(di.p, in, nl_from * n * sizeof(double), hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_mfn_convert, dim3(blocks_of((long)n, 256)), dim3(256), 0, s, static_cast<const double*>(di.d()),
                     dout.d(), nl_from, nl_to, (long)n);
  HP_LAUNCH_CHECK();
  HP_HIP(hipMemcpyAsync(out, dout.p, nl_to * n * sizeof(double), hipMemcpyDeviceToHost, s));
  HP_HIP(hipStreamSynchronize(s));
  return HP_OK;
}

}  // extern "C"
