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
#include <type_traits>
#include <vector>

#include "hp_chol.hpp"
#include "hp_internal.hpp"

namespace hp {

template <class A>
int chol_exact(Mat a, long n, long nb, bool unblocked, Mat S, ErrRec* err, double* inv_h, double* inv_l, hipStream_t s);
template <class A>
int chol_panel_split(Mat a, Mat S, long c0, long jb, long n, ErrRec* err, double* inv, hipStream_t s,
                     hipStream_t side, hipEvent_t* evq, hipEvent_t evdone);
template <class A>
int fwd_exact(Mat l, double* yh, double* yl, long n, hipStream_t s);
template <class A>
int rank_update_sumsub(Mat a, Mat S, long ri0, long M, long cj0, long N, long ka0, long K, ErrRec* err, hipStream_t s);
template <class A>
int rank_update_sumsub_below(Mat a, Mat S, long ri0, long M, long cj0, long N, long ka0, long K, ErrRec* err,
                             hipStream_t s);

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

template <class A>
__device__ __forceinline__ typename A::T Vg(double (*Vh)[BW + 1], double (*Vl)[BW + 1], int i, int c) {
  if constexpr (A::LIMBS == 2) return mkdd(Vh[i][c], Vl[i][c]);
  else return Vh[i][c];
}

// ---------------------------------------------------------------- blocked inverse solves
// The correction solves use precomputed inverses of the 64 x 64 diagonal blocks of L, so
// no block step has a sequential substitution chain (k_inv_solve below).
template <class A>
__global__ __launch_bounds__(64) void k_diag_block_inverse(Mat l, long n, double* ih, double* il) {
  // inv(L_bb) for block b = blockIdx.x, lane c computes column c by forward substitution
  using T = typename A::T;
  __shared__ double Lh[BW][BW + 1], Ll[A::LIMBS == 2 ? BW : 1][BW + 1];
  __shared__ double Vh[BW][BW + 1], Vl[A::LIMBS == 2 ? BW : 1][BW + 1];
  const long b0 = (long)blockIdx.x * BW;
  const int bw = (int)min((long)BW, n - b0);
  const int c = threadIdx.x;
  for (int i = 0; i < bw; ++i)
    if (c <= i) {
      T v = A::ld(l, b0 + i, b0 + c);
      if constexpr (A::LIMBS == 2) { Lh[i][c] = v.hi; Ll[i][c] = v.lo; }
      else Lh[i][c] = v;
    }
  __syncthreads();
  auto Lg = [&](int i, int k) -> T {
    if constexpr (A::LIMBS == 2) return mkdd(Lh[i][k], Ll[i][k]);
    else return Lh[i][k];
  };
  if (c < bw) {
    for (int i = 0; i < bw; ++i) {
      T v = A::zero();
      if (i >= c) {
        T s = (i == c) ? A::one() : A::zero();
        for (int k = c; k < i; ++k) {
          T vk;
          if constexpr (A::LIMBS == 2) vk = mkdd(Vh[k][c], Vl[k][c]);
          else vk = Vh[k][c];
          s = A::sub(s, A::mul(Lg(i, k), vk));
        }
        v = A::div(s, Lg(i, i));
      }
      if constexpr (A::LIMBS == 2) { Vh[i][c] = v.hi; Vl[i][c] = v.lo; }
      else Vh[i][c] = v;
    }
  }
  __syncthreads();
  // store as [block][i][c]
  const size_t base = (size_t)blockIdx.x * BW * BW;
  for (int i = 0; i < BW; ++i) {
    T v = (i < bw && c < bw) ? Vg<A>(Vh, Vl, i, c) : A::zero();
    A::stv(ih, il, (long)(base + (size_t)i * BW + c), v);
  }
}

// ---------------------------------------------------------------- persistent block solves
// L L^T x = y in ONE launch of 2 nblk workgroups (512 threads), with the inverses of the
// 64 x 64 diagonal blocks:
//   workgroup b < nblk:        z_b = inv(L_bb)   (y_b - sum_{j<b} L_bj z_j)
//   workgroup nblk + c:        x_b = inv(L_bb)^T (z_b - sum_{j>b} L_jb^T x_j),  b = nblk-1-c
// Results travel between workgroups as self-validating values: z and x start filled with
// EMPTY (a signalling-NaN pattern that no arithmetic result carries: IEEE operations
// return quiet NaNs), a block's values are published with relaxed agent-scope atomic
// stores, and a consumer polls them with relaxed agent-scope atomic loads until none of a
// block's values is EMPTY.  Every value is its own atomic, so no flags, fences or second
// round trip are needed between a block's stores and its use.  Each of the 8 wavefronts
// polls one of the next 8 dependency blocks per round (64 lanes, one element each), so a
// workgroup that is behind takes up to 8 ready blocks per round.  Every wait is on a
// lower-numbered workgroup: under in-order dispatch the smallest-numbered resident
// workgroup never waits on one that is not running, so the chain always advances,
// whatever the residency.  Each wait is also bounded: past SPIN_LIMIT polls it raises
// *fail and the workgroups leave (the host reports a device error).  The consumer of x
// (k_add_absmax) refills z and x with EMPTY for the next solve.
// Summation: eight lanes per row, four interleaved chains per lane, each a compensated
// multiply-accumulate (dd products with the low cross terms, two_sum of the leading
// parts, the errors summed in a second double), renormalised at the end; these solves only
// feed the refinement, so neither the reference's order nor its operations are needed.
constexpr int SPIN_LIMIT = 1 << 22;
constexpr unsigned long long EMPTY = 0x7FF0DEADBEEF0001ull;

__device__ __forceinline__ double ld_coherent(const double* p) {
  return __builtin_bit_cast(double, __hip_atomic_load((const unsigned long long*)p, __ATOMIC_RELAXED,
                                                      __HIP_MEMORY_SCOPE_AGENT));
}
__device__ __forceinline__ void st_coherent(double* p, double v) {
  __hip_atomic_store((unsigned long long*)p, __builtin_bit_cast(unsigned long long, v), __ATOMIC_RELAXED,
                     __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ bool is_empty(double v) { return __builtin_bit_cast(unsigned long long, v) == EMPTY; }

// compensated multiply-accumulate: a (h + l, unnormalised) += x y
struct AccDD {
  double h, l;
};
__device__ __forceinline__ void cacc_mac(AccDD& a, dd x, dd y) {
  const double p = x.hi * y.hi;
  double e = __builtin_fma(x.hi, y.hi, -p);
  e = __builtin_fma(x.hi, y.lo, e);
  e = __builtin_fma(x.lo, y.hi, e);
  double s, t;
  two_sum(a.h, p, s, t);
  a.h = s;
  a.l = a.l + (t + e);
}
__device__ __forceinline__ void cacc_add(AccDD& a, AccDD b) {
  double s, t;
  two_sum(a.h, b.h, s, t);
  a.h = s;
  a.l = a.l + (b.l + t);
}
__device__ __forceinline__ void cacc_mac(double& a, double x, double y) { a = __builtin_fma(x, y, a); }
__device__ __forceinline__ void cacc_add(double& a, double b) { a = a + b; }
__device__ __forceinline__ dd cacc_value(AccDD a) { return renorm(a.h, a.l); }
__device__ __forceinline__ double cacc_value(double a) { return a; }
__device__ __forceinline__ AccDD cacc_shfl_xor(AccDD a, int off) {
  return AccDD{__shfl_xor(a.h, off, 64), __shfl_xor(a.l, off, 64)};
}
__device__ __forceinline__ double cacc_shfl_xor(double a, int off) { return __shfl_xor(a, off, 64); }

template <class A>
using AccOf = typename std::conditional<A::LIMBS == 2, AccDD, double>::type;

// the four chains of a lane, then the eight lanes 8r..8r+7 that share a row
template <class A>
__device__ __forceinline__ typename A::T row_total(AccOf<A> (&p)[4]) {
  cacc_add(p[0], p[1]);
  cacc_add(p[2], p[3]);
  cacc_add(p[0], p[2]);
#pragma unroll
  for (int off = 1; off < 8; off <<= 1) cacc_add(p[0], cacc_shfl_xor(p[0], off));
  return cacc_value(p[0]);
}

template <class A>
__global__ __launch_bounds__(512) void k_inv_solve(Mat l, const double* __restrict__ ih, const double* __restrict__ il,
                                                   const double* yh, const double* yl, double* zh, double* zl,
                                                   double* xh, double* xl, long n, int nblk, int* fail,
                                                   unsigned long long* red) {
  using T = typename A::T;
  using Acc = AccOf<A>;
  constexpr int TPL = 8;  // terms per lane
  constexpr int NW = 8;   // wavefronts, one polled block each
  __shared__ double Vh[NW][BW], Vl[A::LIMBS == 2 ? NW : 1][BW];
  __shared__ double Th[BW], Tl[BW], Oh[BW], Ol[BW];
  __shared__ int s_valid[NW], s_abort;
  const int t = threadIdx.x, part = t & 7, kb = part * TPL, r = t >> 3, w = t >> 6, lane = t & 63;
  const bool fwd = (int)blockIdx.x < nblk;
  const int b = fwd ? (int)blockIdx.x : 2 * nblk - 1 - (int)blockIdx.x;
  const long b0 = (long)b * BW;
  const int bw = (int)min((long)BW, n - b0);
  const bool live = r < bw;
  if (blockIdx.x == 0 && t < 2) red[t] = 0;  // the max-reductions of the kernel that follows
  // this lane's terms of the inverse (fwd: row r, bwd: column r), loaded ahead of any wait
  const size_t base = (size_t)b * BW * BW;
  T iv[TPL];
#pragma unroll
  for (int u = 0; u < TPL; ++u) {
    const int k = kb + u;
    iv[u] = A::ldv(ih, il, (long)(base + (fwd ? (size_t)r * BW + k : (size_t)k * BW + r)));
  }
  // Polls blocks first, first + step, ... (at most min(NW, count)) of (ph, pl) into the
  // slots Vh[q] until the first is complete; returns how many leading blocks are, or 0
  // when the wait was abandoned.  Called by the whole workgroup.
  auto poll = [&](const double* ph, const double* pl, int first, int step, int count) -> int {
    const int m = min(NW, count);
    for (int spins = 0;; ++spins) {
      if (w < m) {
        const long e = (long)(first + w * step) * BW + lane;
        bool v = true;
        double h = 0.0, lo = 0.0;
        if (e < n) {
          h = ld_coherent(ph + e);
          v = !is_empty(h);
          if constexpr (A::LIMBS == 2) {
            lo = ld_coherent(pl + e);
            v = v && !is_empty(lo);
          }
        }
        Vh[w][lane] = h;
        if constexpr (A::LIMBS == 2) Vl[w][lane] = lo;
        const bool all = __all(v);
        if (lane == 0) s_valid[w] = all;
      }
      if (t == 0) {
        s_abort = 0;
        if ((spins & 255) == 255 && __hip_atomic_load(fail, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) s_abort = 1;
        if (spins > SPIN_LIMIT) {
          __hip_atomic_store(fail, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          s_abort = 1;
        }
      }
      __syncthreads();
      int ready = 0;
      while (ready < m && s_valid[ready]) ++ready;
      if (ready > 0) return ready;
      if (s_abort) return 0;
      __syncthreads();  // every wavefront has read s_valid before the next poll rewrites it
      __builtin_amdgcn_s_sleep(1);
    }
  };
  Acc p[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) p[q] = Acc{};
  // this lane's factor entries of dependency block j: fwd L[b0+r][j0+kb+u] (row r of L_bj),
  // bwd L[j0+kb+u][b0+r] (column r of L_jb); zero outside the matrix
  auto load_l = [&](int j, T (&lv)[TPL]) {
    const long j0 = (long)j * BW;
    const int bwj = (int)min((long)BW, n - j0);
#pragma unroll
    for (int u = 0; u < TPL; ++u) {
      const int k = kb + u;
      lv[u] = (live && k < bwj) ? (fwd ? A::ld(l, b0 + r, j0 + k) : A::ld(l, j0 + k, b0 + r)) : A::zero();
    }
  };
  const double* vh = fwd ? zh : xh;  // the dependencies' results
  const double* vl = fwd ? zl : xl;
  int next = fwd ? 0 : nblk - 1;  // next dependency to take: fwd ascending below b, bwd descending above b
  const int step = fwd ? 1 : -1;
  int remaining = fwd ? b : nblk - 1 - b;
  T lv[TPL];
  while (remaining > 0) {
    load_l(next, lv);  // the factor does not depend on the results: in flight during the poll
    const int ready = poll(vh, vl, next, step, remaining);
    if (ready == 0) return;
    for (int q = 0; q < ready; ++q) {
      if (q > 0) load_l(next + q * step, lv);
#pragma unroll
      for (int u = 0; u < TPL; ++u) {
        T vv;
        if constexpr (A::LIMBS == 2) vv = mkdd(Vh[q][kb + u], Vl[q][kb + u]);
        else vv = Vh[q][kb + u];
        cacc_mac(p[u & 3], lv[u], vv);  // lanes past the matrix hold zero factor entries
      }
    }
    next += ready * step;
    remaining -= ready;
    __syncthreads();  // the slots are rewritten by the next poll
  }
  T own = A::zero();  // bwd: z_b of the forward pass
  if (!fwd) {
    if (poll(zh, zl, b, 1, 1) == 0) return;
    if constexpr (A::LIMBS == 2) own = mkdd(Vh[0][r], Vl[0][r]);
    else own = Vh[0][r];
  }
  const T sum = row_total<A>(p);
  if (part == 0) {
    T v = A::zero();
    if (live) v = A::sub(fwd ? A::ldv(yh, yl, b0 + r) : own, sum);
    if constexpr (A::LIMBS == 2) { Th[r] = v.hi; Tl[r] = v.lo; }
    else Th[r] = v;
  }
  __syncthreads();
  Acc q[4];
#pragma unroll
  for (int u = 0; u < 4; ++u) q[u] = Acc{};
#pragma unroll
  for (int u = 0; u < TPL; ++u) {
    T tv;
    if constexpr (A::LIMBS == 2) tv = mkdd(Th[kb + u], Tl[kb + u]);
    else tv = Th[kb + u];
    cacc_mac(q[u & 3], iv[u], tv);
  }
  const T o = row_total<A>(q);
  if (part == 0) {
    if constexpr (A::LIMBS == 2) { Oh[r] = o.hi; Ol[r] = o.lo; }
    else Oh[r] = o;
  }
  __syncthreads();
  if (t < bw) {  // publish the block: one coherent store per value
    double* oh = fwd ? zh : xh;
    double* ol = fwd ? zl : xl;
    st_coherent(oh + b0 + t, Oh[t]);
    if constexpr (A::LIMBS == 2) st_coherent(ol + b0 + t, Ol[t]);
  }
}

// x += dx (dd += dd, or dd += f64 when dl is null), and red[0] = max|dx.hi|, red[1] =
// max|x.hi| after the update (bit patterns of non-negative doubles order as integers);
// red is cleared by the solve kernel that precedes this one on the stream.  dx and the
// solve's intermediate z (zh, zl; zl may be null) are refilled with EMPTY for the next solve.
__global__ __launch_bounds__(256) void k_add_absmax(double* xh, double* xl, double* dh, double* dl, long n,
                                                    unsigned long long* red, double* zh, double* zl) {
  __shared__ double m0[256], m1[256];
  const double empty = __builtin_bit_cast(double, EMPTY);
  double a0 = 0.0, a1 = 0.0;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const double d0 = dh[i];
    dd r = dl ? dd_add(mkdd(xh[i], xl[i]), mkdd(d0, dl[i])) : dd_add_f(mkdd(xh[i], xl[i]), d0);
    xh[i] = r.hi;
    xl[i] = r.lo;
    a0 = fmax(a0, fabs(d0));
    a1 = fmax(a1, fabs(r.hi));
    dh[i] = empty;
    if (dl) dl[i] = empty;
    zh[i] = empty;
    if (zl) zl[i] = empty;
  }
  m0[threadIdx.x] = a0;
  m1[threadIdx.x] = a1;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      m0[threadIdx.x] = fmax(m0[threadIdx.x], m0[threadIdx.x + s]);
      m1[threadIdx.x] = fmax(m1[threadIdx.x], m1[threadIdx.x + s]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    atomicMax(red, __builtin_bit_cast(unsigned long long, m0[0]));
    atomicMax(red + 1, __builtin_bit_cast(unsigned long long, m1[0]));
  }
}
__global__ void k_fill_empty(double* p, long n) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = __builtin_bit_cast(double, EMPTY);
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

// inv(L_bb) for all 64-row diagonal blocks: nblk x 64 x 64 values per plane
template <class A>
int diag_inverses(Mat l, long n, double* ih, double* il, hipStream_t s) {
  long nblk = (n + BW - 1) / BW;
  hipLaunchKernelGGL(k_diag_block_inverse<A>, dim3((unsigned)nblk), dim3(64), 0, s, l, n, ih, il);
  HP_LAUNCH_CHECK();
  return HP_OK;
}
// L L^T x = y through the persistent kernel; z (n) is scratch, z and x hold EMPTY on entry
// (k_fill_empty, or k_add_absmax after the previous solve), fail belongs to the caller
template <class A>
int solve_inv_persistent(Mat l, const double* ih, const double* il, const double* yh, const double* yl, double* zh,
                         double* zl, double* xh, double* xl, long n, int* fail, unsigned long long* red,
                         hipStream_t s) {
  const int nblk = (int)((n + BW - 1) / BW);
  hipLaunchKernelGGL(k_inv_solve<A>, dim3((unsigned)(2 * nblk)), dim3(512), 0, s, l, ih, il, yh, yl, zh, zl, xh, xl, n,
                     nblk, fail, red);
  HP_LAUNCH_CHECK();
  return HP_OK;
}

}  // namespace

// dd blocked Cholesky with the trailing update P P^T on the INT8 matrix cores.  Per panel:
// the panel on FP64 VALU (diagonal sub-blocks with dd reciprocal-square-root pivots, the
// rows below solved on a side stream as each sub-block is ready), the next panel's
// columns updated on VALU (its diagonal block first, the rows below on the side stream),
// then the rest of the trailing matrix updated by one split SYRK
// (lower tiles only, the subtraction fused into the CRT) on a second stream while the next
// panel is factored.  Every element receives the panels' updates in panel order.  All
// updates subtract in place: S (a partial-sum plane in the other factorisations) is not
// read or written.
// Updates of the split-dd factorisation whose summation order is free:
//   a(ri0 + i, cj0 + j) -= sum_k a(ri0 + i, c0 + k) a(cj0 + j, c0 + k),  i < M, j < N, k < K
// (TRI: ri0 == cj0, M == N, j <= i only): the next panel's diagonal block at each panel
// boundary, and within a panel the later columns of each sub-block's rows, both within the
// panel (TRI) and below it (the side stream).  Each 16 x 16 tile takes a workgroup of 1024
// threads.  The tile's 16 + 16 rows come through LDS in 64-deep k chunks (coalesced
// 512-byte row pieces, the next chunk loaded into registers while the current one is
// summed); thread group g (threadIdx / 256) sums the 16 k of its quarter of each chunk with
// the compensated multiply-accumulate, the quarters are added in a fixed order, and the dd
// total is subtracted.  The reference-order update (one 256-long dd chain per element on
// 136 workgroups of 256) took ~40 us at M = N = K = 256, as did one reading its rows
// straight from global memory 8 bytes at a time.
constexpr int SYD_KC = 64;
template <bool TRI>
__global__ __launch_bounds__(1024) void k_syrk_split(Mat a, long ri0, long M, long cj0, long N, long c0, long K,
                                                     const ErrRec* err) {
  __shared__ double Ph[2][16][SYD_KC + 1], Pl[2][16][SYD_KC + 1];  // [0]: rows i, [1]: rows j
  __shared__ double ph[4][256], pl[4][256];
  if (err->flag) return;
  const long b = blockIdx.x;
  long ti, tc;
  if constexpr (TRI) {  // triangular tile index: b = ti (ti + 1) / 2 + tc, tc <= ti
    ti = (long)((__builtin_sqrt(8.0 * (double)b + 1.0) - 1.0) * 0.5);
    while ((ti + 1) * (ti + 2) / 2 <= b) ++ti;
    while (ti * (ti + 1) / 2 > b) --ti;
    tc = b - ti * (ti + 1) / 2;
  } else {
    const long ntc = (N + 15) / 16;
    ti = b / ntc;
    tc = b % ntc;
  }
  const int t = threadIdx.x & 255, g = threadIdx.x >> 8;
  const int ty = t >> 4, tx = t & 15;
  const long i = ti * 16 + ty, j = tc * 16 + tx;
  const bool live = i < M && j < N && (!TRI || j <= i);
  // loader: thread T stages row lr (0..31: 16 rows i, then 16 rows j) at k offset lk, hi and lo
  const int T = threadIdx.x, lr = T >> 5, lk = 2 * (T & 31);
  const long lrow = lr < 16 ? ti * 16 + lr : tc * 16 + (lr - 16);
  const bool rok = lrow < (lr < 16 ? M : N);
  const long grow = (lr < 16 ? ri0 : cj0) + lrow;
  double v[4];
  auto fetch = [&](long k0) {
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      const long k = k0 + lk + q;
      const bool ok = rok && k < K;
      const long o = grow * a.ld + c0 + k;
      v[2 * q] = ok ? a.hi[o] : 0.0;
      v[2 * q + 1] = ok ? a.lo[o] : 0.0;
    }
  };
  AccDD acc{0.0, 0.0};
  fetch(0);
  for (long k0 = 0; k0 < K; k0 += SYD_KC) {
    __syncthreads();  // the previous chunk has been read
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      Ph[lr >> 4][lr & 15][lk + q] = v[2 * q];
      Pl[lr >> 4][lr & 15][lk + q] = v[2 * q + 1];
    }
    __syncthreads();
    if (k0 + SYD_KC < K) fetch(k0 + SYD_KC);
    if (live) {
#pragma unroll
      for (int kk = 16 * g; kk < 16 * g + 16; ++kk)
        cacc_mac(acc, mkdd(Ph[0][ty][kk], Pl[0][ty][kk]), mkdd(Ph[1][tx][kk], Pl[1][tx][kk]));
    }
  }
  ph[g][t] = acc.h;
  pl[g][t] = acc.l;
  __syncthreads();
  if (g == 0 && live) {
    AccDD sum{ph[0][t], pl[0][t]};
#pragma unroll
    for (int q = 1; q < 4; ++q) cacc_add(sum, AccDD{ph[q][t], pl[q][t]});
    const long o = (ri0 + i) * a.ld + cj0 + j;
    const dd r = dd_sub(mkdd(a.hi[o], a.lo[o]), cacc_value(sum));
    a.hi[o] = r.hi;
    a.lo[o] = r.lo;
  }
}

// the TRI form on the block at (t0, t0) of width w, on s
int diag_syrk_split(Mat a, long t0, long w, long c0, long K, ErrRec* err, hipStream_t s) {
  if (w <= 0 || K <= 0) return HP_OK;
  const long nt = (w + 15) / 16;
  hipLaunchKernelGGL(k_syrk_split<true>, dim3((unsigned)(nt * (nt + 1) / 2)), dim3(1024), 0, s, a, t0, w, t0, w, c0,
                     K, (const ErrRec*)err);
  HP_LAUNCH_CHECK();
  return HP_OK;
}
// the rectangular form (rows ri0.., columns cj0..), on s
int rect_syrk_split(Mat a, long ri0, long M, long cj0, long N, long c0, long K, ErrRec* err, hipStream_t s) {
  if (M <= 0 || N <= 0 || K <= 0) return HP_OK;
  const long nb = ((M + 15) / 16) * ((N + 15) / 16);
  if (nb > 0x7fffffffL) {
    set_error("split factorisation: update of %ld x %ld too large", M, N);
    return HP_EINVAL;
  }
  hipLaunchKernelGGL(k_syrk_split<false>, dim3((unsigned)nb), dim3(1024), 0, s, a, ri0, M, cj0, N, c0, K,
                     (const ErrRec*)err);
  HP_LAUNCH_CHECK();
  return HP_OK;
}

int chol_split(Mat a, long n, long nb, Mat S, ErrRec* err, double*, double*, hipStream_t s) {
  if (nb <= 0) nb = 256;
  nb = std::max<long>(64, (nb / 64) * 64);
  const int Q = (int)(nb / 64);
  if (Q > 8) {
    set_error("split factorisation: panel width %ld above 512", nb);
    return HP_EINVAL;
  }
  DevBuf invb;
  int rc;
  if ((rc = invb.alloc((size_t)Q * 2 * 64 * sizeof(double)))) return rc;
  hipStream_t s2 = nullptr, s3 = nullptr;
  hipEvent_t evA = nullptr, evB = nullptr, evP = nullptr, evq[16] = {};
  bool evB_live = false;
  struct Guard {
    hipStream_t *s2, *s3;
    hipEvent_t *a, *b, *p, *q;
    ~Guard() {
      for (hipStream_t* x : {s2, s3})
        if (*x) { (void)hipStreamSynchronize(*x); (void)hipStreamDestroy(*x); }
      for (hipEvent_t* e : {a, b, p})
        if (*e) (void)hipEventDestroy(*e);
      for (int i = 0; i < 16; ++i)
        if (q[i]) (void)hipEventDestroy(q[i]);
    }
  } guard{&s2, &s3, &evA, &evB, &evP, evq};
  HP_HIP(hipStreamCreateWithFlags(&s2, hipStreamNonBlocking));
  HP_HIP(hipStreamCreateWithFlags(&s3, hipStreamNonBlocking));
  for (hipEvent_t* e : {&evA, &evB, &evP}) HP_HIP(hipEventCreateWithFlags(e, hipEventDisableTiming));
  for (int q = 0; q < Q; ++q) {
    HP_HIP(hipEventCreateWithFlags(&evq[q], hipEventDisableTiming));
    HP_HIP(hipEventCreateWithFlags(&evq[8 + q], hipEventDisableTiming));
  }
  if ((rc = chol_panel_split<ArDD>(a, S, 0, std::min(nb, n), n, err, invb.d(), s, s3, evq, evP))) return rc;
  for (long c0 = 0; c0 < n; c0 += nb) {
    long jb = std::min(nb, n - c0);
    long rem = n - c0 - jb;
    if (rem <= 0) break;
    long t0 = c0 + jb, w2 = std::min(nb, rem);
    if (evB_live) {
      HP_HIP(hipStreamWaitEvent(s, evB, 0));
      HP_HIP(hipStreamWaitEvent(s3, evB, 0));
    }
    // next panel's columns take this panel's update: its diagonal block on s (FP64 VALU),
    // the rows below on s3 (where the next panel's rows below are solved) as a split GEMM
    // in the second workspace slot, beside the trailing SYRK's
    if ((rc = diag_syrk_split(a, t0, w2, c0, jb, err, s))) return rc;
    if (rem > w2) {
      const long r1 = t0 + w2;
      SplitOperand P{a.hi + r1 * a.ld + c0, a.lo + r1 * a.ld + c0, a.ld, true};
      SplitOperand Q{a.hi + t0 * a.ld + c0, a.lo + t0 * a.ld + c0, a.ld, true};
      SplitOpts o;
      o.subtract = true;
      o.slot = 1;
      o.variant = 4064;
      if ((rc = split_gemm_general(P, Q, a.hi + r1 * a.ld + t0, a.lo + r1 * a.ld + t0, (size_t)a.ld,
                                   (size_t)(rem - w2), (size_t)w2, (size_t)jb, 0, s3, o)))
        return rc;
    }
    HP_HIP(hipEventRecord(evA, s));
    // next panel on s (and s3), concurrently with the rest update below
    if ((rc = chol_panel_split<ArDD>(a, S, t0, w2, n, err, invb.d(), s, s3, evq, evP))) return rc;
    if (rem > w2) {
      HP_HIP(hipStreamWaitEvent(s2, evA, 0));
      long r0 = t0 + w2, m = n - r0;
      SplitOperand P{a.hi + r0 * a.ld + c0, a.lo + r0 * a.ld + c0, a.ld, true};
      SplitOpts opt;
      opt.lower_only = opt.subtract = opt.same_operand = true;
      if ((rc = split_gemm_general(P, P, a.hi + r0 * a.ld + r0, a.lo + r0 * a.ld + r0, (size_t)a.ld, (size_t)m,
                                   (size_t)m, (size_t)jb, 0, s2, opt)))
        return rc;
      HP_HIP(hipEventRecord(evB, s2));
      evB_live = true;
    }
  }
  if (evB_live) HP_HIP(hipStreamWaitEvent(s, evB, 0));
  HP_HIP(hipStreamSynchronize(s));  // invb is freed on return
  return HP_OK;
}

std::atomic<int> g_split_panel{256};  // panel width of the split-dd factorisation (process-wide)

// x = S^-1 b by corrections through a given factor L (f64 or dd by kind) and the inverses of
// its 64 x 64 diagonal blocks: r = b - S x in quad-double, x += L^-T L^-1 r, until the
// correction's max-norm falls to tol relative to x (converged = 1), or the next one,
// predicted from the last two by the geometric contraction, is ten times below tol while
// the corrections shrink at least 4x per step (converged = 2).  Stagnation (after four
// corrections, one above half the smallest so far, as lib.rs:102-106) or max_iter steps
// end the loop unconverged (converged = 0).
int refine_loop(int kind, Mat Lm, const double* ivh, const double* ivl, const double* sh, const double* sl,
                const double* bh, const double* bl, long n, double tol, int max_iter, double* xh, double* xl,
                int* iters, double* last_out, int* converged, hipStream_t s) {
  int rc;
  DevBuf vec, sc;
  if ((rc = vec.alloc(6 * n * sizeof(double))) || (rc = sc.alloc(64))) return rc;
  double* rh = vec.d();
  double* rlo = rh + n;
  double* dh = rh + 2 * n;  // the correction (dd, or f64 in dh for the f64 factor)
  double* dl = rh + 3 * n;
  double* zh = rh + 4 * n;  // the forward solve's result
  double* zl = rh + 5 * n;
  unsigned long long* red = (unsigned long long*)sc.p;  // max|dx|, max|x|
  int* fail = (int*)((char*)sc.p + 16);
  HP_HIP(hipMemsetAsync(sc.p, 0, 64, s));
  HP_HIP(hipMemsetAsync(xh, 0, n * sizeof(double), s));
  HP_HIP(hipMemsetAsync(xl, 0, n * sizeof(double), s));
  hipLaunchKernelGGL(k_fill_empty, dim3(g1(4 * n)), dim3(256), 0, s, dh, 4 * n);  // dx and z (contiguous)
  HP_LAUNCH_CHECK();
  double last = 1.0, best = HUGE_VAL;
  int it = 0;
  *converged = HP_NOT_CONVERGED;
  for (; it < max_iter; ++it) {
    const double* yh_ = bh;  // the right-hand side of this correction: b, then r = b - S x
    const double* yl_ = bl;
    if (it > 0) {
      hipLaunchKernelGGL(k_resid_qd, dim3((unsigned)((n + 3) / 4)), dim3(256), 0, s, sh, sl, xh, xl, bh, bl, rh, rlo, n);
      HP_LAUNCH_CHECK();
      yh_ = rh;
      yl_ = rlo;
    }
    if (kind == HP_FACTOR_F64) {
      hipLaunchKernelGGL(k_dd_to_f64, dim3(g1(n)), dim3(256), 0, s, yh_, yl_, dh, n);  // f64 right-hand side
      HP_LAUNCH_CHECK();
      if ((rc = solve_inv_persistent<ArF64>(Lm, ivh, nullptr, dh, nullptr, zh, nullptr, dl, nullptr, n, fail, red, s)))
        return rc;
      hipLaunchKernelGGL(k_add_absmax, dim3(std::min<unsigned>(g1(n), 1024)), dim3(256), 0, s, xh, xl, dl, nullptr, n,
                         red, zh, nullptr);
    } else {
      if ((rc = solve_inv_persistent<ArDD>(Lm, ivh, ivl, yh_, yl_, zh, zl, dh, dl, n, fail, red, s))) return rc;
      hipLaunchKernelGGL(k_add_absmax, dim3(std::min<unsigned>(g1(n), 1024)), dim3(256), 0, s, xh, xl, dh, dl, n, red,
                         zh, zl);
    }
    HP_LAUNCH_CHECK();
    struct { unsigned long long m[2]; int fail; } h{};
    HP_HIP(hipMemcpyAsync(&h, sc.p, sizeof(h), hipMemcpyDeviceToHost, s));
    HP_HIP(hipStreamSynchronize(s));
    if (h.fail) {
      set_error("refinement: a block solve's wait exceeded its bound");
      return HP_EDEVICE;
    }
    const double dn = __builtin_bit_cast(double, h.m[0]), xn = __builtin_bit_cast(double, h.m[1]);
    const double prev = last;
    last = xn > 0 ? dn / xn : dn;
    if (last <= tol) {
      *converged = HP_CONVERGED_MEASURED;
      ++it;
      break;
    }
    // the corrections shrink geometrically by the factor's contraction rho ~ kappa u_f:
    // stop once the predicted next one, last * (last / prev), is ten times below tol
    if (it >= 1 && prev > 0 && last <= 0.25 * prev && last * (last / prev) <= 0.1 * tol) {
      *converged = HP_CONVERGED_PREDICTED;
      ++it;
      break;
    }
    if (it >= 4 && last > 0.5 * best) { ++it; break; }  // stagnation
    best = std::min(best, last);
  }
  *iters = it;
  *last_out = last;
  return HP_OK;
}

// dd solve with a given dd factor (reference-order or split), refined as above
int solve_with_dd_factor(Mat Lm, const double* sh, const double* sl, const double* bh, const double* bl, long n,
                         double tol, int max_iter, double* xh, double* xl, int* iters, double* last, hipStream_t s) {
  const long nblk = (n + BW - 1) / BW;
  DevBuf binv;
  int rc;
  if ((rc = binv.alloc((size_t)nblk * BW * BW * 2 * sizeof(double)))) return rc;
  double* ivh = binv.d();
  double* ivl = binv.d() + (size_t)nblk * BW * BW;
  if ((rc = diag_inverses<ArDD>(Lm, n, ivh, ivl, s))) return rc;
  int conv = 0;
  rc = refine_loop(HP_FACTOR_DD, Lm, ivh, ivl, sh, sl, bh, bl, n, tol, max_iter, xh, xl, iters, last, &conv, s);
  if (rc == HP_OK) HP_HIP(hipStreamSynchronize(s));  // binv is freed on return
  return rc;
}


int solve_refined(int kind, const double* sh, const double* sl, const double* bh, const double* bl, long n, double tol,
                  int max_iter, double* xh, double* xl, hp_refine_stats* st, hipStream_t s) {
  using clk = std::chrono::steady_clock;
  std::memset(st, 0, sizeof(*st));
  st->factor_kind = kind;
  if (n == 0) return HP_OK;
  const size_t nn = (size_t)n * n;
  DevBuf L, Sw, inv, err;
  int rc;
  const int limbs = (kind == HP_FACTOR_F64) ? 1 : 2;
  DevBuf binv;
  const long nblk = (n + BW - 1) / BW;
  // the split factorisation works by subtraction in place: no partial-sum plane
  const size_t sw_elems = kind == HP_FACTOR_DD_SPLIT ? 0 : nn * limbs;
  if ((rc = L.alloc(nn * limbs * sizeof(double))) || (rc = Sw.alloc(sw_elems * sizeof(double))) ||
      (rc = inv.alloc(2 * 64 * sizeof(double))) || (rc = err.alloc(sizeof(ErrRec))) ||
      (rc = binv.alloc((size_t)nblk * BW * BW * limbs * sizeof(double))))
    return rc;
  double* ivh = binv.d();
  double* ivl = limbs == 2 ? binv.d() + (size_t)nblk * BW * BW : nullptr;
  HP_HIP(hipMemsetAsync(err.p, 0, sizeof(ErrRec), s));
  HP_HIP(hipStreamSynchronize(s));
  auto t0 = clk::now();
  Mat Lm{L.d(), limbs == 2 ? L.d() + nn : nullptr, n};
  Mat Sm{Sw.d(), limbs == 2 && sw_elems ? Sw.d() + nn : nullptr, n};
  HP_HIP(hipMemcpyAsync(Lm.hi, sh, nn * sizeof(double), hipMemcpyDeviceToDevice, s));
  if (limbs == 2) HP_HIP(hipMemcpyAsync(Lm.lo, sl, nn * sizeof(double), hipMemcpyDeviceToDevice, s));
  ErrRec* er = (ErrRec*)err.p;
  switch (kind) {
    case HP_FACTOR_F64: rc = chol_exact<ArF64>(Lm, n, 64, false, Sm, er, inv.d(), inv.d() + 64, s); break;
    case HP_FACTOR_DD: rc = chol_exact<ArDD>(Lm, n, 64, false, Sm, er, inv.d(), inv.d() + 64, s); break;
    case HP_FACTOR_DD_SPLIT: rc = chol_split(Lm, n, g_split_panel.load(), Sm, er, inv.d(), inv.d() + 64, s); break;
    default: set_error("solve_refined: unknown factor kind %d", kind); return HP_EINVAL;
  }
  if (rc) return rc;
  ErrRec e;
  HP_HIP(hipMemcpyAsync(&e, er, sizeof(e), hipMemcpyDeviceToHost, s));
  HP_HIP(hipStreamSynchronize(s));
  if (!e.flag) {
    rc = (kind == HP_FACTOR_F64) ? diag_inverses<ArF64>(Lm, n, ivh, ivl, s) : diag_inverses<ArDD>(Lm, n, ivh, ivl, s);
    if (rc) return rc;
    HP_HIP(hipStreamSynchronize(s));
  }
  auto t1 = clk::now();
  st->factor_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
  if (e.flag) {
    set_error("Cholesky failed at pivot %lld (value=%.4e)", e.pivot, e.value);
    st->pivot = (size_t)e.pivot;
    return HP_ENOTSPD;
  }
  int it = 0;
  double last = 1.0;
  int conv = 0;
  if ((rc = refine_loop(kind, Lm, ivh, ivl, sh, sl, bh, bl, n, tol, max_iter, xh, xl, &it, &last, &conv, s)))
    return rc;
  auto t2 = clk::now();
  st->converged = conv;
  st->iterations = it;
  st->last_correction = last;
  st->refine_ms = std::chrono::duration<double, std::milli>(t2 - t1).count();
  return HP_OK;
}

}  // namespace hp

namespace hp {
// inverses of the 64 x 64 diagonal blocks of a dd factor, stored [block][i][c]
int diag_block_inverses_dd(Mat l, long n, double* ih, double* il, hipStream_t s) {
  return diag_inverses<ArDD>(l, n, ih, il, s);
}
}  // namespace hp
