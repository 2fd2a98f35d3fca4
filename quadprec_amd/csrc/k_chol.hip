// k_chol.hip -- reference-order (bit-exact) Cholesky factorisations, triangular solves and
// matvec for double-double and binary64.
//
// The reference's unblocked Cholesky (crates/quad/src/linalg.rs:42-96) computes every
// L[i][c] as (a[i][c] - s) * inv(L[c][c]) with s = sum_{k<c} L[i][k]*L[c][k] accumulated in
// increasing k from zero.  This file runs the same arithmetic right-looking: each (i, c)
// keeps its partial sum S[i][c] in HBM, and once column k is final every S[i][c], c > k,
// receives L[i][k]*L[c][k] -- still in increasing k, still one dd add per term -- so every
// output bit equals the reference's, while the O(n^3) part runs as parallel tiles:
//
//   k_diag_factor   a <=64-wide diagonal block, column by column, in one workgroup (LDS)
//   k_row_recur     the rows below a 64-wide panel: one wavefront per row, one lane per
//                   panel column, the finished entry broadcast with v_readlane each step
//   k_rank_update   the trailing lower triangle: 64x64 register-blocked tiles, the panel
//                   staged through LDS, per element the k-ordered dd MAC chain
//
// cholesky_blocked (linalg.rs:238-293) uses the same kernels: its diagonal block is the
// unblocked algorithm on the block, its trsm_right_lt (linalg.rs:224-235) is k_row_recur
// in SUB mode (s = b; s -= b_ik*l_jk; b = s*recip(l_jj)), and its trailing update is
// k_rank_update in SUMSUB mode (s = sum from zero; a -= s).
//
// Binary64 instantiations give the solver crate's cholesky_f64 / solve_cholesky_f64
// (crates/solver/src/lib.rs:169-213).
#include "hp_chol.hpp"
#include "hp_internal.hpp"

namespace hp {
namespace {

constexpr int DB = 64;  // diagonal block / panel width handled by one workgroup

__device__ __forceinline__ double rl(double v, int lane) {  // v_readlane of a double
  long long b = __builtin_bit_cast(long long, v);
  int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), lane);
  int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
  return __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned int)lo);
}
__device__ __forceinline__ dd rl(dd v, int lane) { return mkdd(rl(v.hi, lane), rl(v.lo, lane)); }

template <class A>
struct Lds2 {  // a 64x65 tile of T in LDS as hi / lo planes
  double h[DB][DB + 1];
  double l[A::LIMBS == 2 ? DB : 1][DB + 1];
};
template <class A>
__device__ __forceinline__ typename A::T lget(Lds2<A>& s, int i, int j) {
  if constexpr (A::LIMBS == 2) return mkdd(s.h[i][j], s.l[i][j]);
  else return s.h[i][j];
}
template <class A>
__device__ __forceinline__ void lset(Lds2<A>& s, int i, int j, typename A::T v) {
  if constexpr (A::LIMBS == 2) { s.h[i][j] = v.hi; s.l[i][j] = v.lo; }
  else s.h[i][j] = v;
}

// ---------------------------------------------------------------- diagonal block
// Factor a[r0:r0+jb, r0:r0+jb] (jb <= 64) with the unblocked algorithm.  S_INIT: the
// partial sums start from S (unblocked emulation across panels) instead of zero.  On
// success the lower triangle is written back and inv[j] = recip(L[j][j]); on a
// breakdown the block is left as it was and err records the pivot (linalg.rs:54-57).
// One barrier per column.  Between barriers j and j+1, wavefront 0 finishes column j+1
// alone (lane i holds row i): it adds column j's product to S(i, j+1), takes the pivot on
// lane j+1, broadcasts recip(L[j+1][j+1]) with v_readlane and scales the column; the
// other seven wavefronts add column j's products to the columns beyond j+1 (walked
// through a packed-triangle index table, so every lane has work).  Every element still
// receives its products in increasing k, one dd add each, and is scaled exactly as in
// the unblocked algorithm.  The reciprocals stay in LDS until the block is done, so no
// global store is pending at a barrier.
template <class A, bool S_INIT>
__global__ __launch_bounds__(512) void k_diag_factor(Mat a, Mat S, long r0, int jb, ErrRec* err, double* inv_hi,
                                                     double* inv_lo) {
  using T = typename A::T;
  extern __shared__ __align__(16) unsigned char smem[];
  Lds2<A>& L = *reinterpret_cast<Lds2<A>*>(smem);
  Lds2<A>& Sa = *reinterpret_cast<Lds2<A>*>(smem + sizeof(Lds2<A>));
  __shared__ int fail;
  __shared__ double invs_h[DB], invs_l[DB];
  // packed triangle: entry e = q(q+1)/2 + p, 0 <= p <= q, is the element (jb-1-p, jb-1-q);
  // the columns beyond j+1 form the prefix q <= jb-3-j
  __shared__ unsigned short tri[DB * (DB - 1) / 2];
  const int tid = threadIdx.x, lane = tid & 63;
  if (err->flag) return;
  for (int idx = tid; idx < jb * jb; idx += blockDim.x) {
    int i = idx / jb, c = idx % jb;
    if (c <= i) {
      lset<A>(L, i, c, A::ld(a, r0 + i, r0 + c));
      lset<A>(Sa, i, c, S_INIT ? A::ld(S, r0 + i, r0 + c) : A::zero());
    }
  }
  for (int e = tid; e < DB * (DB - 1) / 2; e += blockDim.x) {
    int q = (int)((__builtin_sqrtf(8.0f * e + 1.0f) - 1.0f) * 0.5f);
    while (q * (q + 1) / 2 > e) --q;
    while ((q + 1) * (q + 2) / 2 <= e) ++q;
    tri[e] = (unsigned short)((e - q * (q + 1) / 2) | (q << 8));
  }
  if (tid == 0) fail = 0;
  __syncthreads();
  // wavefront 0: finish column c whose partial sums (lane i: row i) are in sc
  T lcol = A::zero();
  auto finish_column = [&](int c, T sc) {
    T inv = A::zero();
    int bad = 0;
    if (lane == c) {
      T diag = A::sub(lget<A>(L, c, c), sc);
      if (A::nonpos(diag)) {
        err->flag = 1;
        err->pivot = r0 + c;
        err->value = A::f64(diag);
        fail = 1;
        bad = 1;
      } else {
        const T ljj = A::sqrt(diag);
        inv = A::recip(ljj);
        lset<A>(L, c, c, ljj);
        if constexpr (A::LIMBS == 2) { invs_h[c] = inv.hi; invs_l[c] = inv.lo; }
        else invs_h[c] = inv;
      }
    }
    if (__builtin_amdgcn_readlane(bad, c)) return;
    inv = rl(inv, c);
    if (lane > c && lane < jb) {
      lcol = A::mul(A::sub(lget<A>(L, lane, c), sc), inv);
      lset<A>(L, lane, c, lcol);
    }
  };
  if (tid < 64) finish_column(0, lane < jb ? lget<A>(Sa, lane, 0) : A::zero());
  __syncthreads();
  for (int j = 0; j + 1 < jb; ++j) {
    if (fail) return;
    if (tid < 64) {
      const int c = j + 1;
      T sc = A::zero();
      const T ljc = rl(lcol, c);  // L(j+1, j)
      if (lane >= c && lane < jb) sc = A::add(lget<A>(Sa, lane, c), A::mul(lcol, ljc));
      finish_column(c, sc);
    } else {
      constexpr int NTR = 512 - 64;  // threads of the trailing wavefronts
      const int m = jb - j - 2;  // columns j+2 .. jb-1
      const int cnt = m * (m + 1) / 2;
      for (int e0 = tid - 64; e0 < cnt; e0 += 2 * NTR) {
        T sv[2], li[2], lc[2];
        int ii[2], cc[2];
#pragma unroll
        for (int u = 0; u < 2; ++u) {  // loads of both elements first (index clamped in range)
          const unsigned short pq = tri[min(e0 + u * NTR, cnt - 1)];
          ii[u] = jb - 1 - (pq & 255);
          cc[u] = jb - 1 - (pq >> 8);
          sv[u] = lget<A>(Sa, ii[u], cc[u]);
          li[u] = lget<A>(L, ii[u], j);
          lc[u] = lget<A>(L, cc[u], j);
        }
#pragma unroll
        for (int u = 0; u < 2; ++u)
          if (e0 + u * NTR < cnt) lset<A>(Sa, ii[u], cc[u], A::add(sv[u], A::mul(li[u], lc[u])));
      }
    }
    __syncthreads();
  }
  if (fail) return;
  for (int idx = tid; idx < jb * jb; idx += blockDim.x) {
    int i = idx / jb, c = idx % jb;
    if (c <= i) A::st(a, r0 + i, r0 + c, lget<A>(L, i, c));
  }
  if (tid < jb) {
    inv_hi[tid] = invs_h[tid];
    if constexpr (A::LIMBS == 2) inv_lo[tid] = invs_l[tid];
  }
}

// The same block factor with the partial sums starting from zero (cholesky_blocked's
// diagonal block) and held in registers instead of an LDS plane: element e of the packed
// triangle (e = q(q+1)/2 + p, 0 <= p <= q, is (jb-1-p, jb-1-q)) belongs to trailing thread
// e mod 448 for the whole block (the columns still active form a prefix of the packing, so
// the work stays spread evenly), and reaches wavefront 0 through a two-column LDS buffer
// when its column is next but one.  Same operations in the same order as k_diag_factor
// (bit-identical); one LDS tile (66.5 KB for dd) instead of two, so the block fits on a CU
// beside the trailing updates of the other stream.
template <class A>
__global__ __launch_bounds__(512) void k_diag_factor_rs(Mat a, long r0, int jb, ErrRec* err, double* inv_hi,
                                                        double* inv_lo) {
  using T = typename A::T;
  constexpr int OWN = 5;  // 5 * 448 >= 62 * 63 / 2
  extern __shared__ __align__(16) unsigned char smem[];
  Lds2<A>& L = *reinterpret_cast<Lds2<A>*>(smem);
  __shared__ double Hh[2][DB], Hl[A::LIMBS == 2 ? 2 : 1][DB];
  __shared__ int fail;
  __shared__ double invs_h[DB], invs_l[DB];
  const int tid = threadIdx.x, lane = tid & 63;
  if (err->flag) return;
  for (int idx = tid; idx < jb * jb; idx += blockDim.x) {
    int i = idx / jb, c = idx % jb;
    if (c <= i) lset<A>(L, i, c, A::ld(a, r0 + i, r0 + c));
  }
  if (tid == 0) fail = 0;
  const int cnt_all = (jb - 2) * (jb - 1) / 2;
  T sv[OWN];
  int ei[OWN], ec[OWN];
  if (tid >= 64) {
#pragma unroll
    for (int u = 0; u < OWN; ++u) {
      const int e = tid - 64 + u * (512 - 64);
      int q = (int)((__builtin_sqrtf(8.0f * (float)e + 1.0f) - 1.0f) * 0.5f);
      while (q * (q + 1) / 2 > e) --q;
      while ((q + 1) * (q + 2) / 2 <= e) ++q;
      ei[u] = jb - 1 - (e - q * (q + 1) / 2);
      ec[u] = jb - 1 - q;
      if (e >= cnt_all) {
        ei[u] = 0;
        ec[u] = -1;  // no element
      }
      sv[u] = A::zero();
    }
  }
  __syncthreads();
  // wavefront 0: finish column c whose partial sums (lane i: row i) are in sc
  T lcol = A::zero();
  auto finish_column = [&](int c, T sc) {
    T inv = A::zero();
    int bad = 0;
    if (lane == c) {
      T diag = A::sub(lget<A>(L, c, c), sc);
      if (A::nonpos(diag)) {
        err->flag = 1;
        err->pivot = r0 + c;
        err->value = A::f64(diag);
        fail = 1;
        bad = 1;
      } else {
        const T ljj = A::sqrt(diag);
        inv = A::recip(ljj);
        lset<A>(L, c, c, ljj);
        if constexpr (A::LIMBS == 2) { invs_h[c] = inv.hi; invs_l[c] = inv.lo; }
        else invs_h[c] = inv;
      }
    }
    if (__builtin_amdgcn_readlane(bad, c)) return;
    inv = rl(inv, c);
    if (lane > c && lane < jb) {
      lcol = A::mul(A::sub(lget<A>(L, lane, c), sc), inv);
      lset<A>(L, lane, c, lcol);
    }
  };
  if (tid < 64) finish_column(0, A::zero());
  __syncthreads();
  for (int j = 0; j + 1 < jb; ++j) {
    if (fail) return;
    if (tid < 64) {
      const int c = j + 1;
      T sc = A::zero();
      const T ljc = rl(lcol, c);  // L(j+1, j)
      if (lane >= c && lane < jb) {
        T sin;
        if (c == 1) sin = A::zero();
        else if constexpr (A::LIMBS == 2) sin = mkdd(Hh[c & 1][lane], Hl[c & 1][lane]);
        else sin = Hh[c & 1][lane];
        sc = A::add(sin, A::mul(lcol, ljc));
      }
      finish_column(c, sc);
    } else {
      T li[OWN], lc[OWN];
#pragma unroll
      for (int u = 0; u < OWN; ++u) {  // every operand read first
        li[u] = lget<A>(L, ei[u], j);
        lc[u] = lget<A>(L, max(ec[u], 0), j);
      }
#pragma unroll
      for (int u = 0; u < OWN; ++u) {
        if (ec[u] >= j + 2) {  // columns beyond j+1 take column j
          sv[u] = A::add(sv[u], A::mul(li[u], lc[u]));
          if (ec[u] == j + 2) {  // complete but for column j+1, which wavefront 0 adds
            if constexpr (A::LIMBS == 2) { Hh[ec[u] & 1][ei[u]] = sv[u].hi; Hl[ec[u] & 1][ei[u]] = sv[u].lo; }
            else Hh[ec[u] & 1][ei[u]] = sv[u];
          }
        }
      }
    }
    __syncthreads();
  }
  if (fail) return;
  for (int idx = tid; idx < jb * jb; idx += blockDim.x) {
    int i = idx / jb, c = idx % jb;
    if (c <= i) A::st(a, r0 + i, r0 + c, lget<A>(L, i, c));
  }
  if (tid < jb) {
    inv_hi[tid] = invs_h[tid];
    if constexpr (A::LIMBS == 2) inv_lo[tid] = invs_l[tid];
  }
}

// ---------------------------------------------------------------- diagonal block, split-dd
// The split-dd factorisation's diagonal block (reference order not required).  a[r0:r0+jb,
// r0:r0+jb] (jb <= 64) is reduced column by column: every element takes one compensated
// multiply-subtract per column, T(i, c) -= L(i, j) L(c, j), and column c is finished as
// L(i, c) = T(i, c) * inv(L(c, c)) with the pivot's reciprocal square root from
// rsqrt_dd_fast.  Wavefront 0 (raised wave priority) finishes column j+1 while the seven
// trailing wavefronts apply column j to the columns beyond j+1.  Each trailing thread
// keeps its elements of those columns in registers for the whole block: element e of the
// packed triangle (e = q(q+1)/2 + p, 0 <= p <= q, is (jb-1-p, jb-1-q)) belongs to thread
// e mod 448, so the columns still active after column j (a prefix of the packing) stay
// spread evenly over the threads.  When an element's column is next but one, its owner
// hands it to wavefront 0 through a two-column LDS buffer.  Per column the chain is one
// multiply-subtract, the reciprocal square root, one broadcast and one multiply; the
// trailing work is ~12 FP64 operations per element and column, operands read from the
// finished columns of L in LDS.  On a breakdown the block is left as it was and err
// records the pivot (linalg.rs:54-57).
struct DDacc {  // h + l, not normalised
  double h, l;
};
// a -= x y: the product's leading part by two_sum into a.h, its FMA error and the two cross
// terms into a.l
__device__ __forceinline__ void dd_fms_acc(DDacc& a, dd x, dd y) {
  const double p = x.hi * y.hi;
  double e = __builtin_fma(x.hi, y.hi, -p);
  e = __builtin_fma(x.hi, y.lo, e);
  e = __builtin_fma(x.lo, y.hi, e);
  double s, t;
  two_sum(a.h, -p, s, t);
  a.h = s;
  a.l = a.l + (t - e);
}
__device__ __forceinline__ dd dd_of(DDacc a) {
  dd r;
  two_sum(a.h, a.l, r.hi, r.lo);
  return r;
}
// inv = 1/sqrt(d) and root = sqrt(d) for d > 0, relative error below 2^-103 (2^-103.2 and
// 2^-102.9 worst over 2e6 samples of every binade, subnormals included, against binary128):
// d is scaled by 4^-k into [1, 4) (exact), y = v_rsq_f64 refined by two binary64 Newton
// steps (each at least doubles the correct bits, so a seed better than 2^-14 reaches
// binary64 rounding), then y (1 + e/2 + 3e^2/8) with e = 1 - d y^2 formed exactly enough:
// y^2 = p + pe by TwoProd, 1 - d.hi p by one FMA (|result| ~ 2^-52, rounding ~2^-105), the
// small terms d.hi pe and d.lo p subtracted by FMA.
__device__ __forceinline__ void rsqrt_dd_fast(dd d, dd& root, dd& inv) {
  const int k = (__builtin_amdgcn_frexp_exp(d.hi) - 1) >> 1;  // d.hi = m 2^(2k + {1, 2}) ...
  const dd s = mkdd(__builtin_ldexp(d.hi, -2 * k), __builtin_ldexp(d.lo, -2 * k));  // s.hi in [1, 4)
  double y = __builtin_amdgcn_rsq(s.hi);
  y = y * __builtin_fma(-(0.5 * s.hi) * y, y, 1.5);
  y = y * __builtin_fma(-(0.5 * s.hi) * y, y, 1.5);
  const double p = y * y;
  const double pe = __builtin_fma(y, y, -p);
  double e = __builtin_fma(-s.hi, p, 1.0);
  e = __builtin_fma(-s.hi, pe, e);
  e = __builtin_fma(-s.lo, p, e);
  dd r;
  quick_two_sum(y, y * (e * __builtin_fma(0.375, e, 0.5)), r.hi, r.lo);
  inv = mkdd(__builtin_ldexp(r.hi, -k), __builtin_ldexp(r.lo, -k));
  root = dd_mul(d, inv);
}

constexpr int DF_THREADS = 1024;
constexpr int DF_NTR = DF_THREADS - 64;  // trailing threads
constexpr int DF_OWN = 3;                // elements per trailing thread: 3 * 960 >= 62 * 63 / 2
__device__ __forceinline__ dd ld2(const double2* p) {
  const double2 v = *p;
  return mkdd(v.x, v.y);
}
__global__ __launch_bounds__(DF_THREADS) void k_diag_factor_split(Mat a, long r0, int jb, ErrRec* err, double* inv_hi,
                                                                  double* inv_lo) {
  // the block, then the finished columns of L, as (hi, lo) pairs: one 16-byte LDS access per value
  extern __shared__ __align__(16) unsigned char smem[];
  double2(*L)[DB + 1] = reinterpret_cast<double2(*)[DB + 1]>(smem);
  __shared__ double2 H[2][DB];  // column hand-off to wavefront 0, by column parity
  __shared__ int fail;
  __shared__ double invs_h[DB], invs_l[DB];
  const int tid = threadIdx.x, lane = tid & 63;
  if (err->flag) return;
  for (int idx = tid; idx < jb * jb; idx += blockDim.x) {
    int i = idx / jb, c = idx % jb;
    if (c <= i) L[i][c] = make_double2(a.hi[(r0 + i) * a.ld + r0 + c], a.lo[(r0 + i) * a.ld + r0 + c]);
  }
  if (tid == 0) fail = 0;
  __syncthreads();
  // trailing threads: their elements of columns 2 .. jb-1 into registers
  const int cnt_all = (jb - 2) * (jb - 1) / 2;
  DDacc sv[DF_OWN];
  int ei[DF_OWN], ec[DF_OWN];
  dd t0 = mkdd(0.0, 0.0), t1 = mkdd(0.0, 0.0);
  if (tid >= 64) {
#pragma unroll
    for (int u = 0; u < DF_OWN; ++u) {
      const int e = tid - 64 + u * DF_NTR;
      int q = (int)((__builtin_sqrtf(8.0f * (float)e + 1.0f) - 1.0f) * 0.5f);
      while (q * (q + 1) / 2 > e) --q;
      while ((q + 1) * (q + 2) / 2 <= e) ++q;
      ei[u] = jb - 1 - (e - q * (q + 1) / 2);
      ec[u] = jb - 1 - q;
      sv[u] = DDacc{0.0, 0.0};
      if (e < cnt_all) {
        const dd v = ld2(&L[ei[u]][ec[u]]);
        sv[u] = DDacc{v.hi, v.lo};
      } else {
        ei[u] = 0;
        ec[u] = -1;  // no element
      }
    }
  } else {
    __builtin_amdgcn_s_setprio(3);
    if (lane < jb) t0 = ld2(&L[lane][0]);
    if (lane < jb && jb > 1) t1 = ld2(&L[lane][1]);
  }
  // wavefront 0: finish column c from t = T(lane, c) after every earlier column.  Lane c
  // takes the reciprocal square root; one multiply by it then gives the root on lane c and
  // the column below it on the other lanes.
  dd lcol = mkdd(0.0, 0.0);
  auto finish_column = [&](int c, dd t) {
    dd inv = mkdd(0.0, 0.0);
    if (lane == c) {
      if (ArDD::nonpos(t)) {
        err->flag = 1;
        err->pivot = r0 + c;
        err->value = ArDD::f64(t);
        fail = 1;
      } else {
        dd root;
        rsqrt_dd_fast(t, root, inv);
        invs_h[c] = inv.hi;
        invs_l[c] = inv.lo;
      }
    }
    inv = rl(inv, c);
    const dd v = dd_mul(t, inv);
    if (lane >= c && lane < jb) L[lane][c] = make_double2(v.hi, v.lo);
    if (lane > c) lcol = v;
  };
  if (tid < 64) finish_column(0, t0);
  __syncthreads();
  for (int j = 0; j + 1 < jb; ++j) {
    if (fail) return;
    if (tid < 64) {
      const int c = j + 1;
      const dd ljc = rl(lcol, c);  // L(j+1, j)
      dd t = mkdd(0.0, 0.0);
      if (lane >= c && lane < jb) {
        const dd tin = c == 1 ? t1 : ld2(&H[c & 1][lane]);
        DDacc acc{tin.hi, tin.lo};
        dd_fms_acc(acc, lcol, ljc);
        t = dd_of(acc);
      }
      finish_column(c, t);
    } else {
      dd li[DF_OWN], lc[DF_OWN];
#pragma unroll
      for (int u = 0; u < DF_OWN; ++u) {  // every operand read first
        li[u] = ld2(&L[ei[u]][j]);
        lc[u] = ld2(&L[max(ec[u], 0)][j]);
      }
#pragma unroll
      for (int u = 0; u < DF_OWN; ++u) {
        if (ec[u] >= j + 2) {  // columns beyond j+1 take column j
          dd_fms_acc(sv[u], li[u], lc[u]);
          if (ec[u] == j + 2) {  // final but for column j+1, which wavefront 0 applies
            const dd v = dd_of(sv[u]);
            H[ec[u] & 1][ei[u]] = make_double2(v.hi, v.lo);
          }
        }
      }
    }
    __syncthreads();
  }
  if (fail) return;
  for (int idx = tid; idx < jb * jb; idx += blockDim.x) {
    int i = idx / jb, c = idx % jb;
    if (c <= i) {
      const double2 v = L[i][c];
      a.hi[(r0 + i) * a.ld + r0 + c] = v.x;
      a.lo[(r0 + i) * a.ld + r0 + c] = v.y;
    }
  }
  if (tid < jb) {
    inv_hi[tid] = invs_h[tid];
    inv_lo[tid] = invs_l[tid];
  }
}

// ---------------------------------------------------------------- rows below a panel
// Rows [row_begin, row_end) x panel columns [c0, c0+w), w <= 64, L = the factored
// diagonal block a[c0:c0+w, c0:c0+w], inv[k] = recip(L[k][k]).
//   SUB (trsm_right_lt, linalg.rs:224-235):  t = b_ij; t = t - b_ik*l_jk (k<j); b_ij = t*inv_j
//   ACC (unblocked, linalg.rs:78-82):        t = S_ij; t = t + l_ik*l_jk (k<j); l_ij = (a_ij - t)*inv_j
// One wavefront per row, lane j = panel column j; the rows of a workgroup share the L tile
// (16 for binary64; 4 for dd, so that a small set of rows, whose chains are latency-bound,
// spreads over more CUs).
template <class A>
constexpr int row_recur_rows() { return A::LIMBS == 2 ? 4 : 16; }
// the block's strict lower triangle, packed: (i, k), k < i, at i (i - 1) / 2 + k
constexpr int TRI_S = DB * (DB - 1) / 2;
template <class A>
struct LdsTri {
  double h[TRI_S];
  double l[A::LIMBS == 2 ? TRI_S : 1];
};
template <class A>
__device__ __forceinline__ typename A::T tget(const LdsTri<A>& s, int e) {
  if constexpr (A::LIMBS == 2) return mkdd(s.h[e], s.l[e]);
  else return s.h[e];
}
// CF (split-dd factorisation only, where the order is free; dd, SUB): each lane keeps its
// entry as an unnormalised pair and subtracts b l_jk with the compensated multiply-subtract
// (the product's leading part by two_sum, its FMA error and the cross terms in the second
// double, ~12 FP64 operations against 29 for a dd multiply and subtract); the pair is
// renormalised by two_sum when its column is next.
template <class A, bool ACC, bool CF = false>
__global__ __launch_bounds__(64 * row_recur_rows<A>()) void k_row_recur(Mat a, Mat S, long c0, int w, long row_begin, long row_end,
                                                   const double* inv_hi, const double* inv_lo, const ErrRec* err) {
  static_assert(!CF || (A::LIMBS == 2 && !ACC), "the compensated form is the dd SUB mode");
  using T = typename A::T;
  extern __shared__ __align__(16) unsigned char smem[];
  LdsTri<A>& L = *reinterpret_cast<LdsTri<A>*>(smem);  // 31.5 KB for dd: fits beside the trailing updates
  if (err->flag) return;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  for (int idx = tid; idx < w * w; idx += blockDim.x) {
    int i = idx / w, c = idx % w;
    if (c < i) {
      const T v = A::ld(a, c0 + i, c0 + c);
      if constexpr (A::LIMBS == 2) { L.h[i * (i - 1) / 2 + c] = v.hi; L.l[i * (i - 1) / 2 + c] = v.lo; }
      else L.h[i * (i - 1) / 2 + c] = v;
    }
  }
  __syncthreads();
  T inv = A::zero();
  if (lane < w) inv = A::ldv(inv_hi, inv_lo, lane);
  constexpr int RPB = row_recur_rows<A>();
  const int rowoff = lane * (lane - 1) / 2;  // this lane's row of the packed triangle
  for (long row = row_begin + (long)blockIdx.x * RPB + wv; row < row_end; row += (long)gridDim.x * RPB) {
    if constexpr (CF) {
      double th = 0.0, tl = 0.0;
      dd fin = mkdd(0.0, 0.0);
      if (lane < w) {
        const dd v = A::ld(a, row, c0 + lane);
        th = v.hi;
        tl = v.lo;
      }
      dd lk = tget<A>(L, lane > 0 ? rowoff : 0);
      for (int k = 0; k < w; ++k) {
        const dd lnext = tget<A>(L, k + 1 < lane ? rowoff + k + 1 : 0);
        dd tk;
        two_sum(th, tl, tk.hi, tk.lo);
        const dd bk = A::mul(tk, inv);  // meaningful on lane k
        if (lane == k) fin = bk;
        const dd b = rl(bk, k);
        const double p = b.hi * lk.hi;
        double e = __builtin_fma(b.hi, lk.hi, -p);
        e = __builtin_fma(b.hi, lk.lo, e);
        e = __builtin_fma(b.lo, lk.hi, e);
        double sh, st;
        two_sum(th, -p, sh, st);
        const double nl = tl + (st - e);
        const bool up = lane > k && lane < w;
        th = up ? sh : th;
        tl = up ? nl : tl;
        lk = lnext;
      }
      if (lane < w) A::st(a, row, c0 + lane, fin);
      continue;
    }
    T t = A::zero(), av = A::zero(), fin = A::zero();
    if (lane < w) {
      if (ACC) { t = A::ld(S, row, c0 + lane); av = A::ld(a, row, c0 + lane); }
      else t = A::ld(a, row, c0 + lane);
    }
    // the factor entry of the next step is read from LDS one step ahead, and the update is
    // taken by a select (same operations, every lane), so neither the LDS latency nor a
    // branch sits on the step's dependent chain; lanes at or left of the pivot read entry 0
    T lk = tget<A>(L, lane > 0 ? rowoff : 0);
    for (int k = 0; k < w; ++k) {
      const T lnext = tget<A>(L, k + 1 < lane ? rowoff + k + 1 : 0);
      T bk = ACC ? A::mul(A::sub(av, t), inv) : A::mul(t, inv);  // meaningful on lane k
      if (lane == k) fin = bk;
      T b = rl(bk, k);
      const T p = A::mul(b, lk);
      const T u = ACC ? A::add(t, p) : A::sub(t, p);
      if constexpr (A::LIMBS == 2) {
        const bool up = lane > k && lane < w;
        t.hi = up ? u.hi : t.hi;
        t.lo = up ? u.lo : t.lo;
      } else {
        t = (lane > k && lane < w) ? u : t;
      }
      lk = lnext;
    }
    if (lane < w) A::st(a, row, c0 + lane, fin);
  }
}

// ---------------------------------------------------------------- rank-k updates
// out(i, j), i in [ri0, ri0+M), j in [cj0, cj0+N), with operands
//   p(i, k) = a(i, ka0+k),  q(j, k) = a(bj0 + j - cj0, ka0+k),  k in [0, K):
//   SUMSUB (cholesky_blocked trailing, linalg.rs:268-279): s = 0; s += p*q (k order); a -= s
//   ACC    (unblocked emulation, linalg.rs:78-82):         S += p*q (k order)
//   SUBINC (trsm_right_lt across 64-column sub-panels, linalg.rs:228-231): a -= p*q (k order)
// TRI restricts to j <= i (ri0 == cj0).  64x64 tiles, 4x4 register block per thread, the
// operands staged through LDS in 16-deep k slabs.
enum { RU_SUMSUB = 0, RU_ACC = 1, RU_SUBINC = 2 };
constexpr int TB = 64, TK = 16;

// RB x RB outputs per thread on a 16 x 16 thread grid: tiles of TB = 16 RB.  Skinny
// updates (the next panel's columns, the sub-panel updates) use RB = 1 or 2 so that
// enough workgroups exist; each element's chain is the same either way.
template <class A, int MODE, bool TRI, int RB>
__global__ __launch_bounds__(256) void k_rank_update(Mat a, Mat S, long ri0, long M, long cj0, long N, long bj0,
                                                     long ka0, long K, const ErrRec* err) {
  using T = typename A::T;
  constexpr int TBr = 16 * RB;
  const int ti = blockIdx.y, tc = blockIdx.x;
  if (TRI && tc > ti) return;
  if (err->flag) return;
  __shared__ double As[2][TK][TBr], Bs[2][TK][TBr];
  const int t = threadIdx.x, tx = t & 15, ty = t >> 4;
  const long i0 = (long)ti * TBr, j0 = (long)tc * TBr;  // tile origin relative to ri0 / cj0
  T acc[RB][RB];
#pragma unroll
  for (int r = 0; r < RB; ++r)
#pragma unroll
    for (int c = 0; c < RB; ++c) {
      long li = i0 + ty * RB + r, lj = j0 + tx * RB + c;
      bool in = li < M && lj < N && (!TRI || ri0 + li >= cj0 + lj);
      T v = A::zero();
      if (in && MODE == RU_ACC) v = A::ld(S, ri0 + li, cj0 + lj);
      if (in && MODE == RU_SUBINC) v = A::ld(a, ri0 + li, cj0 + lj);
      acc[r][c] = v;
    }
  for (long k0 = 0; k0 < K; k0 += TK) {
    {  // thread t stages tile row t / (16/RB), k (t % (16/RB)) * RB .. + RB-1 of both operands
      constexpr int TPR = 16 / RB;
      int r = t / TPR, kc = (t % TPR) * RB;
#pragma unroll
      for (int q = 0; q < RB; ++q) {
        long kk = k0 + kc + q;
        T va = (i0 + r < M && kk < K) ? A::ld(a, ri0 + i0 + r, ka0 + kk) : A::zero();
        T vb = (j0 + r < N && kk < K) ? A::ld(a, bj0 + j0 + r, ka0 + kk) : A::zero();
        if constexpr (A::LIMBS == 2) {
          As[0][kc + q][r] = va.hi; As[1][kc + q][r] = va.lo;
          Bs[0][kc + q][r] = vb.hi; Bs[1][kc + q][r] = vb.lo;
        } else {
          As[0][kc + q][r] = va; Bs[0][kc + q][r] = vb;
        }
      }
    }
    __syncthreads();
    const int kmax = (int)min((long)TK, K - k0);
    for (int kk = 0; kk < kmax; ++kk) {
      T av[RB], bv[RB];
#pragma unroll
      for (int r = 0; r < RB; ++r) {
        if constexpr (A::LIMBS == 2) {
          av[r] = mkdd(As[0][kk][ty * RB + r], As[1][kk][ty * RB + r]);
          bv[r] = mkdd(Bs[0][kk][tx * RB + r], Bs[1][kk][tx * RB + r]);
        } else {
          av[r] = As[0][kk][ty * RB + r];
          bv[r] = Bs[0][kk][tx * RB + r];
        }
      }
#pragma unroll
      for (int r = 0; r < RB; ++r)
#pragma unroll
        for (int c = 0; c < RB; ++c) {
          T p = A::mul(av[r], bv[c]);
          acc[r][c] = (MODE == RU_SUBINC) ? A::sub(acc[r][c], p) : A::add(acc[r][c], p);
        }
    }
    __syncthreads();
  }
#pragma unroll
  for (int r = 0; r < RB; ++r)
#pragma unroll
    for (int c = 0; c < RB; ++c) {
      long li = i0 + ty * RB + r, lj = j0 + tx * RB + c;
      if (li < M && lj < N && (!TRI || ri0 + li >= cj0 + lj)) {
        long gi = ri0 + li, gj = cj0 + lj;
        if (MODE == RU_ACC) A::st(S, gi, gj, acc[r][c]);
        else if (MODE == RU_SUBINC) A::st(a, gi, gj, acc[r][c]);
        else A::st(a, gi, gj, A::sub(A::ld(a, gi, gj), acc[r][c]));
      }
    }
}

// inv[j] = recip(L[c0+j][c0+j]) for j < w (trsm_right_lt, linalg.rs:226)
template <class A>
__global__ void k_diag_recip(Mat a, long c0, int w, double* inv_h, double* inv_l) {
  int j = threadIdx.x;
  if (j < w) A::stv(inv_h, inv_l, j, A::recip(A::ld(a, c0 + j, c0 + j)));
}

// ---------------------------------------------------------------- helpers
template <class A>
__global__ void k_zero_lower(Mat S, long r0, long len) {  // S[r0+i][r0+j] = 0 for j <= i
  long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  long total = len * len;
  for (; idx < total; idx += (long)gridDim.x * blockDim.x) {
    long i = idx / len, j = idx % len;
    if (j <= i) A::st(S, r0 + i, r0 + j, A::zero());
  }
}
__global__ void k_zero_upper_f64(double* l, long n) {  // cholesky_f64 zeroes the upper triangle (lib.rs:190-192)
  long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  for (; idx < n * n; idx += (long)gridDim.x * blockDim.x)
    if (idx % n > idx / n) l[idx] = 0.0;
}

// ---------------------------------------------------------------- forward solve
// linalg.rs:99-108: y_i = y_i - l_ij*y_j (j ascending), then y_i = y_i / l_ii.
// Blocks of 64 rows: the diagonal block in one wavefront, then the rows below receive
// the block's contributions in j order.
template <class A>
__global__ __launch_bounds__(64) void k_fwd_diag(Mat l, double* yh, double* yl, long j0, int bw) {
  using T = typename A::T;
  extern __shared__ __align__(16) unsigned char smem[];
  Lds2<A>& L = *reinterpret_cast<Lds2<A>*>(smem);
  const int lane = threadIdx.x;
  for (int i = 0; i < bw; ++i)
    if (lane <= i) lset<A>(L, i, lane, A::ld(l, j0 + i, j0 + lane));
  __syncthreads();
  T y = lane < bw ? A::ldv(yh, yl, j0 + lane) : A::zero();
  for (int j = 0; j < bw; ++j) {
    T yj = A::div(y, lget<A>(L, lane < bw ? lane : 0, lane < bw ? lane : 0));  // meaningful on lane j
    if (lane == j) y = yj;
    T b = rl(yj, j);
    if (lane > j && lane < bw) y = A::sub(y, A::mul(lget<A>(L, lane, j), b));
  }
  if (lane < bw) A::stv(yh, yl, j0 + lane, y);
}
template <class A>
__global__ __launch_bounds__(256) void k_fwd_update(Mat l, double* yh, double* yl, long j0, int bw, long n) {
  using T = typename A::T;
  __shared__ double Th[64][DB + 1], Tl[A::LIMBS == 2 ? 64 : 1][DB + 1];
  __shared__ double Yh[DB], Yl[DB];
  const int tid = threadIdx.x;
  const long rbase = j0 + bw + (long)blockIdx.x * 64;
  if (tid < bw) { Yh[tid] = yh[j0 + tid]; if constexpr (A::LIMBS == 2) Yl[tid] = yl[j0 + tid]; }
  for (int idx = tid; idx < 64 * bw; idx += 256) {
    int r = idx / bw, c = idx % bw;
    long gi = rbase + r;
    T v = gi < n ? A::ld(l, gi, j0 + c) : A::zero();
    if constexpr (A::LIMBS == 2) { Th[r][c] = v.hi; Tl[r][c] = v.lo; }
    else Th[r][c] = v;
  }
  __syncthreads();
  if (tid < 64 && rbase + tid < n) {
    long gi = rbase + tid;
    T y = A::ldv(yh, yl, gi);
    for (int j = 0; j < bw; ++j) {
      T lv, yv;
      if constexpr (A::LIMBS == 2) { lv = mkdd(Th[tid][j], Tl[tid][j]); yv = mkdd(Yh[j], Yl[j]); }
      else { lv = Th[tid][j]; yv = Yh[j]; }
      y = A::sub(y, A::mul(lv, yv));
    }
    A::stv(yh, yl, gi, y);
  }
}

// ---------------------------------------------------------------- backward solve
// linalg.rs:111-120: x_i = x_i - l_ji*x_j for j = i+1..n-1 ascending, then x_i / l_ii.
// The chain of row i begins with x_{i+1}, the last value computed before it, so the
// n(n-1)/2 dependent dd subtractions cannot be reordered without changing the result.
// One workgroup: the products l_ji*x_j of a chunk are formed in parallel into LDS and
// thread 0 runs the subtraction chain.
template <class A>
__global__ __launch_bounds__(256) void k_bwd_exact(Mat l, double* xh, double* xl, long n) {
  using T = typename A::T;
  constexpr int G = 8;  // the chain's group of products (read ahead up to one group past a chunk)
  __shared__ double Ph[256 + G], Pl[A::LIMBS == 2 ? 256 + G : 1];
  const int tid = threadIdx.x;
  for (long i = n - 1; i >= 0; --i) {
    T s = A::zero();
    if (tid == 0) s = A::ldv(xh, xl, i);
    for (long jc = i + 1; jc < n; jc += 256) {
      long j = jc + tid;
      if (j < n) {
        T p = A::mul(A::ld(l, j, i), A::ldv(xh, xl, j));
        if constexpr (A::LIMBS == 2) { Ph[tid] = p.hi; Pl[tid] = p.lo; }
        else Ph[tid] = p;
      }
      __syncthreads();
      if (tid == 0) {  // the chain, in groups of G products: each group is read from LDS
                       // while the previous one is subtracted (two register sets, no copies)
        const int cnt = (int)min((long)256, n - jc);
        auto rd = [&](int q) -> T {
          if constexpr (A::LIMBS == 2) return mkdd(Ph[q], Pl[q]);
          else return Ph[q];
        };
        T ga[G], gb[G];
#pragma unroll
        for (int u = 0; u < G; ++u) ga[u] = rd(u);
        int q = 0;
        for (; q + 2 * G <= cnt; q += 2 * G) {
#pragma unroll
          for (int u = 0; u < G; ++u) gb[u] = rd(q + G + u);
          __builtin_amdgcn_sched_barrier(0);
#pragma unroll
          for (int u = 0; u < G; ++u) s = A::sub(s, ga[u]);
#pragma unroll
          for (int u = 0; u < G; ++u) ga[u] = rd(q + 2 * G + u);  // q + 2G + u < cnt + G
          __builtin_amdgcn_sched_barrier(0);
#pragma unroll
          for (int u = 0; u < G; ++u) s = A::sub(s, gb[u]);
        }
        if (q + G <= cnt) {
#pragma unroll
          for (int u = 0; u < G; ++u) s = A::sub(s, ga[u]);
          q += G;
        }
        for (; q < cnt; ++q) s = A::sub(s, rd(q));
      }
      __syncthreads();
    }
    if (tid == 0) {
      T x = A::div(s, A::ld(l, i, i));
      A::stv(xh, xl, i, x);
      __threadfence_block();
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------- matvec
// linalg.rs:129-147: y_i = sum_j a_ij*x_j, sequential from zero.  64 rows per
// wavefront, the row tile staged through LDS so HBM reads stay coalesced.
template <class A>
__global__ __launch_bounds__(64) void k_matvec(Mat a, const double* xh, const double* xl, double* yh, double* yl,
                                               long n) {
  using T = typename A::T;
  __shared__ double Th[64][65], Tl[A::LIMBS == 2 ? 64 : 1][65];
  __shared__ double Xh[64], Xl[64];
  const int lane = threadIdx.x;
  const long r0 = (long)blockIdx.x * 64;
  T s = A::zero();
  for (long c0 = 0; c0 < n; c0 += 64) {
    int cw = (int)min((long)64, n - c0);
    for (int r = 0; r < 64; ++r) {
      long gi = r0 + r;
      T v = (gi < n && lane < cw) ? A::ld(a, gi, c0 + lane) : A::zero();
      if constexpr (A::LIMBS == 2) { Th[r][lane] = v.hi; Tl[r][lane] = v.lo; }
      else Th[r][lane] = v;
    }
    if (lane < cw) { Xh[lane] = xh[c0 + lane]; if constexpr (A::LIMBS == 2) Xl[lane] = xl[c0 + lane]; }
    __syncthreads();
    for (int q = 0; q < cw; ++q) {
      T av, xv;
      if constexpr (A::LIMBS == 2) { av = mkdd(Th[lane][q], Tl[lane][q]); xv = mkdd(Xh[q], Xl[q]); }
      else { av = Th[lane][q]; xv = Xh[q]; }
      s = A::add(s, A::mul(av, xv));
    }
    __syncthreads();
  }
  if (r0 + lane < n) A::stv(yh, yl, r0 + lane, s);
}

}  // namespace

// ================================================================ host orchestration
template <class A>
static int set_lds_attrs() {
  const int big = (int)(2 * sizeof(Lds2<A>));
  const int one = (int)sizeof(Lds2<A>);
  HP_HIP(hipFuncSetAttribute((const void*)k_diag_factor<A, true>, hipFuncAttributeMaxDynamicSharedMemorySize, big));
  HP_HIP(hipFuncSetAttribute((const void*)k_diag_factor_rs<A>, hipFuncAttributeMaxDynamicSharedMemorySize, one));
  if constexpr (A::LIMBS == 2)
    HP_HIP(hipFuncSetAttribute((const void*)k_diag_factor_split, hipFuncAttributeMaxDynamicSharedMemorySize, one));
  HP_HIP(hipFuncSetAttribute((const void*)k_row_recur<A, true>, hipFuncAttributeMaxDynamicSharedMemorySize, one));
  HP_HIP(hipFuncSetAttribute((const void*)k_row_recur<A, false>, hipFuncAttributeMaxDynamicSharedMemorySize, one));
  if constexpr (A::LIMBS == 2)
    HP_HIP(hipFuncSetAttribute((const void*)k_row_recur<A, false, true>, hipFuncAttributeMaxDynamicSharedMemorySize,
                               one));
  HP_HIP(hipFuncSetAttribute((const void*)k_fwd_diag<A>, hipFuncAttributeMaxDynamicSharedMemorySize, one));
  return HP_OK;
}
template <class A>
static int ensure_attrs() {
  static int rc = set_lds_attrs<A>();  // once per process (function attributes are global)
  return rc;
}

template <class A, int MODE, bool TRI, int RB>
static int rank_update_rb(Mat a, Mat S, long ri0, long M, long cj0, long N, long bj0, long ka0, long K, ErrRec* err,
                          hipStream_t s) {
  constexpr long TBr = 16 * RB;
  dim3 grid((unsigned)((N + TBr - 1) / TBr), (unsigned)((M + TBr - 1) / TBr));
  hipLaunchKernelGGL((k_rank_update<A, MODE, TRI, RB>), grid, dim3(256), 0, s, a, S, ri0, M, cj0, N, bj0, ka0, K,
                     (const ErrRec*)err);
  HP_LAUNCH_CHECK();
  return HP_OK;
}
// the register block is chosen so that the launch has enough workgroups to fill the chip
template <class A, int MODE, bool TRI>
static int rank_update(Mat a, Mat S, long ri0, long M, long cj0, long N, long bj0, long ka0, long K, ErrRec* err,
                       hipStream_t s) {
  if (M <= 0 || N <= 0) return HP_OK;
  auto tiles = [&](long tb) {
    long tm = (M + tb - 1) / tb, tn = (N + tb - 1) / tb;
    return TRI ? tm * (tn + 1) / 2 : tm * tn;
  };
  if (tiles(64) >= 1024) return rank_update_rb<A, MODE, TRI, 4>(a, S, ri0, M, cj0, N, bj0, ka0, K, err, s);
  if (tiles(32) >= 1024) return rank_update_rb<A, MODE, TRI, 2>(a, S, ri0, M, cj0, N, bj0, ka0, K, err, s);
  return rank_update_rb<A, MODE, TRI, 1>(a, S, ri0, M, cj0, N, bj0, ka0, K, err, s);
}

template <class A, bool ACC, bool CF = false>
static int row_recur(Mat a, Mat S, long c0, int w, long rb, long re, double* inv_h, double* inv_l, ErrRec* err,
                     hipStream_t s) {
  if (re <= rb) return HP_OK;
  constexpr int RPB = row_recur_rows<A>();
  unsigned g = (unsigned)std::min<long>((re - rb + RPB - 1) / RPB, 8192 / RPB);
  hipLaunchKernelGGL((k_row_recur<A, ACC, CF>), dim3(g), dim3(64 * RPB), sizeof(LdsTri<A>), s, a, S, c0, w, rb, re,
                     inv_h, inv_l, (const ErrRec*)err);
  HP_LAUNCH_CHECK();
  return HP_OK;
}

// Unblocked Cholesky (linalg.rs:42-96) of a[r0:r0+len, r0:r0+len], S = partial sums.
template <class A>
static int factor_unblocked_region(Mat a, Mat S, long r0, long len, ErrRec* err, double* inv_h, double* inv_l,
                                   hipStream_t s) {
  {
    long total = len * len;
    unsigned g = (unsigned)std::min<long>((total + 255) / 256, 65536);
    hipLaunchKernelGGL(k_zero_lower<A>, dim3(g ? g : 1), dim3(256), 0, s, S, r0, len);
    HP_LAUNCH_CHECK();
  }
  for (long c0 = r0; c0 < r0 + len; c0 += DB) {
    int w = (int)std::min<long>(DB, r0 + len - c0);
    hipLaunchKernelGGL((k_diag_factor<A, true>), dim3(1), dim3(512), 2 * sizeof(Lds2<A>), s, a, S, c0, w, err,
                       inv_h, inv_l);
    HP_LAUNCH_CHECK();
    long rb = c0 + w, re = r0 + len;
    int rc;
    if ((rc = row_recur<A, true>(a, S, c0, w, rb, re, inv_h, inv_l, err, s))) return rc;
    if ((rc = rank_update<A, RU_ACC, true>(a, S, rb, re - rb, rb, re - rb, rb, c0, w, err, s))) return rc;
  }
  return HP_OK;
}

// One block column of cholesky_blocked (linalg.rs:247-261): the diagonal block
// [c0, c0+jb) by the unblocked algorithm, then trsm_right_lt on rows [c0+jb, n).
template <class A>
int chol_panel_exact(Mat a, Mat S, long c0, long jb, long n, ErrRec* err, double* inv_h, double* inv_l,
                     hipStream_t s) {
  int rc = ensure_attrs<A>();
  if (rc) return rc;
  if (jb <= DB) {
    hipLaunchKernelGGL((k_diag_factor_rs<A>), dim3(1), dim3(512), sizeof(Lds2<A>), s, a, c0, (int)jb, err, inv_h,
                       inv_l);
    HP_LAUNCH_CHECK();
  } else if ((rc = factor_unblocked_region<A>(a, S, c0, jb, err, inv_h, inv_l, s))) {
    return rc;
  }
  long rem = n - c0 - jb;
  if (rem <= 0) return HP_OK;
  for (long q0 = 0; q0 < jb; q0 += DB) {
    int w = (int)std::min<long>(DB, jb - q0);
    if (q0 > 0 && (rc = rank_update<A, RU_SUBINC, false>(a, S, c0 + jb, rem, c0 + q0, w, c0 + q0, c0, q0, err, s)))
      return rc;
    if (jb > DB) {
      hipLaunchKernelGGL(k_diag_recip<A>, dim3(1), dim3(64), 0, s, a, c0 + q0, w, inv_h, inv_l);
      HP_LAUNCH_CHECK();
    }
    if ((rc = row_recur<A, false>(a, S, c0 + q0, w, c0 + jb, n, inv_h, inv_l, err, s))) return rc;
  }
  return HP_OK;
}

// The split-dd factorisation's panel (reference order not required): each 64-wide
// diagonal sub-block by k_diag_factor_split, the panel's rows below it solved and updated
// by subtraction (t = a; t -= l_ik l_jk; l_ij = t / l_jj), as in trsm_right_lt.  The rows
// below the panel are solved on a side stream, sub-panel q as soon as its diagonal block
// is factored, overlapping the factor of the next diagonal sub-block; each solved
// sub-panel then updates the panel's later columns of those rows right away, so after the
// last diagonal block only its own row solve remains.  The reciprocals of sub-panel q live
// in slot q of inv (2 x 64 doubles per slot).  evq: 16 events, [q] after sub-panel q's
// diagonal block, [8 + q] after its rows within the panel; evdone is recorded on `side`
// once the panel is complete and `s` waits for it.
// k_solve.hip: order-free updates a(r, c) -= sum_k a(r, c0 + k) a(c, c0 + k) of the lower
// triangle of the block at (t0, t0), width w, or of the rows ri0.. x columns cj0..
int diag_syrk_split(Mat a, long t0, long w, long c0, long K, ErrRec* err, hipStream_t s);
int rect_syrk_split(Mat a, long ri0, long M, long cj0, long N, long c0, long K, ErrRec* err, hipStream_t s);

template <class A>
int chol_panel_split(Mat a, Mat S, long c0, long jb, long n, ErrRec* err, double* inv, hipStream_t s,
                     hipStream_t side, hipEvent_t* evq, hipEvent_t evdone) {
  static_assert(A::LIMBS == 2, "the split-dd panel is double-double");
  int rc = ensure_attrs<A>();
  if (rc) return rc;
  const long rem = n - c0 - jb;
  for (long q0 = 0; q0 < jb; q0 += DB) {
    const int w = (int)std::min<long>(DB, jb - q0);
    const int q = (int)(q0 / DB);
    double* ih = inv + (size_t)q * 2 * DB;
    double* il = ih + DB;
    hipLaunchKernelGGL(k_diag_factor_split, dim3(1), dim3(DF_THREADS), sizeof(Lds2<A>), s, a, c0 + q0, w, err, ih, il);
    HP_LAUNCH_CHECK();
    HP_HIP(hipEventRecord(evq[q], s));
    const long rb = c0 + q0 + w, re = c0 + jb;
    if (re > rb) {
      if ((rc = row_recur<A, false, true>(a, S, c0 + q0, w, rb, re, ih, il, err, s))) return rc;
      HP_HIP(hipEventRecord(evq[8 + q], s));
      // the panel's later columns of those rows (lower triangle), its summation order free:
      // partial sums in parallel instead of one 64-long dd chain per element
      if ((rc = diag_syrk_split(a, rb, re - rb, c0 + q0, w, err, s))) return rc;
    }
    if (rem > 0) {
      HP_HIP(hipStreamWaitEvent(side, evq[q], 0));
      if ((rc = row_recur<A, false, true>(a, S, c0 + q0, w, c0 + jb, n, ih, il, err, side))) return rc;
      if (re > rb) {  // L[rb:re, sub-panel q] comes from the row solve on s
        HP_HIP(hipStreamWaitEvent(side, evq[8 + q], 0));
        if ((rc = rect_syrk_split(a, c0 + jb, rem, rb, re - rb, c0 + q0, w, err, side))) return rc;
      }
    }
  }
  HP_HIP(hipEventRecord(evdone, side));
  HP_HIP(hipStreamWaitEvent(s, evdone, 0));
  return HP_OK;
}
template int chol_panel_split<ArDD>(Mat, Mat, long, long, long, ErrRec*, double*, hipStream_t, hipStream_t,
                                    hipEvent_t*, hipEvent_t);

// Second stream and events for the factorisations' lookahead (default priority).
struct Lookahead {
  hipStream_t s2 = nullptr;
  hipEvent_t evA = nullptr, evB = nullptr;
  bool evB_live = false;
  int init() {
    HP_HIP(hipStreamCreateWithFlags(&s2, hipStreamNonBlocking));
    HP_HIP(hipEventCreateWithFlags(&evA, hipEventDisableTiming));
    HP_HIP(hipEventCreateWithFlags(&evB, hipEventDisableTiming));
    return HP_OK;
  }
  ~Lookahead() {
    if (s2) { (void)hipStreamSynchronize(s2); (void)hipStreamDestroy(s2); }
    if (evA) (void)hipEventDestroy(evA);
    if (evB) (void)hipEventDestroy(evB);
  }
};

template <class A>
int rank_update_sumsub(Mat a, Mat S, long ri0, long M, long cj0, long N, long ka0, long K, ErrRec* err,
                       hipStream_t s) {
  return rank_update<A, RU_SUMSUB, true>(a, S, ri0, M, cj0, N, cj0, ka0, K, err, s);
}
// the same update on a block strictly below the diagonal (rows ri0.., columns cj0..)
template <class A>
int rank_update_sumsub_below(Mat a, Mat S, long ri0, long M, long cj0, long N, long ka0, long K, ErrRec* err,
                             hipStream_t s) {
  return rank_update<A, RU_SUMSUB, false>(a, S, ri0, M, cj0, N, cj0, ka0, K, err, s);
}

// cholesky_blocked with lookahead: after panel k, the next panel's columns receive panel
// k's update first (stream s); the rest of the trailing update runs on a second stream
// while panel k+1 is factored.  Every element still receives its updates in panel order
// (the strip update of step k+1 waits for the rest update of step k), so the result is
// bit-identical to the sequential loop of linalg.rs:247-291.
template <class A>
int chol_exact(Mat a, long n, long nb, bool unblocked, Mat S, ErrRec* err, double* inv_h, double* inv_l,
               hipStream_t s) {
  int rc = ensure_attrs<A>();
  if (rc) return rc;
  if (n == 0) return HP_OK;
  if (nb == 0) nb = 64;  // linalg.rs:241
  if (unblocked || nb >= n) return factor_unblocked_region<A>(a, S, 0, n, err, inv_h, inv_l, s);
  Lookahead la;
  if ((rc = la.init())) return rc;
  for (long jb0 = 0; jb0 < n; jb0 += nb) {
    long jb = std::min(nb, n - jb0);
    if ((rc = chol_panel_exact<A>(a, S, jb0, jb, n, err, inv_h, inv_l, s))) return rc;
    long rem = n - jb0 - jb;
    if (rem <= 0) continue;
    long t0 = jb0 + jb, w2 = std::min(nb, rem);
    if (la.evB_live) HP_HIP(hipStreamWaitEvent(s, la.evB, 0));
    // next panel's columns (linalg.rs:263-288 restricted to them)
    if ((rc = rank_update<A, RU_SUMSUB, true>(a, S, t0, rem, t0, w2, t0, jb0, jb, err, s))) return rc;
    if (rem > w2) {
      HP_HIP(hipEventRecord(la.evA, s));
      HP_HIP(hipStreamWaitEvent(la.s2, la.evA, 0));
      long r0 = t0 + w2;
      if ((rc = rank_update<A, RU_SUMSUB, true>(a, S, r0, n - r0, r0, n - r0, r0, jb0, jb, err, la.s2))) return rc;
      HP_HIP(hipEventRecord(la.evB, la.s2));
      la.evB_live = true;
    }
  }
  if (la.evB_live) HP_HIP(hipStreamWaitEvent(s, la.evB, 0));
  return HP_OK;
}

template <class A>
int fwd_exact(Mat l, double* yh, double* yl, long n, hipStream_t s) {
  int rc = ensure_attrs<A>();
  if (rc) return rc;
  for (long j0 = 0; j0 < n; j0 += DB) {
    int bw = (int)std::min<long>(DB, n - j0);
    hipLaunchKernelGGL(k_fwd_diag<A>, dim3(1), dim3(64), sizeof(Lds2<A>), s, l, yh, yl, j0, bw);
    HP_LAUNCH_CHECK();
    long rem = n - j0 - bw;
    if (rem > 0) {
      hipLaunchKernelGGL(k_fwd_update<A>, dim3((unsigned)((rem + 63) / 64)), dim3(256), 0, s, l, yh, yl, j0, bw, n);
      HP_LAUNCH_CHECK();
    }
  }
  return HP_OK;
}

template <class A>
int bwd_exact(Mat l, double* xh, double* xl, long n, hipStream_t s) {
  if (n == 0) return HP_OK;
  hipLaunchKernelGGL(k_bwd_exact<A>, dim3(1), dim3(256), 0, s, l, xh, xl, n);
  HP_LAUNCH_CHECK();
  return HP_OK;
}

template <class A>
int matvec_exact(Mat a, const double* xh, const double* xl, double* yh, double* yl, long n, hipStream_t s) {
  if (n == 0) return HP_OK;
  hipLaunchKernelGGL(k_matvec<A>, dim3((unsigned)((n + 63) / 64)), dim3(64), 0, s, a, xh, xl, yh, yl, n);
  HP_LAUNCH_CHECK();
  return HP_OK;
}

int zero_upper_f64(double* l, long n, hipStream_t s) {
  unsigned g = (unsigned)std::min<long>((n * n + 255) / 256, 65536);
  hipLaunchKernelGGL(k_zero_upper_f64, dim3(g ? g : 1), dim3(256), 0, s, l, n);
  HP_LAUNCH_CHECK();
  return HP_OK;
}

template int chol_exact<ArDD>(Mat, long, long, bool, Mat, ErrRec*, double*, double*, hipStream_t);
template int chol_panel_exact<ArDD>(Mat, Mat, long, long, long, ErrRec*, double*, double*, hipStream_t);
template int chol_panel_exact<ArF64>(Mat, Mat, long, long, long, ErrRec*, double*, double*, hipStream_t);
template int rank_update_sumsub<ArDD>(Mat, Mat, long, long, long, long, long, long, ErrRec*, hipStream_t);
template int rank_update_sumsub_below<ArDD>(Mat, Mat, long, long, long, long, long, long, ErrRec*, hipStream_t);
template int chol_exact<ArF64>(Mat, long, long, bool, Mat, ErrRec*, double*, double*, hipStream_t);
template int fwd_exact<ArDD>(Mat, double*, double*, long, hipStream_t);
template int fwd_exact<ArF64>(Mat, double*, double*, long, hipStream_t);
template int bwd_exact<ArDD>(Mat, double*, double*, long, hipStream_t);
template int bwd_exact<ArF64>(Mat, double*, double*, long, hipStream_t);
template int matvec_exact<ArDD>(Mat, const double*, const double*, double*, double*, long, hipStream_t);
template int matvec_exact<ArF64>(Mat, const double*, const double*, double*, double*, long, hipStream_t);

}  // namespace hp
