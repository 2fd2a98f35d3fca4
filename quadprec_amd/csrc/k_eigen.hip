// k_eigen.hip -- symmetric dd eigensolver, canonical orthogonalization and the
// generalized eigenproblem F C = S C eps on the GPU (the consumer of the overlap matrix:
// crates/quad/src/linalg.rs:299-362, crates/solver/src/ortho.rs:61-248).
//
// The reference rotates one (p, q) pair at a time (cyclic sweeps in linalg.rs, the
// largest |a_pq| in ortho.rs), a serial chain of O(n) updates.  Here every step of a sweep
// applies n/2 disjoint rotations at once (round-robin order, every pair once per sweep):
//   * per pair, the reference's rotation (tau = (a_qq - a_pp) / 2 a_pq,
//     t = sign(tau) / (|tau| + sqrt(1 + tau^2)), c = 1 / sqrt(1 + t^2), s = t c) and the
//     new diagonal a_pp - t a_pq, a_qq + t a_pq (linalg.rs:323-333);
//   * k_jac_step: every 2 x 2 block (I, J) becomes R_I^T X R_J (the diagonal blocks take
//     the new diagonal and exact zeros), the lower blocks in the mirrored order of
//     operations so that A stays exactly symmetric; the eigenvector matrix takes V R.  The
//     rotations are formed inside the step by the threads that use them, one launch per step.
// The matrix is kept in "pair layout": the two members of pair i sit at positions 2i and
// 2i+1, so each step reads adjacent columns, and the apply kernel writes its results
// straight into the next step's layout (the circle rotation sigma).  Odd n is padded by a
// zero row and column, whose rotations are the identity.
//
// Rotations differ in order from the reference, so the results agree to dd rounding, not
// bit for bit: eigenvalues to a few units of 2^-104 |A|, eigenvectors up to sign (and up
// to rotations inside a degenerate eigenspace).
#include <algorithm>
#include <atomic>
#include <climits>
#include <cmath>
#include <cstring>
#include <numeric>
#include <vector>

#include "hp_arith.hpp"
#include "hp_internal.hpp"

namespace hp {
namespace {

// position of an element after one round-robin step (circle method: the member at
// position 0 stays, the others move one place around the circle t1 .. t_{m-1}, b_{m-1} .. b0)
__host__ __device__ __forceinline__ long rr_sigma(long pos, long m) {
  if (m == 1 || pos == 0) return pos;
  const long i = pos >> 1;
  if ((pos & 1) == 0) return i <= m - 2 ? 2 * i + 2 : 2 * m - 1;
  return i == 0 ? 2 : 2 * i - 1;
}

struct Rot {
  double ch, cl, sh, sl;  // c, s
  double ph, pl, qh, ql;  // new a_pp, a_qq
  int active;             // 0: identity (a_pq is zero or negligible), block left as is
  int pad;
};

// The reference's rotation of a pair (linalg.rs:323-333): c, s and the new diagonal; false
// (the identity) when a_pq is zero, or below 1e-150 of the diagonal gap (|tau| > 1e150,
// where tau^2 would overflow and the rotation is the identity to dd precision)
__device__ __forceinline__ bool jac_rotation(dd app, dd aqq, dd apq, dd& c, dd& s, dd& np, dd& nq) {
  c = mkdd(1.0);
  s = mkdd(0.0);
  np = app;
  nq = aqq;
  if (apq.hi == 0.0) return false;
  const dd one = mkdd(1.0);
  const dd tau = dd_div(dd_sub(aqq, app), dd_mul(mkdd(2.0), apq));
  if (!(__builtin_fabs(tau.hi) <= 1e150)) return false;
  const dd tau_abs = dd_abs(tau);
  const dd t_denom = dd_add(tau_abs, dd_sqrt(dd_add(one, dd_mul(tau, tau))));
  const dd t = tau.hi >= 0.0 ? dd_div(one, t_denom) : dd_div(dd_neg(one), t_denom);
  c = dd_div(one, dd_sqrt(dd_add(one, dd_mul(t, t))));
  s = dd_mul(t, c);
  np = dd_sub(app, dd_mul(t, apq));
  nq = dd_add(aqq, dd_mul(t, apq));
  return true;
}

// the rotation of pair i (positions 2i, 2i+1) from A's diagonal 2 x 2 block
__device__ __forceinline__ Rot pair_rot(const double* __restrict__ ah, const double* __restrict__ al, long N2,
                                        long i) {
  const long p = 2 * i, q = 2 * i + 1;
  const dd app = mkdd(ah[p * N2 + p], al[p * N2 + p]);
  const dd aqq = mkdd(ah[q * N2 + q], al[q * N2 + q]);
  const dd apq = mkdd(ah[p * N2 + q], al[p * N2 + q]);
  Rot r;
  r.pad = 0;
  dd c, s, np, nq;
  r.active = jac_rotation(app, aqq, apq, c, s, np, nq) ? 1 : 0;
  r.ch = c.hi; r.cl = c.lo; r.sh = s.hi; r.sl = s.lo;
  r.ph = np.hi; r.pl = np.lo; r.qh = nq.hi; r.ql = nq.lo;
  return r;
}

// One round-robin step, one launch.  z = 0: A (pair layout) -> B (next layout); a block
// takes the 2 x 2 blocks (I, J) of 4 WI pairs I by 64 pairs J (lane = J, each wavefront WI
// of the I).  z = 1: V -> W; a block takes 64 pairs by 4 WR rows (lane = pair, each
// wavefront WR rows).
// The rotations a block uses are formed once, one per lane in parallel, and shared through
// LDS.  An upper block (I < J) takes its columns by R_J and then its rows by R_I; a lower
// block takes its rows by R_I and then its columns by R_J, which is the upper computation
// transposed operation for operation, so A stays exactly symmetric without mirror writes.
// WI pairs I per wavefront (z = 0), WR rows per wavefront (z = 1); HOIST: load the blocks
// before forming the rotations (hides the load latency when few blocks run; costs registers)
// rot (optional): the z = 0 blocks of the first row of blocks store the step's rotations
// (c.hi, c.lo, s.hi, s.lo per pair) for k_jac_vbatch
template <int WI, int WR, bool HOIST>
__global__ __launch_bounds__(256) void k_jac_step(const double* __restrict__ ah, const double* __restrict__ al,
                                                  double* __restrict__ bh, double* __restrict__ bl,
                                                  const double* __restrict__ vh, const double* __restrict__ vl,
                                                  double* __restrict__ wh, double* __restrict__ wl, long N2,
                                                  double* __restrict__ rot) {
  constexpr int ST_I = 4 * WI, ST_R = 4 * WR;
  __shared__ Rot rs[64 + ST_I];
  const long m = N2 / 2;
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  if (blockIdx.z == 1) {
    const long i0 = (long)blockIdx.x * 64, r0 = (long)blockIdx.y * ST_R;
    if (r0 >= N2) return;  // the whole block
    const long i = i0 + lane;
    dd v0[HOIST ? WR : 1], v1[HOIST ? WR : 1];  // the rotations do not depend on them
    if constexpr (HOIST) {
#pragma unroll
      for (int k = 0; k < WR; ++k) {
        const long r = r0 + w * WR + k;
        const bool ok = i < m && r < N2;
        v0[k] = ok ? mkdd(vh[r * N2 + 2 * i], vl[r * N2 + 2 * i]) : mkdd(0.0);
        v1[k] = ok ? mkdd(vh[r * N2 + 2 * i + 1], vl[r * N2 + 2 * i + 1]) : mkdd(0.0);
      }
    }
    if (t < 64 && i0 + t < m) rs[t] = pair_rot(ah, al, N2, i0 + t);
    __syncthreads();
    if (i >= m) return;
    const Rot R = rs[lane];
    const dd c = mkdd(R.ch, R.cl), s = mkdd(R.sh, R.sl);
    const long o0 = rr_sigma(2 * i, m), o1 = rr_sigma(2 * i + 1, m);
#pragma unroll
    for (int k = 0; k < WR; ++k) {
      const long r = r0 + w * WR + k;
      if (r >= N2) continue;
      const dd a0 = HOIST ? v0[HOIST ? k : 0] : mkdd(vh[r * N2 + 2 * i], vl[r * N2 + 2 * i]);
      const dd a1 = HOIST ? v1[HOIST ? k : 0] : mkdd(vh[r * N2 + 2 * i + 1], vl[r * N2 + 2 * i + 1]);
      const dd w0 = dd_sub(dd_mul(c, a0), dd_mul(s, a1));
      const dd w1 = dd_add(dd_mul(s, a0), dd_mul(c, a1));
      wh[r * N2 + o0] = w0.hi; wl[r * N2 + o0] = w0.lo;
      wh[r * N2 + o1] = w1.hi; wl[r * N2 + o1] = w1.lo;
    }
    return;
  }
  const long J0 = (long)blockIdx.x * 64, I0 = (long)blockIdx.y * ST_I;
  if (I0 >= m) return;  // the whole block
  const long J = J0 + lane;
  const long q0 = 2 * J, q1 = q0 + 1;
  dd xs[HOIST ? WI : 1][4];  // the rotations do not depend on them
  if constexpr (HOIST) {
#pragma unroll
    for (int k = 0; k < WI; ++k) {
      const long I = I0 + w * WI + k;
      const bool ok = I < m && J < m;
      const long p0 = 2 * I, p1 = p0 + 1;
      xs[k][0] = ok ? mkdd(ah[p0 * N2 + q0], al[p0 * N2 + q0]) : mkdd(0.0);
      xs[k][1] = ok ? mkdd(ah[p0 * N2 + q1], al[p0 * N2 + q1]) : mkdd(0.0);
      xs[k][2] = ok ? mkdd(ah[p1 * N2 + q0], al[p1 * N2 + q0]) : mkdd(0.0);
      xs[k][3] = ok ? mkdd(ah[p1 * N2 + q1], al[p1 * N2 + q1]) : mkdd(0.0);
    }
  }
  if (t < 64) {
    if (J0 + t < m) {
      rs[t] = pair_rot(ah, al, N2, J0 + t);
      if (rot && blockIdx.y == 0) {
        double* q = rot + 4 * (J0 + t);
        q[0] = rs[t].ch; q[1] = rs[t].cl; q[2] = rs[t].sh; q[3] = rs[t].sl;
      }
    }
  } else if (t < 64 + ST_I) {
    if (I0 + t - 64 < m) rs[t] = pair_rot(ah, al, N2, I0 + t - 64);
  }
  __syncthreads();
  if (J >= m) return;
  const Rot RJ = rs[lane];
  const dd cJ = mkdd(RJ.ch, RJ.cl), sJ = mkdd(RJ.sh, RJ.sl);
  const long nq0 = rr_sigma(q0, m), nq1 = rr_sigma(q1, m);
#pragma unroll
  for (int k = 0; k < WI; ++k) {
    const long I = I0 + w * WI + k;
    if (I >= m) continue;
    const Rot RI = rs[64 + w * WI + k];
    const long p0 = 2 * I, p1 = p0 + 1;
    dd x00, x01, x10, x11;
    if constexpr (HOIST) {
      x00 = xs[k][0]; x01 = xs[k][1]; x10 = xs[k][2]; x11 = xs[k][3];
    } else {
      x00 = mkdd(ah[p0 * N2 + q0], al[p0 * N2 + q0]);
      x01 = mkdd(ah[p0 * N2 + q1], al[p0 * N2 + q1]);
      x10 = mkdd(ah[p1 * N2 + q0], al[p1 * N2 + q0]);
      x11 = mkdd(ah[p1 * N2 + q1], al[p1 * N2 + q1]);
    }
    const long np0 = rr_sigma(p0, m), np1 = rr_sigma(p1, m);
    dd y00, y01, y10, y11;
    if (I == J) {
      if (RI.active) {
        x00 = mkdd(RI.ph, RI.pl);
        x11 = mkdd(RI.qh, RI.ql);
        x01 = mkdd(0.0);
        x10 = mkdd(0.0);
      }
      y00 = x00; y01 = x01; y10 = x10; y11 = x11;
    } else {
      const dd cI = mkdd(RI.ch, RI.cl), sI = mkdd(RI.sh, RI.sl);
      if (I < J) {  // columns q0, q1 by rotation J, then rows p0, p1 by rotation I
        const dd c00 = dd_sub(dd_mul(cJ, x00), dd_mul(sJ, x01)), c01 = dd_add(dd_mul(sJ, x00), dd_mul(cJ, x01));
        const dd c10 = dd_sub(dd_mul(cJ, x10), dd_mul(sJ, x11)), c11 = dd_add(dd_mul(sJ, x10), dd_mul(cJ, x11));
        y00 = dd_sub(dd_mul(cI, c00), dd_mul(sI, c10));
        y10 = dd_add(dd_mul(sI, c00), dd_mul(cI, c10));
        y01 = dd_sub(dd_mul(cI, c01), dd_mul(sI, c11));
        y11 = dd_add(dd_mul(sI, c01), dd_mul(cI, c11));
      } else {  // rows p0, p1 by rotation I, then columns q0, q1 by rotation J
        const dd r00 = dd_sub(dd_mul(cI, x00), dd_mul(sI, x10)), r10 = dd_add(dd_mul(sI, x00), dd_mul(cI, x10));
        const dd r01 = dd_sub(dd_mul(cI, x01), dd_mul(sI, x11)), r11 = dd_add(dd_mul(sI, x01), dd_mul(cI, x11));
        y00 = dd_sub(dd_mul(cJ, r00), dd_mul(sJ, r01));
        y01 = dd_add(dd_mul(sJ, r00), dd_mul(cJ, r01));
        y10 = dd_sub(dd_mul(cJ, r10), dd_mul(sJ, r11));
        y11 = dd_add(dd_mul(sJ, r10), dd_mul(cJ, r11));
      }
    }
    bh[np0 * N2 + nq0] = y00.hi; bl[np0 * N2 + nq0] = y00.lo;
    bh[np0 * N2 + nq1] = y01.hi; bl[np0 * N2 + nq1] = y01.lo;
    bh[np1 * N2 + nq0] = y10.hi; bl[np1 * N2 + nq0] = y10.lo;
    bh[np1 * N2 + nq1] = y11.hi; bl[np1 * N2 + nq1] = y11.lo;
  }
}

// V <- V R_1 ... R_cnt for cnt consecutive steps' rotations (as stored by k_jac_step),
// each step followed by its layout permutation: the operations of k_jac_step's V part,
// step after step, on VR rows of V held in LDS (two buffers), so V is read and written
// once per cnt steps instead of once per step.  Rows are updated in place.
template <int VR>
__global__ __launch_bounds__(256) void k_jac_vbatch(double* __restrict__ vh, double* __restrict__ vl,
                                                    const double* __restrict__ rot, int cnt, long N2) {
  extern __shared__ double vsm[];  // [2 buffers][hi, lo][VR][N2]
  const long m = N2 / 2, r0 = (long)blockIdx.x * VR;
  const int t = threadIdx.x;
  auto at = [&](int buf, int part, int rr, long c) -> double& { return vsm[((size_t)(buf * 2 + part) * VR + rr) * N2 + c]; };
  for (long idx = t; idx < (long)VR * N2; idx += 256) {
    const int rr = (int)(idx / N2);
    const long c = idx % N2, row = r0 + rr;
    at(0, 0, rr, c) = row < N2 ? vh[row * N2 + c] : 0.0;
    at(0, 1, rr, c) = row < N2 ? vl[row * N2 + c] : 0.0;
  }
  __syncthreads();
  int cur = 0;
  for (int st = 0; st < cnt; ++st) {
    const double* rs = rot + (size_t)st * 4 * m;
    for (long idx = t; idx < (long)VR * m; idx += 256) {
      const int rr = (int)(idx / m);
      const long i = idx % m;
      const dd c = mkdd(rs[4 * i], rs[4 * i + 1]), s = mkdd(rs[4 * i + 2], rs[4 * i + 3]);
      const dd a0 = mkdd(at(cur, 0, rr, 2 * i), at(cur, 1, rr, 2 * i));
      const dd a1 = mkdd(at(cur, 0, rr, 2 * i + 1), at(cur, 1, rr, 2 * i + 1));
      const dd w0 = dd_sub(dd_mul(c, a0), dd_mul(s, a1));
      const dd w1 = dd_add(dd_mul(s, a0), dd_mul(c, a1));
      const long o0 = rr_sigma(2 * i, m), o1 = rr_sigma(2 * i + 1, m);
      at(cur ^ 1, 0, rr, o0) = w0.hi; at(cur ^ 1, 1, rr, o0) = w0.lo;
      at(cur ^ 1, 0, rr, o1) = w1.hi; at(cur ^ 1, 1, rr, o1) = w1.lo;
    }
    __syncthreads();
    cur ^= 1;
  }
  for (long idx = t; idx < (long)VR * N2; idx += 256) {
    const int rr = (int)(idx / N2);
    const long c = idx % N2, row = r0 + rr;
    if (row < N2) {
      vh[row * N2 + c] = at(cur, 0, rr, c);
      vl[row * N2 + c] = at(cur, 1, rr, c);
    }
  }
}

// off-diagonal size of A: sum of hi^2 over a < b (for the Frobenius stop test) and
// max |hi + lo| (for the largest-element stop test)
__global__ __launch_bounds__(256) void k_jac_off(const double* __restrict__ ah, const double* __restrict__ al, long N2,
                                                 double* __restrict__ sum, unsigned long long* __restrict__ mx) {
  __shared__ double ss[4];
  __shared__ double sm[4];
  const long a = blockIdx.y;
  const long b = (long)blockIdx.x * blockDim.x + threadIdx.x;
  double v2 = 0.0, va = 0.0;
  if (b > a && b < N2) {
    const double h = ah[a * N2 + b];
    v2 = h * h;
    va = __builtin_fabs(h + al[a * N2 + b]);
  }
  for (int o = 32; o > 0; o >>= 1) {
    v2 += __shfl_xor(v2, o, 64);
    va = fmax(va, __shfl_xor(va, o, 64));
  }
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    ss[w] = v2;
    sm[w] = va;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const double t = ss[0] + ss[1] + ss[2] + ss[3];
    const double u = fmax(fmax(sm[0], sm[1]), fmax(sm[2], sm[3]));
    if (t > 0.0) atomicAdd(sum, t);
    if (u > 0.0) atomicMax(mx, (unsigned long long)__double_as_longlong(u));  // >= 0: ordered as integers
  }
}

// B (N2 x N2, zero padded) <- A (n x n); V <- identity
__global__ void k_jac_init(const double* __restrict__ ah, const double* __restrict__ al, long n, long N2,
                           double* __restrict__ bh, double* __restrict__ bl, double* __restrict__ vh,
                           double* __restrict__ vl) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= N2 * N2) return;
  const long r = idx / N2, c = idx % N2;
  const bool in = r < n && c < n;
  bh[idx] = in ? ah[r * n + c] : 0.0;
  bl[idx] = in ? al[r * n + c] : 0.0;
  vh[idx] = r == c ? 1.0 : 0.0;
  vl[idx] = 0.0;
}

// outputs from the pair layout: evals[o] = A[pos(o)][pos(o)]; vec_rows: evecs[o][r] =
// V[r][pos(o)] (row o = eigenvector o), else evecs[r][o] (column o)
__global__ void k_jac_extract(const double* __restrict__ ah, const double* __restrict__ al,
                              const double* __restrict__ vh, const double* __restrict__ vl, long n, long N2,
                              const long* __restrict__ pos, double* __restrict__ eh, double* __restrict__ el,
                              double* __restrict__ xh, double* __restrict__ xl, int vec_rows) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n * n) return;
  const long o = idx / n, r = idx % n;
  const long po = pos[o];
  const long out = vec_rows ? o * n + r : r * n + o;
  xh[out] = vh[r * N2 + po];
  xl[out] = vl[r * N2 + po];
  if (r == 0) {
    eh[o] = ah[po * N2 + po];
    el[o] = al[po * N2 + po];
  }
}

unsigned blocks(long n, int t = 256) { return (unsigned)((n + t - 1) / t); }

// ---- block Jacobi (large n).  A is cut into 16 x 16 blocks, nb of them per side (n padded
// with zero rows and columns to a multiple of 32, whose rotations are the identity).  An
// outer sweep runs nb - 1 steps of the round-robin order over the blocks; at a step each of
// the nb / 2 block pairs (P, R) takes its 32 x 32 subproblem [A_PP A_PR; A_RP A_RR] through
// one inner sweep of the scalar parallel Jacobi (31 steps of 16 disjoint rotations, their
// angles to binary64 accuracy: jac_rotation_fast) in LDS, accumulating them into an orthogonal Q
// (k_bj_sub); then A <- Q^T A Q and V <- V Q with Q block diagonal over the pairs, as
// 32 x 32 dd products (k_bj_apply, k_bj_vapply).  Every index pair meets in at least one
// subproblem per outer sweep, as in a scalar sweep, and a sweep makes nb - 1 passes over A
// instead of n - 1.
constexpr int BJ_B = 16, BJ_S = 2 * BJ_B;

// round-robin (circle method) pairing of n players, round t in [0, n-1), match k in
// [0, n/2): every pair exactly once in n - 1 rounds; p < q
__host__ __device__ __forceinline__ void rr_match(int t, int k, int n, int& p, int& q) {
  int a, b;
  if (k == 0) {
    a = t;
    b = n - 1;
  } else {
    a = (t + k) % (n - 1);
    b = (t - k + (n - 1)) % (n - 1);
  }
  p = a < b ? a : b;
  q = a < b ? b : a;
}

// The block form's rotation: the angle to binary64 accuracy (tau from the dd difference of
// the diagonal, t in binary64), then c = 1 / sqrt(1 + t^2) and s = t c in dd, so the
// rotation is orthogonal to dd precision and annihilates a_pq to about 2^-53 of itself; the
// pair's 2 x 2 block is then rotated in full (not replaced by the closed-form diagonal).
// A Jacobi rotation need not take the exact angle: any orthogonal rotation keeps the
// similarity exact, and the sweeps drive the remaining a_pq down.  One dd division and one
// dd square root instead of three and two.
__device__ __forceinline__ bool jac_rotation_fast(dd app, dd aqq, dd apq, dd& c, dd& s) {
  c = mkdd(1.0);
  s = mkdd(0.0);
  if (apq.hi == 0.0) return false;
  const dd d = dd_sub(aqq, app);
  const double tau = (d.hi + d.lo) / (2.0 * (apq.hi + apq.lo));
  if (!(__builtin_fabs(tau) <= 1e150)) return false;
  const double t = (tau >= 0.0 ? 1.0 : -1.0) / (__builtin_fabs(tau) + sqrt(1.0 + tau * tau));
  const dd td = mkdd(t);
  c = dd_div(mkdd(1.0), dd_sqrt(dd_add(mkdd(1.0), dd_mul(td, td))));
  s = dd_mul(td, c);
  return true;
}

// global index of local row i of pair (P, R)'s subproblem
__device__ __forceinline__ long bj_g(int i, int P, int R) { return i < BJ_B ? (long)P * BJ_B + i : (long)R * BJ_B + i - BJ_B; }

// 2 x 2 block (rows p0, q0; columns p1, q1) of a 32 x 32 LDS matrix through the rotations
// (c0, s0) on the rows and (c1, s1) on the columns: R0^T X R1 with R = [c s; -s c];
// rows_first computes (R0^T X) R1, else R0^T (X R1): the two are transposes of each other
// operation for operation, which keeps a symmetric matrix exactly symmetric
__device__ __forceinline__ void bj_rot2(dd& x00, dd& x01, dd& x10, dd& x11, dd c0, dd s0, dd c1, dd s1,
                                         bool rows_first) {
  if (rows_first) {
    const dd y00 = dd_sub(dd_mul(c0, x00), dd_mul(s0, x10)), y01 = dd_sub(dd_mul(c0, x01), dd_mul(s0, x11));
    const dd y10 = dd_add(dd_mul(s0, x00), dd_mul(c0, x10)), y11 = dd_add(dd_mul(s0, x01), dd_mul(c0, x11));
    x00 = dd_sub(dd_mul(c1, y00), dd_mul(s1, y01));
    x01 = dd_add(dd_mul(s1, y00), dd_mul(c1, y01));
    x10 = dd_sub(dd_mul(c1, y10), dd_mul(s1, y11));
    x11 = dd_add(dd_mul(s1, y10), dd_mul(c1, y11));
  } else {
    const dd y00 = dd_sub(dd_mul(c1, x00), dd_mul(s1, x01)), y01 = dd_add(dd_mul(s1, x00), dd_mul(c1, x01));
    const dd y10 = dd_sub(dd_mul(c1, x10), dd_mul(s1, x11)), y11 = dd_add(dd_mul(s1, x10), dd_mul(c1, x11));
    x00 = dd_sub(dd_mul(c0, y00), dd_mul(s0, y10));
    x01 = dd_sub(dd_mul(c0, y01), dd_mul(s0, y11));
    x10 = dd_add(dd_mul(s0, y00), dd_mul(c0, y10));
    x11 = dd_add(dd_mul(s0, y01), dd_mul(c0, y11));
  }
}

// one inner sweep of pair w's subproblem; Q (row-major 32 x 32 dd) to q_h / q_l + w * 1024.
// S is double-buffered (each step reads one copy and writes the other), and every Q entry
// is read and written by the one thread that owns it, so a step takes two barriers.
__global__ __launch_bounds__(256) void k_bj_sub(const double* __restrict__ ah, const double* __restrict__ al, long Nb,
                                                const int2* __restrict__ pairs, double* __restrict__ q_h,
                                                double* __restrict__ q_l) {
  __shared__ double Sh[2][BJ_S][BJ_S + 1], Sl[2][BJ_S][BJ_S + 1];
  __shared__ double Qh[BJ_S][BJ_S + 1], Ql[BJ_S][BJ_S + 1];
  __shared__ double rc[4][BJ_B];  // c.hi, c.lo, s.hi, s.lo per match
  __shared__ int ract[BJ_B];
  const int tid = threadIdx.x;
  const int2 pr = pairs[blockIdx.x];
  for (int e = tid; e < BJ_S * BJ_S; e += 256) {
    const int i = e / BJ_S, j = e % BJ_S;
    const long o = bj_g(i, pr.x, pr.y) * Nb + bj_g(j, pr.x, pr.y);
    Sh[0][i][j] = ah[o];
    Sl[0][i][j] = al[o];
    Qh[i][j] = i == j ? 1.0 : 0.0;
    Ql[i][j] = 0.0;
  }
  __syncthreads();
  const int I = tid >> 4, J = tid & 15;  // this thread's 2 x 2 block (matches I, J)
  for (int t = 0; t < BJ_S - 1; ++t) {
    const int cur = t & 1, nxt = cur ^ 1;
    if (tid < BJ_B) {
      int p, q;
      rr_match(t, tid, BJ_S, p, q);
      dd c, sn;
      ract[tid] = jac_rotation_fast(mkdd(Sh[cur][p][p], Sl[cur][p][p]), mkdd(Sh[cur][q][q], Sl[cur][q][q]),
                                    mkdd(Sh[cur][p][q], Sl[cur][p][q]), c, sn) ? 1 : 0;
      rc[0][tid] = c.hi; rc[1][tid] = c.lo; rc[2][tid] = sn.hi; rc[3][tid] = sn.lo;
    }
    __syncthreads();
    int p0, q0, p1, q1;
    rr_match(t, I, BJ_S, p0, q0);
    rr_match(t, J, BJ_S, p1, q1);
    dd x00 = mkdd(Sh[cur][p0][p1], Sl[cur][p0][p1]), x01 = mkdd(Sh[cur][p0][q1], Sl[cur][p0][q1]);
    dd x10 = mkdd(Sh[cur][q0][p1], Sl[cur][q0][p1]), x11 = mkdd(Sh[cur][q0][q1], Sl[cur][q0][q1]);
    if (ract[I] || ract[J]) {
      bj_rot2(x00, x01, x10, x11, mkdd(rc[0][I], rc[1][I]), mkdd(rc[2][I], rc[3][I]), mkdd(rc[0][J], rc[1][J]),
              mkdd(rc[2][J], rc[3][J]), I > J);
      if (I == J) x10 = x01;  // the pair's own block: symmetric exactly
    }
    Sh[nxt][p0][p1] = x00.hi; Sl[nxt][p0][p1] = x00.lo; Sh[nxt][p0][q1] = x01.hi; Sl[nxt][p0][q1] = x01.lo;
    Sh[nxt][q0][p1] = x10.hi; Sl[nxt][q0][p1] = x10.lo; Sh[nxt][q0][q1] = x11.hi; Sl[nxt][q0][q1] = x11.lo;
    // Q <- Q R: rows 2J, 2J+1 of the columns of match I (this thread's entries only)
    if (ract[I]) {
      const dd cq = mkdd(rc[0][I], rc[1][I]), sq = mkdd(rc[2][I], rc[3][I]);
#pragma unroll
      for (int e = 0; e < 2; ++e) {
        const int r = 2 * J + e;
        const dd a = mkdd(Qh[r][p0], Ql[r][p0]), b = mkdd(Qh[r][q0], Ql[r][q0]);
        const dd na = dd_sub(dd_mul(cq, a), dd_mul(sq, b)), nb2 = dd_add(dd_mul(sq, a), dd_mul(cq, b));
        Qh[r][p0] = na.hi; Ql[r][p0] = na.lo;
        Qh[r][q0] = nb2.hi; Ql[r][q0] = nb2.lo;
      }
    }
    __syncthreads();
  }
  double* oh = q_h + (size_t)blockIdx.x * BJ_S * BJ_S;
  double* ol = q_l + (size_t)blockIdx.x * BJ_S * BJ_S;
  for (int e = tid; e < BJ_S * BJ_S; e += 256) {
    oh[e] = Qh[e / BJ_S][e % BJ_S];
    ol[e] = Ql[e / BJ_S][e % BJ_S];
  }
}

// X <- Q_i^T X Q_j for the 32 x 32 block of A on pair i's rows and pair j's columns, i <= j
// (the linear block index over the upper triangle of pairs), the transpose written to the
// mirror (i < j); the diagonal blocks keep their upper triangle and mirror it
__global__ __launch_bounds__(256) void k_bj_apply(double* __restrict__ ah, double* __restrict__ al, long Nb,
                                                  const int2* __restrict__ pairs, const double* __restrict__ q_h,
                                                  const double* __restrict__ q_l) {
  __shared__ double Xh[BJ_S][BJ_S + 1], Xl[BJ_S][BJ_S + 1];
  __shared__ double Ah[BJ_S][BJ_S + 1], Al[BJ_S][BJ_S + 1];  // Q_i
  __shared__ double Bh[BJ_S][BJ_S + 1], Bl[BJ_S][BJ_S + 1];  // Q_j
  long j = (long)((__builtin_sqrt(8.0 * (double)blockIdx.x + 1.0) - 1.0) / 2.0);
  while ((j + 1) * (j + 2) / 2 <= (long)blockIdx.x) ++j;
  while (j * (j + 1) / 2 > (long)blockIdx.x) --j;
  const long i = (long)blockIdx.x - j * (j + 1) / 2;  // i <= j
  const int2 pi = pairs[i], pj = pairs[j];
  const int tid = threadIdx.x;
  for (int e = tid; e < BJ_S * BJ_S; e += 256) {
    const int r = e / BJ_S, c = e % BJ_S;
    const long o = bj_g(r, pi.x, pi.y) * Nb + bj_g(c, pj.x, pj.y);
    Xh[r][c] = ah[o];
    Xl[r][c] = al[o];
    Ah[r][c] = q_h[i * BJ_S * BJ_S + e];
    Al[r][c] = q_l[i * BJ_S * BJ_S + e];
    Bh[r][c] = q_h[j * BJ_S * BJ_S + e];
    Bl[r][c] = q_l[j * BJ_S * BJ_S + e];
  }
  __syncthreads();
  const int r = tid >> 3, c0 = (tid & 7) * 4;  // this thread's four entries of row r
  dd acc[4];
#pragma unroll
  for (int e = 0; e < 4; ++e) acc[e] = mkdd(0.0);
  for (int k = 0; k < BJ_S; ++k) {  // T = X Q_j
    const dd x = mkdd(Xh[r][k], Xl[r][k]);
#pragma unroll
    for (int e = 0; e < 4; ++e) acc[e] = dd_add(acc[e], dd_mul(x, mkdd(Bh[k][c0 + e], Bl[k][c0 + e])));
  }
  __syncthreads();
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    Xh[r][c0 + e] = acc[e].hi;
    Xl[r][c0 + e] = acc[e].lo;
    acc[e] = mkdd(0.0);
  }
  __syncthreads();
  for (int k = 0; k < BJ_S; ++k) {  // Z = Q_i^T T
    const dd qa = mkdd(Ah[k][r], Al[k][r]);
#pragma unroll
    for (int e = 0; e < 4; ++e) acc[e] = dd_add(acc[e], dd_mul(qa, mkdd(Xh[k][c0 + e], Xl[k][c0 + e])));
  }
  const long gr = bj_g(r, pi.x, pi.y);
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    const int c = c0 + e;
    const long gc = bj_g(c, pj.x, pj.y);
    if (i < j || r <= c) {
      ah[gr * Nb + gc] = acc[e].hi;
      al[gr * Nb + gc] = acc[e].lo;
      if (i < j || r < c) {
        ah[gc * Nb + gr] = acc[e].hi;
        al[gc * Nb + gr] = acc[e].lo;
      }
    }
  }
}

// V <- V Q_j on pair j's 32 columns for rows 32 blockIdx.x .. +31
__global__ __launch_bounds__(256) void k_bj_vapply(double* __restrict__ vh, double* __restrict__ vl, long Nb,
                                                   const int2* __restrict__ pairs, const double* __restrict__ q_h,
                                                   const double* __restrict__ q_l) {
  __shared__ double Th[BJ_S][BJ_S + 1], Tl[BJ_S][BJ_S + 1];
  __shared__ double Bh[BJ_S][BJ_S + 1], Bl[BJ_S][BJ_S + 1];
  const long j = blockIdx.y, r0 = (long)blockIdx.x * BJ_S;
  const int2 pj = pairs[j];
  const int tid = threadIdx.x;
  for (int e = tid; e < BJ_S * BJ_S; e += 256) {
    const int r = e / BJ_S, c = e % BJ_S;
    const long o = (r0 + r) * Nb + bj_g(c, pj.x, pj.y);
    Th[r][c] = vh[o];
    Tl[r][c] = vl[o];
    Bh[r][c] = q_h[j * BJ_S * BJ_S + e];
    Bl[r][c] = q_l[j * BJ_S * BJ_S + e];
  }
  __syncthreads();
  const int r = tid >> 3, c0 = (tid & 7) * 4;
  dd acc[4];
#pragma unroll
  for (int e = 0; e < 4; ++e) acc[e] = mkdd(0.0);
  for (int k = 0; k < BJ_S; ++k) {
    const dd x = mkdd(Th[r][k], Tl[r][k]);
#pragma unroll
    for (int e = 0; e < 4; ++e) acc[e] = dd_add(acc[e], dd_mul(x, mkdd(Bh[k][c0 + e], Bl[k][c0 + e])));
  }
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    const long o = (r0 + r) * Nb + bj_g(c0 + e, pj.x, pj.y);
    vh[o] = acc[e].hi;
    vl[o] = acc[e].lo;
  }
}

// X[i][k] = V[i][sel[k]] * recip(sqrt(lambda_k)) (ortho.rs:185-190), V's columns the eigenvectors
__global__ void k_ortho_x(const double* __restrict__ vh, const double* __restrict__ vl, long n,
                          const long* __restrict__ sel, const double* __restrict__ lam, long m, double* __restrict__ xh,
                          double* __restrict__ xl) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n * m) return;
  const long i = idx / m, k = idx % m;
  const dd inv_sqrt = dd_recip(dd_sqrt(mkdd(lam[2 * k], lam[2 * k + 1])));
  const dd v = mkdd(vh[i * n + sel[k]], vl[i * n + sel[k]]);
  const dd x = dd_mul(v, inv_sqrt);
  xh[idx] = x.hi;
  xl[idx] = x.lo;
}
// F'[i][j] = F'[j][i] = (F'[i][j] + F'[j][i]) * 0.5 for i < j (ortho.rs:218-225)
__global__ void k_symmetrize(double* __restrict__ h, double* __restrict__ l, long m) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= m * m) return;
  const long i = idx / m, j = idx % m;
  if (j <= i) return;
  const dd avg = dd_mul(dd_add(mkdd(h[i * m + j], l[i * m + j]), mkdd(h[j * m + i], l[j * m + i])), mkdd(0.5));
  h[i * m + j] = avg.hi; l[i * m + j] = avg.lo;
  h[j * m + i] = avg.hi; l[j * m + i] = avg.lo;
}
// out[i][nk] = in[i][order[nk]] (m x m)
__global__ void k_gather_cols(const double* __restrict__ ih, const double* __restrict__ il, long m,
                              const long* __restrict__ order, double* __restrict__ oh, double* __restrict__ ol) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= m * m) return;
  const long i = idx / m, k = idx % m;
  oh[idx] = ih[i * m + order[k]];
  ol[idx] = il[i * m + order[k]];
}

}  // namespace

// Symmetric eigendecomposition of the n x n dd matrix (device planes, row-major).
//   stop_frobenius: stop when sqrt(sum_{a<b} a_ab^2) < 1e-30, tested before each sweep,
//                   with at most max_sweeps - 1 sweeps run (linalg.rs's loop); HP_ENOCONV
//                   otherwise;
//   else:           sweep until max |a_ab| < 1e-60 (ortho.rs's test), at most max_sweeps.
// evals: n dd (diagonal order of the original indices); evecs: n x n dd, row o =
// eigenvector o when vec_rows, else column o.
// smallest n that takes the block form (hp_eigen_set_block_min)
constexpr long BJ_DEFAULT_MIN = 768;
std::atomic<long> g_bj_min{BJ_DEFAULT_MIN};

// The block form of jacobi_eigen_dev (same stop rules, outputs and sweep count meaning).
int block_jacobi_dev(const double* a_h, const double* a_l, long n, bool stop_frobenius, long max_sweeps,
                     double* e_h, double* e_l, double* x_h, double* x_l, bool vec_rows, long* sweeps_run,
                     hipStream_t s) {
  const long Nb = (n + BJ_S - 1) / BJ_S * BJ_S, nb = Nb / BJ_B, mp = nb / 2, steps = nb - 1;
  if (Nb > 65536) {
    set_error("jacobi_eigen: n = %ld too large (at most 65536)", n);
    return HP_EINVAL;
  }
  const size_t nn = (size_t)Nb * Nb;
  DevBuf A, V, Q, red, posb, pairb;
  int rc;
  if ((rc = A.alloc(2 * nn * sizeof(double))) || (rc = V.alloc(2 * nn * sizeof(double))) ||
      (rc = Q.alloc(4 * (size_t)mp * BJ_S * BJ_S * sizeof(double))) || (rc = red.alloc(16)) ||
      (rc = posb.alloc(n * sizeof(long))) || (rc = pairb.alloc((size_t)steps * mp * sizeof(int2))))
    return rc;
  double *ah = A.d(), *al = ah + nn, *vh = V.d(), *vl = vh + nn;
  // two Q buffers (steps alternate), so the V update of a step can run on a second stream
  // beside the next step's subproblems
  const size_t qsz = (size_t)mp * BJ_S * BJ_S;
  double* qbh[2] = {Q.d(), Q.d() + 2 * qsz};
  double* qbl[2] = {Q.d() + qsz, Q.d() + 3 * qsz};
  std::vector<int2> pr((size_t)steps * mp);
  for (long t = 0; t < steps; ++t)
    for (long k = 0; k < mp; ++k) {
      int p, q;
      rr_match((int)t, (int)k, (int)nb, p, q);
      pr[t * mp + k] = make_int2(p, q);
    }
  HP_HIP(hipMemcpyAsync(pairb.p, pr.data(), pr.size() * sizeof(int2), hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_jac_init, dim3(blocks((long)nn)), dim3(256), 0, s, a_h, a_l, n, Nb, ah, al, vh, vl);
  HP_LAUNCH_CHECK();
  auto off = [&](double* fro, double* mxv) -> int {
    HP_HIP(hipMemsetAsync(red.p, 0, 16, s));
    unsigned long long* mx = reinterpret_cast<unsigned long long*>(red.d() + 1);
    hipLaunchKernelGGL(k_jac_off, dim3(blocks(Nb), (unsigned)Nb), dim3(256), 0, s, ah, al, Nb, red.d(), mx);
    HP_LAUNCH_CHECK();
    double host[2];
    HP_HIP(hipMemcpyAsync(host, red.p, 16, hipMemcpyDeviceToHost, s));
    HP_HIP(hipStreamSynchronize(s));
    *fro = std::sqrt(host[0]);
    std::memcpy(mxv, &host[1], 8);
    return HP_OK;
  };
  // one outer sweep as a graph, replayed (everything is in place)
  struct Graph {
    hipStream_t cs = nullptr, vs = nullptr;
    std::vector<hipEvent_t> ev;
    hipGraphExec_t exec = nullptr;
    ~Graph() {
      if (exec) (void)hipGraphExecDestroy(exec);
      if (vs) (void)hipStreamDestroy(vs);
      if (cs) (void)hipStreamDestroy(cs);
      for (hipEvent_t e : ev) (void)hipEventDestroy(e);
    }
  } gr;
  HP_HIP(hipStreamCreateWithFlags(&gr.cs, hipStreamNonBlocking));
  HP_HIP(hipStreamCreateWithFlags(&gr.vs, hipStreamNonBlocking));
  gr.ev.assign(2 * steps + 2, nullptr);
  for (hipEvent_t& e : gr.ev) HP_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  HP_HIP(hipStreamBeginCapture(gr.cs, hipStreamCaptureModeThreadLocal));
  HP_HIP(hipEventRecord(gr.ev[2 * steps], gr.cs));
  HP_HIP(hipStreamWaitEvent(gr.vs, gr.ev[2 * steps], 0));
  const int2* pd = static_cast<const int2*>(pairb.p);
  // step t: Q_t into buffer t & 1; the V update of step t (second stream) reads it, so
  // the subproblems of step t + 2, which rewrite that buffer, wait for it
  for (long t = 0; t < steps; ++t) {
    double *qh = qbh[t & 1], *ql = qbl[t & 1];
    if (t >= 2) HP_HIP(hipStreamWaitEvent(gr.cs, gr.ev[2 * (t - 2) + 1], 0));
    hipLaunchKernelGGL(k_bj_sub, dim3((unsigned)mp), dim3(256), 0, gr.cs, static_cast<const double*>(ah),
                       static_cast<const double*>(al), Nb, pd + t * mp, qh, ql);
    HP_HIP(hipEventRecord(gr.ev[2 * t], gr.cs));
    hipLaunchKernelGGL(k_bj_apply, dim3((unsigned)(mp * (mp + 1) / 2)), dim3(256), 0, gr.cs, ah, al, Nb, pd + t * mp,
                       static_cast<const double*>(qh), static_cast<const double*>(ql));
    HP_HIP(hipStreamWaitEvent(gr.vs, gr.ev[2 * t], 0));
    hipLaunchKernelGGL(k_bj_vapply, dim3((unsigned)(Nb / BJ_S), (unsigned)mp), dim3(256), 0, gr.vs, vh, vl, Nb,
                       pd + t * mp, static_cast<const double*>(qh), static_cast<const double*>(ql));
    HP_HIP(hipEventRecord(gr.ev[2 * t + 1], gr.vs));
  }
  HP_HIP(hipEventRecord(gr.ev[2 * steps + 1], gr.vs));
  HP_HIP(hipStreamWaitEvent(gr.cs, gr.ev[2 * steps + 1], 0));
  {
    hipGraph_t g = nullptr;
    HP_HIP(hipStreamEndCapture(gr.cs, &g));
    hipError_t e = hipGraphInstantiate(&gr.exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    HP_HIP(e);
  }
  bool converged = false;
  long run = 0;
  for (long it = 0; it < max_sweeps; ++it) {
    double fro = 0.0, mxv = 0.0;
    if ((rc = off(&fro, &mxv))) return rc;
    if (stop_frobenius ? fro < 1e-30 : mxv < 1e-60) {
      converged = true;
      break;
    }
    HP_HIP(hipGraphLaunch(gr.exec, s));
    ++run;
  }
  if (!converged && !stop_frobenius) {
    double fro = 0.0, mxv = 0.0;
    if ((rc = off(&fro, &mxv))) return rc;
    converged = mxv < 1e-60;
  }
  if (sweeps_run) *sweeps_run = run;
  if (!converged && stop_frobenius) {
    set_error("Jacobi eigenvalue solver did not converge within max iterations");
    return HP_ENOCONV;
  }
  std::vector<long> pos(n);
  std::iota(pos.begin(), pos.end(), 0L);  // natural order: position = index
  HP_HIP(hipMemcpyAsync(posb.p, pos.data(), n * sizeof(long), hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_jac_extract, dim3(blocks(n * n)), dim3(256), 0, s, ah, al, vh, vl, n, Nb,
                     static_cast<const long*>(posb.p), e_h, e_l, x_h, x_l, vec_rows ? 1 : 0);
  HP_LAUNCH_CHECK();
  HP_HIP(hipStreamSynchronize(s));
  return HP_OK;
}

int jacobi_eigen_dev(const double* a_h, const double* a_l, long n, bool stop_frobenius, long max_sweeps,
                     double* e_h, double* e_l, double* x_h, double* x_l, bool vec_rows, long* sweeps_run,
                     hipStream_t s) {
  if (sweeps_run) *sweeps_run = 0;
  if (n <= 0) return HP_OK;
  if (n >= g_bj_min.load())
    return block_jacobi_dev(a_h, a_l, n, stop_frobenius, max_sweeps, e_h, e_l, x_h, x_l, vec_rows, sweeps_run, s);
  const long N2 = n + (n & 1), m = N2 / 2;
  if (N2 > 65534) {
    set_error("jacobi_eigen: n = %ld too large (at most 65534)", n);
    return HP_EINVAL;
  }
  const size_t nn = (size_t)N2 * N2;
  DevBuf A0, A1, V0, V1, red, posb;
  int rc;
  if ((rc = A0.alloc(2 * nn * sizeof(double))) || (rc = A1.alloc(2 * nn * sizeof(double))) ||
      (rc = V0.alloc(2 * nn * sizeof(double))) || (rc = V1.alloc(2 * nn * sizeof(double))) ||
      (rc = red.alloc(16)) || (rc = posb.alloc(N2 * sizeof(long))))
    return rc;
  double *ah = A0.d(), *al = ah + nn, *bh = A1.d(), *bl = bh + nn;
  double *vh = V0.d(), *vl = vh + nn, *wh = V1.d(), *wl = wh + nn;
  hipLaunchKernelGGL(k_jac_init, dim3(blocks((long)nn)), dim3(256), 0, s, a_h, a_l, n, N2, ah, al, vh, vl);
  HP_LAUNCH_CHECK();
  std::vector<long> layout(N2);  // position -> original index
  std::iota(layout.begin(), layout.end(), 0L);
  std::vector<long> next(N2);
  auto off = [&](double* fro, double* mxv) -> int {
    HP_HIP(hipMemsetAsync(red.p, 0, 16, s));
    double* sum = red.d();
    unsigned long long* mx = reinterpret_cast<unsigned long long*>(red.d() + 1);
    hipLaunchKernelGGL(k_jac_off, dim3(blocks(N2), (unsigned)N2), dim3(256), 0, s, ah, al, N2, sum, mx);
    HP_LAUNCH_CHECK();
    double host[2];
    HP_HIP(hipMemcpyAsync(host, red.p, 16, hipMemcpyDeviceToHost, s));
    HP_HIP(hipStreamSynchronize(s));
    *fro = std::sqrt(host[0]);
    unsigned long long bits;
    std::memcpy(&bits, &host[1], 8);
    double mv;
    std::memcpy(&mv, &bits, 8);
    *mxv = mv;
    return HP_OK;
  };
  // A sweep is N2 - 1 launches; each is captured once into a graph per buffer parity
  // (N2 - 1 is odd, so consecutive sweeps start from opposite buffers) and replayed.
  struct Graphs {
    hipStream_t cs = nullptr;
    hipGraphExec_t exec[2] = {nullptr, nullptr};
    ~Graphs() {
      for (auto& e : exec)
        if (e) (void)hipGraphExecDestroy(e);
      if (cs) (void)hipStreamDestroy(cs);
    }
  } gr;
  // V batching (N2 <= 2048): the steps update A only and store their rotations, and every
  // VB steps (and at the end of a sweep) k_jac_vbatch applies them to V in place: V is
  // read and written once per VB steps instead of once per step.  (A forked stream for
  // the batches, joined by graph edges, measured slower.)  Larger matrices update V inside
  // each step.
  constexpr int VB = 128;
  const bool vbatch = N2 <= 2048;
  const int vr = N2 <= 512 ? 2 : 1;  // rows of V per block: LDS of 32 vr N2 bytes <= 64 KiB
  DevBuf rotb;
  if (vbatch && (rc = rotb.alloc((size_t)VB * 4 * m * sizeof(double)))) return rc;
  auto capture = [&](int parity) -> int {
    if (!gr.cs) HP_HIP(hipStreamCreateWithFlags(&gr.cs, hipStreamNonBlocking));
    double *xh = ah, *xl = al, *yh = bh, *yl = bl, *uh = vh, *ul = vl, *zh = wh, *zl = wl;
    HP_HIP(hipStreamBeginCapture(gr.cs, hipStreamCaptureModeThreadLocal));
    const unsigned gz = vbatch ? 1 : 2;
    for (long step = 0; step < N2 - 1; ++step) {
      double* rp = vbatch ? rotb.d() + (size_t)(step % VB) * 4 * m : nullptr;
      // small matrices: one I per wavefront for enough blocks, loads ahead of the rotations;
      // larger ones: four, which forms fewer copies of each rotation
      if (m <= 256)
        hipLaunchKernelGGL((k_jac_step<1, 4, true>),
                           dim3(blocks(m, 64), vbatch ? blocks(m, 4) : std::max(blocks(m, 4), blocks(N2, 16)), gz),
                           dim3(256), 0, gr.cs, xh, xl, yh, yl, uh, ul, zh, zl, N2, rp);
      else if (m <= 512)
        hipLaunchKernelGGL((k_jac_step<1, 4, false>),
                           dim3(blocks(m, 64), vbatch ? blocks(m, 4) : std::max(blocks(m, 4), blocks(N2, 16)), gz),
                           dim3(256), 0, gr.cs, xh, xl, yh, yl, uh, ul, zh, zl, N2, rp);
      else
        hipLaunchKernelGGL((k_jac_step<4, 16, false>),
                           dim3(blocks(m, 64), vbatch ? blocks(m, 16) : std::max(blocks(m, 16), blocks(N2, 64)), gz),
                           dim3(256), 0, gr.cs, xh, xl, yh, yl, uh, ul, zh, zl, N2, rp);
      std::swap(xh, yh);
      std::swap(xl, yl);
      if (!vbatch) {
        std::swap(uh, zh);
        std::swap(ul, zl);
      } else if ((step + 1) % VB == 0 || step == N2 - 2) {
        const int cnt = (int)(step % VB) + 1;
        const size_t lds = (size_t)32 * vr * N2;
        if (vr == 2)
          hipLaunchKernelGGL(k_jac_vbatch<2>, dim3((unsigned)((N2 + 1) / 2)), dim3(256), lds, gr.cs, vh, vl,
                             rotb.d(), cnt, N2);
        else
          hipLaunchKernelGGL(k_jac_vbatch<1>, dim3((unsigned)N2), dim3(256), lds, gr.cs, vh, vl, rotb.d(), cnt, N2);
        HP_LAUNCH_CHECK();
      }
    }
    hipGraph_t g = nullptr;
    HP_HIP(hipStreamEndCapture(gr.cs, &g));
    hipError_t e = hipGraphInstantiate(&gr.exec[parity], g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    HP_HIP(e);
    return HP_OK;
  };
  int parity = 0;
  auto sweep = [&]() -> int {
    if (!gr.exec[parity]) {
      if (int rc2 = capture(parity)) return rc2;
    }
    HP_HIP(hipGraphLaunch(gr.exec[parity], s));
    std::swap(ah, bh);  // N2 - 1 (odd) swaps
    std::swap(al, bl);
    if (!vbatch) {  // batched V stays in place
      std::swap(vh, wh);
      std::swap(vl, wl);
    }
    parity ^= 1;
    for (long step = 0; step < N2 - 1; ++step) {
      for (long p = 0; p < N2; ++p) next[rr_sigma(p, m)] = layout[p];
      layout.swap(next);
    }
    return HP_OK;
  };
  bool converged = false;
  long run = 0;
  for (long it = 0; it < max_sweeps; ++it) {
    double fro = 0.0, mxv = 0.0;
    if ((rc = off(&fro, &mxv))) return rc;
    if (stop_frobenius ? fro < 1e-30 : mxv < 1e-60) {
      converged = true;
      break;
    }
    if ((rc = sweep())) return rc;
    ++run;
  }
  if (!converged && !stop_frobenius) {
    double fro = 0.0, mxv = 0.0;
    if ((rc = off(&fro, &mxv))) return rc;
    converged = mxv < 1e-60;
  }
  if (sweeps_run) *sweeps_run = run;
  if (!converged && stop_frobenius) {
    set_error("Jacobi eigenvalue solver did not converge within max iterations");
    return HP_ENOCONV;
  }
  std::vector<long> pos(N2);
  for (long p = 0; p < N2; ++p) pos[layout[p]] = p;
  HP_HIP(hipMemcpyAsync(posb.p, pos.data(), N2 * sizeof(long), hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_jac_extract, dim3(blocks(n * n)), dim3(256), 0, s, ah, al, vh, vl, n, N2,
                     static_cast<const long*>(posb.p), e_h, e_l, x_h, x_l, vec_rows ? 1 : 0);
  HP_LAUNCH_CHECK();
  HP_HIP(hipStreamSynchronize(s));
  return HP_OK;
}


// The solver crate's eigensolver stops below max |a_pq| = 1e-60 or after 100 n^2 rotations
// (ortho.rs:73-85); the parallel sweeps stop at the same test, capped at 200 sweeps (each
// sweep is n(n-1)/2 rotations).
constexpr long ORTHO_MAX_SWEEPS = 200;

// canonical_orthogonalize (ortho.rs:158-193) of the n x n dd matrix (device planes):
// X (n x n_orth) is allocated in xbuf (planes hi | lo).  eval_host (optional) receives the
// n eigenvalues (dd, diagonal order).
int canonical_ortho_dev(const double* s_h, const double* s_l, long n, double threshold, DevBuf& xbuf, long* m_out,
                        long* sweeps, hipStream_t s) {
  DevBuf E, V, selb, lamb;
  int rc;
  if ((rc = E.alloc(2 * n * sizeof(double))) || (rc = V.alloc(2 * (size_t)n * n * sizeof(double)))) return rc;
  const size_t nn = (size_t)n * n;
  if ((rc = jacobi_eigen_dev(s_h, s_l, n, false, ORTHO_MAX_SWEEPS, E.d(), E.d() + n, V.d(), V.d() + nn, false, sweeps,
                             s)))
    return rc;
  std::vector<double> ev(2 * n);
  HP_HIP(hipMemcpyAsync(ev.data(), E.p, 2 * n * sizeof(double), hipMemcpyDeviceToHost, s));
  HP_HIP(hipStreamSynchronize(s));
  std::vector<long> sel;
  for (long i = 0; i < n; ++i)
    if (ev[i] + ev[n + i] > threshold) sel.push_back(i);
  std::stable_sort(sel.begin(), sel.end(), [&](long a, long b) { return ev[a] + ev[n + a] < ev[b] + ev[n + b]; });
  const long m = (long)sel.size();
  *m_out = m;
  if ((rc = xbuf.alloc(2 * (size_t)n * std::max<long>(m, 1) * sizeof(double)))) return rc;
  if (m == 0) return HP_OK;
  std::vector<double> lam(2 * m);
  for (long k = 0; k < m; ++k) {
    lam[2 * k] = ev[sel[k]];
    lam[2 * k + 1] = ev[n + sel[k]];
  }
  if ((rc = selb.alloc(m * sizeof(long))) || (rc = lamb.alloc(2 * m * sizeof(double)))) return rc;
  HP_HIP(hipMemcpyAsync(selb.p, sel.data(), m * sizeof(long), hipMemcpyHostToDevice, s));
  HP_HIP(hipMemcpyAsync(lamb.p, lam.data(), 2 * m * sizeof(double), hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_ortho_x, dim3(blocks(n * m)), dim3(256), 0, s, V.d(), V.d() + nn, n,
                     static_cast<const long*>(selb.p), lamb.d(), m, xbuf.d(), xbuf.d() + (size_t)n * m);
  HP_LAUNCH_CHECK();
  HP_HIP(hipStreamSynchronize(s));
  return HP_OK;
}

// solve_gen_eigen (ortho.rs:201-248) on device planes: eigenvalues ascending (m dd into
// ev_h / ev_l), C = X C' (n x m into c_h / c_l; capacity n x n)
int solve_gen_eigen_dev(const double* f_h, const double* f_l, const double* s_h, const double* s_l, long n,
                        double threshold, double* ev_h, double* ev_l, double* c_h, double* c_l, long* m_out,
                        long* sweeps, hipStream_t s) {
  DevBuf X;
  long m = 0;
  long sw[2] = {0, 0};
  int rc;
  if ((rc = canonical_ortho_dev(s_h, s_l, n, threshold, X, &m, &sw[0], s))) return rc;
  *m_out = m;
  if (m == 0) return HP_OK;
  const size_t nm = (size_t)n * m, mm = (size_t)m * m;
  DevBuf T, Fp, E, Vp, Cp, ordb;
  if ((rc = T.alloc(2 * nm * sizeof(double))) || (rc = Fp.alloc(2 * mm * sizeof(double))) ||
      (rc = E.alloc(2 * m * sizeof(double))) || (rc = Vp.alloc(2 * mm * sizeof(double))) ||
      (rc = Cp.alloc(2 * mm * sizeof(double))) || (rc = ordb.alloc(m * sizeof(long))))
    return rc;
  const double *xh = X.d(), *xl = X.d() + nm;
  // tmp = F X (n x m), F' = X^T tmp (m x m): sums in k order as the reference's gemm_rect /
  // gemm_atb (whose skipped zero terms would add exact zeros)
  if ((rc = launch_gemm_dd_plain(0, f_h, f_l, xh, xl, T.d(), T.d() + nm, (size_t)n, (size_t)m, (size_t)n, s)))
    return rc;
  if ((rc = launch_gemm_dd_plain(1, xh, xl, T.d(), T.d() + nm, Fp.d(), Fp.d() + mm, (size_t)m, (size_t)m, (size_t)n,
                                 s)))
    return rc;
  hipLaunchKernelGGL(k_symmetrize, dim3(blocks((long)mm)), dim3(256), 0, s, Fp.d(), Fp.d() + mm, m);
  HP_LAUNCH_CHECK();
  if ((rc = jacobi_eigen_dev(Fp.d(), Fp.d() + mm, m, false, ORTHO_MAX_SWEEPS, E.d(), E.d() + m, Vp.d(), Vp.d() + mm,
                             false, &sw[1], s)))
    return rc;
  if (sweeps) *sweeps = sw[0] + sw[1];
  std::vector<double> ev(2 * m);
  HP_HIP(hipMemcpyAsync(ev.data(), E.p, 2 * m * sizeof(double), hipMemcpyDeviceToHost, s));
  HP_HIP(hipStreamSynchronize(s));
  std::vector<long> order(m);
  std::iota(order.begin(), order.end(), 0L);
  std::stable_sort(order.begin(), order.end(), [&](long a, long b) { return ev[a] + ev[m + a] < ev[b] + ev[m + b]; });
  std::vector<double> sorted(2 * m);
  for (long k = 0; k < m; ++k) {
    sorted[k] = ev[order[k]];
    sorted[m + k] = ev[m + order[k]];
  }
  HP_HIP(hipMemcpyAsync(ev_h, sorted.data(), m * sizeof(double), hipMemcpyHostToDevice, s));
  HP_HIP(hipMemcpyAsync(ev_l, sorted.data() + m, m * sizeof(double), hipMemcpyHostToDevice, s));
  HP_HIP(hipMemcpyAsync(ordb.p, order.data(), m * sizeof(long), hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_gather_cols, dim3(blocks((long)mm)), dim3(256), 0, s, Vp.d(), Vp.d() + mm, m,
                     static_cast<const long*>(ordb.p), Cp.d(), Cp.d() + mm);
  HP_LAUNCH_CHECK();
  if ((rc = launch_gemm_dd_plain(0, xh, xl, Cp.d(), Cp.d() + mm, c_h, c_l, (size_t)n, (size_t)m, (size_t)m, s)))
    return rc;
  HP_HIP(hipStreamSynchronize(s));
  return HP_OK;
}

}  // namespace hp

using namespace hp;

extern "C" {

int hp_eigen_set_block_min(size_t n) {
  g_bj_min.store(n == 0 ? BJ_DEFAULT_MIN : (long)std::min<size_t>(n, (size_t)LONG_MAX));
  return HP_OK;
}

int hp_jacobi_eigen(const double* a, size_t n, size_t max_iter, double* evals, double* evecs) {
  if (n == 0) return HP_OK;
  if (!a || !evals || !evecs) {
    set_error("null pointer argument");
    return HP_EINVAL;
  }
  hipStream_t s = hipStreamPerThread;
  const size_t nn = n * n;
  DevBuf A, E, V;
  int rc;
  if ((rc = A.alloc(2 * nn * sizeof(double))) || (rc = E.alloc(2 * n * sizeof(double))) ||
      (rc = V.alloc(2 * nn * sizeof(double))))
    return rc;
  if ((rc = upload_planes(a, nn, 2, A.d(), s))) return rc;
  if ((rc = jacobi_eigen_dev(A.d(), A.d() + nn, (long)n, true, (long)max_iter, E.d(), E.d() + n, V.d(), V.d() + nn,
                             true, nullptr, s)))
    return rc;
  if ((rc = download_planes(E.d(), n, 2, evals, s))) return rc;
  return download_planes(V.d(), nn, 2, evecs, s);
}

int hp_dev_jacobi_eigen(const double* a_hi, const double* a_lo, size_t n, size_t max_sweeps, int stop_rule,
                        double* evals_hi, double* evals_lo, double* vecs_hi, double* vecs_lo, int vec_rows,
                        size_t* sweeps, void* stream) {
  long run = 0;
  int rc = jacobi_eigen_dev(a_hi, a_lo, (long)n, stop_rule == 0, (long)max_sweeps, evals_hi, evals_lo, vecs_hi, vecs_lo,
                            vec_rows != 0, &run, (hipStream_t)stream);
  if (sweeps) *sweeps = (size_t)run;
  return rc;
}

int hp_canonical_orthogonalize(const double* s_dd, size_t n, double threshold, double* x_dd, size_t* n_orth) {
  if (!n_orth) {
    set_error("null pointer argument");
    return HP_EINVAL;
  }
  *n_orth = 0;
  if (n == 0) return HP_OK;
  if (!s_dd || !x_dd) {
    set_error("null pointer argument");
    return HP_EINVAL;
  }
  hipStream_t s = hipStreamPerThread;
  const size_t nn = n * n;
  DevBuf S, X;
  int rc;
  if ((rc = S.alloc(2 * nn * sizeof(double)))) return rc;
  if ((rc = upload_planes(s_dd, nn, 2, S.d(), s))) return rc;
  long m = 0;
  if ((rc = canonical_ortho_dev(S.d(), S.d() + nn, (long)n, threshold, X, &m, nullptr, s))) return rc;
  *n_orth = (size_t)m;
  if (m == 0) return HP_OK;
  return download_planes(X.d(), n * (size_t)m, 2, x_dd, s);
}

int hp_solve_gen_eigen(const double* f_dd, const double* s_dd, size_t n, double threshold, double* evals_dd,
                       double* c_dd, size_t* n_basis) {
  if (!n_basis) {
    set_error("null pointer argument");
    return HP_EINVAL;
  }
  *n_basis = 0;
  if (n == 0) return HP_OK;
  if (!f_dd || !s_dd || !evals_dd || !c_dd) {
    set_error("null pointer argument");
    return HP_EINVAL;
  }
  hipStream_t s = hipStreamPerThread;
  const size_t nn = n * n;
  DevBuf F, S, E, C;
  int rc;
  if ((rc = F.alloc(2 * nn * sizeof(double))) || (rc = S.alloc(2 * nn * sizeof(double))) ||
      (rc = E.alloc(2 * n * sizeof(double))) || (rc = C.alloc(2 * nn * sizeof(double))))
    return rc;
  if ((rc = upload_planes(f_dd, nn, 2, F.d(), s)) || (rc = upload_planes(s_dd, nn, 2, S.d(), s))) return rc;
  long m = 0;
  if ((rc = solve_gen_eigen_dev(F.d(), F.d() + nn, S.d(), S.d() + nn, (long)n, threshold, E.d(), E.d() + n, C.d(),
                                C.d() + nn, &m, nullptr, s)))
    return rc;
  *n_basis = (size_t)m;
  if (m == 0) return HP_OK;
  // planes of m and n x m values packed at the front of each buffer
  std::vector<double> eh(2 * m);
  HP_HIP(hipMemcpyAsync(eh.data(), E.d(), m * sizeof(double), hipMemcpyDeviceToHost, s));
  HP_HIP(hipMemcpyAsync(eh.data() + m, E.d() + n, m * sizeof(double), hipMemcpyDeviceToHost, s));
  std::vector<double> ch(2 * n * (size_t)m);
  HP_HIP(hipMemcpyAsync(ch.data(), C.d(), n * (size_t)m * sizeof(double), hipMemcpyDeviceToHost, s));
  HP_HIP(hipMemcpyAsync(ch.data() + n * (size_t)m, C.d() + nn, n * (size_t)m * sizeof(double), hipMemcpyDeviceToHost,
                        s));
  HP_HIP(hipStreamSynchronize(s));
  for (long k = 0; k < m; ++k) {
    evals_dd[2 * k] = eh[k];
    evals_dd[2 * k + 1] = eh[m + k];
  }
  const size_t nm = n * (size_t)m;
  for (size_t i = 0; i < nm; ++i) {
    c_dd[2 * i] = ch[i];
    c_dd[2 * i + 1] = ch[nm + i];
  }
  return HP_OK;
}

int hp_dev_solve_gen_eigen(const double* f_hi, const double* f_lo, const double* s_hi, const double* s_lo, size_t n,
                           double threshold, double* evals_hi, double* evals_lo, double* c_hi, double* c_lo,
                           size_t* n_basis, size_t* sweeps, void* stream) {
  long m = 0, sw = 0;
  int rc = solve_gen_eigen_dev(f_hi, f_lo, s_hi, s_lo, (long)n, threshold, evals_hi, evals_lo, c_hi, c_lo, &m, &sw,
                               (hipStream_t)stream);
  if (n_basis) *n_basis = (size_t)m;
  if (sweeps) *sweeps = (size_t)sw;
  return rc;
}

}  // extern "C"
