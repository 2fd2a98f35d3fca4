// k_batch.hip -- many dd SPD factor + solve instances of uneven size in one launch.
//
// A persistent kernel: each workgroup repeatedly takes the next instance from an atomic
// queue ordered by n^3, largest first (longest-processing-time order), copies S into its
// output L, factors it with cholesky_blocked(nb = 64) (crates/quad/src/linalg.rs:238-293,
// same per-element operation order, so each factor is bit-identical to the unbatched
// one), then solves L L^T X = B for the instance's right-hand sides: forward substitution
// in the reference order (linalg.rs:99-108) and a right-looking blocked backward solve.
// The whole batch (queue reset + persistent kernel) is captured once in a hipGraph and
// replayed; the queue is reset by a kernel, not a memset node.
//
// Layout: matrices packed one after another (instance i at sum_{j<i} n_j^2), hi plane then
// lo plane of the total; right-hand sides packed the same way, instance i's r-th vector at
// rhs_off_i + r*n_i.  Per-instance status: 0 ok, p+1 = Cholesky breakdown at pivot p.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstddef>
#include <map>
#include <mutex>
#include <numeric>
#include <vector>

#include "hp_chol.hpp"
#include "hp_internal.hpp"

namespace hp {
namespace {

constexpr int NB = 64;          // block size of the factorisation (cholesky_blocked default)
constexpr int NT = 512;         // threads per workgroup (8 wavefronts)
constexpr int MAXRHS = 8;

struct BatchDesc {
  int n;
  int pad;
  long long mat_off;
  long long rhs_off;
};

__device__ __forceinline__ double rlane(double v, int lane) {
  long long b = __builtin_bit_cast(long long, v);
  int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), lane);
  int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
  return __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned int)lo);
}
__device__ __forceinline__ dd rlane(dd v, int lane) { return mkdd(rlane(v.hi, lane), rlane(v.lo, lane)); }

// LDS, reused between phases
struct DiagLds {
  double Lh[NB][NB + 1], Ll[NB][NB + 1];
  double Sh[NB][NB + 1], Sl[NB][NB + 1];
};
struct TileLds {
  double As[2][2][16][64], Bs[2][2][16][64];  // [half][plane][k][row]
};
union BatchLds {
  DiagLds d;
  TileLds t;
};

__device__ __forceinline__ dd ldm(const double* h, const double* l, long o) { return mkdd(h[o], l[o]); }
__device__ __forceinline__ void stm(double* h, double* l, long o, dd v) { h[o] = v.hi; l[o] = v.lo; }

// ---- the phases of cholesky_blocked(nb = 64) on one instance, shared by both schedules.
// Src is where the not-yet-factored entries are read (S on the panel schedule's first step,
// else the factor L = A itself); results go to A.

// Diagonal block [jb0, jb0 + jb): the unblocked algorithm (linalg.rs:250-255, i.e. 42-96 on
// the block) in the right-looking form that keeps each element's k order: the partial sum
// S_ic starts at zero and accumulates l_ij l_cj (one dd multiply and one dd add) once column
// j is final; column c is finished as l_cc = sqrt(a_cc - S_cc), l_ic = (a_ic - S_ic) *
// recip(l_cc).  One barrier per column: wavefront 0 (lane = row) finishes column j+1 --
// it adds column j's product to S(., j+1) itself, takes the pivot on lane j+1 and
// broadcasts its reciprocal with v_readlane -- while the seven other wavefronts add column
// j's products to the columns beyond j+1.  Their partial sums live in registers: element e
// of the packed triangle (e = q(q+1)/2 + p, 0 <= p <= q, is (jb-1-p, jb-1-q)) belongs to
// thread 64 + e mod 448 (the columns still active form a prefix of the packing, so the work
// stays spread evenly), and an element is handed to wavefront 0 through a two-column LDS
// buffer when its column is next but one.  Returns false (every thread) on a breakdown,
// with status[inst] = pivot + 1.  On success the block's lower triangle is written to A and
// stays in d.Lh / d.Ll (d.Sh / d.Sl are not used).
constexpr int DG_OWN = 5;  // partial sums per trailing thread: 5 * 448 >= 62 * 63 / 2
__device__ __forceinline__ bool bt_diag(const double* Srch, const double* Srcl, double* Ah, double* Al, long n, long jb0, int jb,
                        DiagLds& d, double* /*s_inv*/, int* s_fail, int* status, int inst) {
  __shared__ double Hh[2][NB], Hl[2][NB];  // partial sums of column c for wavefront 0, by the parity of c
  const int tid = threadIdx.x, lane = tid & 63;
  for (int idx = tid; idx < jb * jb; idx += NT) {
    int i = idx / jb, c = idx % jb;
    if (c <= i) {
      long o = (jb0 + i) * n + jb0 + c;
      d.Lh[i][c] = Srch[o];
      d.Ll[i][c] = Srcl[o];
    }
  }
  if (tid == 0) *s_fail = 0;
  const int cnt_all = (jb - 2) * (jb - 1) / 2;
  dd sv[DG_OWN];
  int ei[DG_OWN], ec[DG_OWN];
  if (tid >= 64) {
#pragma unroll
    for (int u = 0; u < DG_OWN; ++u) {
      const int e = tid - 64 + u * (NT - 64);
      int q = (int)((__builtin_sqrtf(8.0f * (float)e + 1.0f) - 1.0f) * 0.5f);
      while (q * (q + 1) / 2 > e) --q;
      while ((q + 1) * (q + 2) / 2 <= e) ++q;
      ei[u] = jb - 1 - (e - q * (q + 1) / 2);
      ec[u] = jb - 1 - q;
      if (e >= cnt_all) {
        ei[u] = 0;
        ec[u] = -1;  // no element
      }
      sv[u] = mkdd(0.0, 0.0);
    }
  }
  __syncthreads();
  dd lcol = mkdd(0.0, 0.0);  // wavefront 0: l(lane, j) of the last finished column
  auto finish_column = [&](int c, dd sc) {
    dd inv = mkdd(0.0, 0.0);
    int bad = 0;
    if (lane == c) {
      const dd diag = dd_sub(mkdd(d.Lh[c][c], d.Ll[c][c]), sc);
      if (diag.hi <= 0.0) {
        *s_fail = 1;
        status[inst] = (int)(jb0 + c) + 1;
        bad = 1;
      } else {
        const dd ljj = dd_sqrt(diag);
        d.Lh[c][c] = ljj.hi;
        d.Ll[c][c] = ljj.lo;
        inv = dd_recip(ljj);
      }
    }
    if (__builtin_amdgcn_readlane(bad, c)) return;
    inv = rlane(inv, c);
    if (lane > c && lane < jb) {
      lcol = dd_mul(dd_sub(mkdd(d.Lh[lane][c], d.Ll[lane][c]), sc), inv);
      d.Lh[lane][c] = lcol.hi;
      d.Ll[lane][c] = lcol.lo;
    }
  };
  if (tid < 64) finish_column(0, mkdd(0.0, 0.0));
  __syncthreads();
  for (int j = 0; j + 1 < jb; ++j) {
    if (*s_fail) return false;
    if (tid < 64) {
      const int c = j + 1;
      const dd ljc = rlane(lcol, c);  // l(j+1, j)
      dd sc = mkdd(0.0, 0.0);
      if (lane >= c && lane < jb) {
        const dd sin = c == 1 ? mkdd(0.0, 0.0) : mkdd(Hh[c & 1][lane], Hl[c & 1][lane]);
        sc = dd_add(sin, dd_mul(lcol, ljc));
      }
      finish_column(c, sc);
    } else {
      dd li[DG_OWN], lc[DG_OWN];
#pragma unroll
      for (int u = 0; u < DG_OWN; ++u) {  // every operand read first
        li[u] = mkdd(d.Lh[ei[u]][j], d.Ll[ei[u]][j]);
        const int cc = max(ec[u], 0);
        lc[u] = mkdd(d.Lh[cc][j], d.Ll[cc][j]);
      }
#pragma unroll
      for (int u = 0; u < DG_OWN; ++u) {
        if (ec[u] >= j + 2) {  // columns beyond j+1 take column j
          sv[u] = dd_add(sv[u], dd_mul(li[u], lc[u]));
          if (ec[u] == j + 2) {  // complete but for column j+1, which wavefront 0 adds
            Hh[ec[u] & 1][ei[u]] = sv[u].hi;
            Hl[ec[u] & 1][ei[u]] = sv[u].lo;
          }
        }
      }
    }
    __syncthreads();
  }
  if (*s_fail) return false;
  for (int idx = tid; idx < jb * jb; idx += NT) {
    int i = idx / jb, c = idx % jb;
    if (c <= i) {
      long o = (jb0 + i) * n + jb0 + c;
      Ah[o] = d.Lh[i][c];
      Al[o] = d.Ll[i][c];
    }
  }
  return true;
}

// trsm_right_lt (linalg.rs:224-235) of rows row0, row0 + step, ... < row1 against the block's
// L in d.Lh / d.Ll: one wavefront per row, lane = column, each product subtracted in k order.
// A wavefront runs four rows' chains interleaved (independent, so their latencies overlap).
__device__ __forceinline__ void bt_trsm_rows(const double* Srch, const double* Srcl, double* Ah, double* Al, long n, long jb0, int jb,
                             const DiagLds& d, long row0, long row1, int step) {
  const int lane = threadIdx.x & 63;
  dd invk = mkdd(0.0);
  if (lane < jb) invk = dd_recip(mkdd(d.Lh[lane][lane], d.Ll[lane][lane]));
  for (long row = row0; row < row1; row += 4L * step) {
    dd t[4], fin[4];
    bool ok[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const long r = row + (long)q * step;
      ok[q] = r < row1;
      t[q] = (ok[q] && lane < jb) ? ldm(Srch, Srcl, r * n + jb0 + lane) : mkdd(0.0);
      fin[q] = mkdd(0.0);
    }
#pragma unroll 1
    for (int k = 0; k < jb; ++k) {
      const dd lk = lane < jb ? mkdd(d.Lh[lane][k], d.Ll[lane][k]) : mkdd(0.0);
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        dd bk = dd_mul(t[q], invk);
        if (lane == k) fin[q] = bk;
        dd b = rlane(bk, k);
        if (lane > k && lane < jb) t[q] = dd_sub(t[q], dd_mul(b, lk));
      }
    }
#pragma unroll
    for (int q = 0; q < 4; ++q)
      if (ok[q] && lane < jb) stm(Ah, Al, (row + (long)q * step) * n + jb0 + lane, fin[q]);
  }
}

// tile (ti, tc), tc <= ti, of the lower triangle of the trailing matrix from t0, row-major
__device__ __forceinline__ void tile_of(long tile, long& ti, long& tc) {
  ti = (long)((__builtin_sqrt(8.0 * (double)tile + 1.0) - 1.0) / 2.0);
  while ((ti + 1) * (ti + 2) / 2 <= tile) ++ti;
  while (ti * (ti + 1) / 2 > tile) --ti;
  tc = tile - ti * (ti + 1) / 2;
}

// Trailing update (linalg.rs:268-279) of one 64 x 64 tile at rows i0, columns c0 by 256
// threads (ht = 0..255, a 16 x 16 grid of 4 x 4 register blocks): s = sum_k p_ik p_ck over
// the panel columns [jb0, jb0 + jb) from zero, then a = a - s (a read from Src).  The panel
// goes through LDS in 16-deep slabs, the next slab's loads in flight (registers) while the
// current one is multiplied.  As / Bs: this group's LDS staging, [plane][k][row].  Every
// thread of the block calls it (the barriers), `active` false for a group without a tile.
__device__ __forceinline__ void bt_tile(const double* Srch, const double* Srcl, double* Ah, double* Al, long n, long jb0, int jb,
                        long i0, long c0, bool active, int ht, double (*As)[16][64], double (*Bs)[16][64]) {
  const int tx = ht & 15, ty = ht >> 4;
  const int lr = ht >> 2, lkc = (ht & 3) * 4;  // this thread's loads: row lr, slab columns lkc..lkc+3
  dd acc[4][4];
#pragma unroll
  for (int r = 0; r < 4; ++r)
#pragma unroll
    for (int c = 0; c < 4; ++c) acc[r][c] = mkdd(0.0);
  const bool oka = active && i0 + lr < n, okb = active && c0 + lr < n;
  const double* pah = Ah + (i0 + lr) * n + jb0 + lkc;  // this thread's loads (valid when oka / okb)
  const double* pal = Al + (i0 + lr) * n + jb0 + lkc;
  const double* pbh = Ah + (c0 + lr) * n + jb0 + lkc;
  const double* pbl = Al + (c0 + lr) * n + jb0 + lkc;
  double ah[4], al[4], bh[4], bl[4];
#pragma unroll
  for (int qq = 0; qq < 4; ++qq) {
    const bool in = lkc + qq < jb;
    ah[qq] = (oka && in) ? pah[qq] : 0.0;
    al[qq] = (oka && in) ? pal[qq] : 0.0;
    bh[qq] = (okb && in) ? pbh[qq] : 0.0;
    bl[qq] = (okb && in) ? pbl[qq] : 0.0;
  }
  for (int k0 = 0; k0 < jb; k0 += 16) {
#pragma unroll
    for (int qq = 0; qq < 4; ++qq) {
      As[0][lkc + qq][lr] = ah[qq];
      As[1][lkc + qq][lr] = al[qq];
      Bs[0][lkc + qq][lr] = bh[qq];
      Bs[1][lkc + qq][lr] = bl[qq];
    }
    __syncthreads();
    if (k0 + 16 < jb) {  // the next slab's loads in flight during this slab's products
#pragma unroll
      for (int qq = 0; qq < 4; ++qq) {
        const bool in = k0 + 16 + lkc + qq < jb;
        ah[qq] = (oka && in) ? pah[k0 + 16 + qq] : 0.0;
        al[qq] = (oka && in) ? pal[k0 + 16 + qq] : 0.0;
        bh[qq] = (okb && in) ? pbh[k0 + 16 + qq] : 0.0;
        bl[qq] = (okb && in) ? pbl[k0 + 16 + qq] : 0.0;
      }
    }
    const int kmax = min(16, jb - k0);
    if (active)
      for (int kk = 0; kk < kmax; ++kk) {
        dd av[4], bv[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          av[r] = mkdd(As[0][kk][ty * 4 + r], As[1][kk][ty * 4 + r]);
          bv[r] = mkdd(Bs[0][kk][tx * 4 + r], Bs[1][kk][tx * 4 + r]);
        }
#pragma unroll
        for (int r = 0; r < 4; ++r)
#pragma unroll
          for (int c = 0; c < 4; ++c) acc[r][c] = dd_add(acc[r][c], dd_mul(av[r], bv[c]));
      }
    __syncthreads();
  }
  if (active) {
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        long gi = i0 + ty * 4 + r, gc = c0 + tx * 4 + c;
        if (gi < n && gc <= gi) {
          long o = gi * n + gc;
          stm(Ah, Al, o, dd_sub(ldm(Srch, Srcl, o), acc[r][c]));
        }
      }
  }
}

// L L^T X = B for nrhs right-hand sides of one factored instance (all NT threads): blocked
// forward and backward substitution, right-looking (each element takes the blocks'
// contributions in block order, within a block in column order).  Each 64-row diagonal
// block is staged in LDS (Dh / Dl) for the wavefronts' substitution chains, which multiply
// by the pivots' dd reciprocals (formed per lane before the chain) instead of dividing; the
// other rows take the block's contributions with two independent element chains
// per thread.  Yh / Yl: the block's solved values.
constexpr int SOLVE_EPT = 2;
__device__ __forceinline__ void bt_solve(const double* Ah, const double* Al, long n, const double* bh, const double* bl, double* xh,
                         double* xl, long ro, int R, double (*Yh)[NB], double (*Yl)[NB], double (*Dh)[NB + 1],
                         double (*Dl)[NB + 1]) {
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  for (long idx = tid; idx < R * n; idx += NT) {
    xh[ro + idx] = bh[ro + idx];
    xl[ro + idx] = bl[ro + idx];
  }
  auto stage = [&](long b0, int bw) {  // L[b0 + i][b0 + c], c <= i, into Dh / Dl
    for (int idx = tid; idx < bw * bw; idx += NT) {
      const int i = idx / bw, c = idx % bw;
      if (c <= i) {
        const long o = (b0 + i) * n + b0 + c;
        Dh[i][c] = Ah[o];
        Dl[i][c] = Al[o];
      }
    }
  };
  __syncthreads();
  for (long b0 = 0; b0 < n; b0 += NB) {
    const int bw = (int)min((long)NB, n - b0);
    stage(b0, bw);
    __syncthreads();
    if (wv < R) {  // wavefront r: diagonal block of right-hand side r, lane = row
      double* Xh = xh + ro + (long)wv * n;
      double* Xl = xl + ro + (long)wv * n;
      dd y = lane < bw ? ldm(Xh, Xl, b0 + lane) : mkdd(0.0);
      const dd rjj = dd_recip(lane < bw ? mkdd(Dh[lane][lane], Dl[lane][lane]) : mkdd(Dh[0][0], Dl[0][0]));
      for (int j = 0; j < bw; ++j) {
        dd yj = dd_mul(y, rjj);
        if (lane == j) y = yj;
        dd b = rlane(yj, j);
        if (lane > j && lane < bw) y = dd_sub(y, dd_mul(mkdd(Dh[lane][j], Dl[lane][j]), b));
      }
      if (lane < bw) {
        stm(Xh, Xl, b0 + lane, y);
        Yh[wv][lane] = y.hi;
        Yl[wv][lane] = y.lo;
      }
    }
    __syncthreads();
    const long rows = n - b0 - bw, tot = (long)R * rows;
    for (long base = tid; base < tot; base += (long)SOLVE_EPT * NT) {
      dd y[SOLVE_EPT];
      long off[SOLVE_EPT];
      int rr[SOLVE_EPT];
#pragma unroll
      for (int e = 0; e < SOLVE_EPT; ++e) {
        const long idx = base + (long)e * NT;
        rr[e] = idx < tot ? (int)(idx / rows) : 0;
        const long i = idx < tot ? b0 + bw + idx % rows : b0 + bw;
        off[e] = i;
        y[e] = idx < tot ? ldm(xh + ro + (long)rr[e] * n, xl + ro + (long)rr[e] * n, i) : mkdd(0.0);
      }
      // each element's row of L is strided by n across the lanes: the next eight columns'
      // loads go out together, ahead of their eight chained steps
#pragma unroll 1
      for (int j0 = 0; j0 < bw; j0 += 8) {
        double lh[SOLVE_EPT][8], ll[SOLVE_EPT][8];
#pragma unroll
        for (int e = 0; e < SOLVE_EPT; ++e)
#pragma unroll
          for (int q = 0; q < 8; ++q) {
            const bool ok = j0 + q < bw;
            const long o = off[e] * n + b0 + j0 + (ok ? q : 0);
            lh[e][q] = ok ? Ah[o] : 0.0;
            ll[e][q] = ok ? Al[o] : 0.0;
          }
#pragma unroll
        for (int q = 0; q < 8; ++q) {
          if (j0 + q < bw) {
#pragma unroll
            for (int e = 0; e < SOLVE_EPT; ++e)
              y[e] = dd_sub(y[e], dd_mul(mkdd(lh[e][q], ll[e][q]), mkdd(Yh[rr[e]][j0 + q], Yl[rr[e]][j0 + q])));
          }
        }
      }
#pragma unroll
      for (int e = 0; e < SOLVE_EPT; ++e)
        if (base + (long)e * NT < tot) stm(xh + ro + (long)rr[e] * n, xl + ro + (long)rr[e] * n, off[e], y[e]);
    }
    __syncthreads();
  }
  for (long b0 = ((n - 1) / NB) * NB; b0 >= 0; b0 -= NB) {
    const int bw = (int)min((long)NB, n - b0);
    stage(b0, bw);
    __syncthreads();
    if (wv < R) {
      double* Xh = xh + ro + (long)wv * n;
      double* Xl = xl + ro + (long)wv * n;
      dd y = lane < bw ? ldm(Xh, Xl, b0 + lane) : mkdd(0.0);
      const dd rjj = dd_recip(lane < bw ? mkdd(Dh[lane][lane], Dl[lane][lane]) : mkdd(Dh[0][0], Dl[0][0]));
      for (int j = bw - 1; j >= 0; --j) {
        dd xj = dd_mul(y, rjj);  // lane j's: its own pivot
        if (lane == j) y = xj;
        dd b = rlane(xj, j);
        if (lane < j) y = dd_sub(y, dd_mul(mkdd(Dh[j][lane], Dl[j][lane]), b));
      }
      if (lane < bw) {
        stm(Xh, Xl, b0 + lane, y);
        Yh[wv][lane] = y.hi;
        Yl[wv][lane] = y.lo;
      }
    }
    __syncthreads();
    const long tot = (long)R * b0;
    for (long base = tid; base < tot; base += (long)SOLVE_EPT * NT) {
      dd y[SOLVE_EPT];
      long ii[SOLVE_EPT];
      int rr[SOLVE_EPT];
#pragma unroll
      for (int e = 0; e < SOLVE_EPT; ++e) {
        const long idx = base + (long)e * NT;
        rr[e] = idx < tot ? (int)(idx / b0) : 0;
        ii[e] = idx < tot ? idx % b0 : 0;
        y[e] = idx < tot ? ldm(xh + ro + (long)rr[e] * n, xl + ro + (long)rr[e] * n, ii[e]) : mkdd(0.0);
      }
#pragma unroll 1
      for (int j = bw - 1; j >= 0; --j) {
#pragma unroll
        for (int e = 0; e < SOLVE_EPT; ++e)
          y[e] = dd_sub(y[e], dd_mul(ldm(Ah, Al, (b0 + j) * n + ii[e]), mkdd(Yh[rr[e]][j], Yl[rr[e]][j])));
      }
#pragma unroll
      for (int e = 0; e < SOLVE_EPT; ++e)
        if (base + (long)e * NT < tot) stm(xh + ro + (long)rr[e] * n, xl + ro + (long)rr[e] * n, ii[e], y[e]);
    }
    __syncthreads();
  }
}

__global__ void k_batch_reset(int* counter) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *counter = 0;
}

// ---- schedule 1: one workgroup per instance (persistent, instances from an atomic queue,
// largest n^3 first); each workgroup copies S to L and runs the whole factor and solve.
__global__ __launch_bounds__(NT) void k_batch_chol_solve(const BatchDesc* __restrict__ desc, const int* __restrict__ order,
                                                         int count, int* __restrict__ counter,
                                                         const double* __restrict__ sh, const double* __restrict__ sl,
                                                         double* __restrict__ lh, double* __restrict__ ll,
                                                         const double* __restrict__ bh, const double* __restrict__ bl,
                                                         double* __restrict__ xh, double* __restrict__ xl, int nrhs,
                                                         int* __restrict__ status) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  BatchLds& lds = *reinterpret_cast<BatchLds*>(smem_raw);
  __shared__ int s_idx, s_fail;
  __shared__ double s_inv[2];
  __shared__ double Yh[MAXRHS][NB], Yl[MAXRHS][NB];
  const int tid = threadIdx.x, wv = tid >> 6;

  for (;;) {
    if (tid == 0) s_idx = atomicAdd(counter, 1);
    __syncthreads();
    const int q = s_idx;
    __syncthreads();
    if (q >= count) break;
    const int inst = order[q];
    const long n = desc[inst].n;
    const long mo = desc[inst].mat_off, ro = desc[inst].rhs_off;
    const double* Sh = sh + mo;
    const double* Sl = sl + mo;
    double* Ah = lh + mo;
    double* Al = ll + mo;
    // L = S (the factorisation works in place on the copy)
    for (long i = tid; i < n * n; i += NT) {
      Ah[i] = Sh[i];
      Al[i] = Sl[i];
    }
    __syncthreads();
    bool ok = true;
    for (long jb0 = 0; jb0 < n && ok; jb0 += NB) {
      const int jb = (int)min((long)NB, n - jb0);
      ok = bt_diag(Ah, Al, Ah, Al, n, jb0, jb, lds.d, s_inv, &s_fail, status, inst);
      if (!ok) break;
      bt_trsm_rows(Ah, Al, Ah, Al, n, jb0, jb, lds.d, jb0 + jb + wv, n, NT / 64);
      __syncthreads();
      // trailing tiles, two at a time (one per 256-thread half)
      const long t0 = jb0 + jb, rem = n - t0;
      if (rem > 0) {
        const long T = (rem + 63) / 64;
        const long ntiles = T * (T + 1) / 2;
        const int half = tid >> 8, ht = tid & 255;
        for (long base = 0; base < ntiles; base += 2) {
          const long tile = base + half;
          long ti = 0, tc = 0;
          if (tile < ntiles) tile_of(tile, ti, tc);
          bt_tile(Ah, Al, Ah, Al, n, jb0, jb, t0 + ti * 64, t0 + tc * 64, tile < ntiles, ht, lds.t.As[half],
                  lds.t.Bs[half]);
        }
      }
      __syncthreads();
    }
    if (!ok) {
      __syncthreads();
      continue;
    }
    if (tid == 0) status[inst] = 0;
    bt_solve(Ah, Al, n, bh, bl, xh, xl, ro, nrhs, Yh, Yl, lds.d.Lh, lds.d.Ll);
  }
}

// ---- schedule 2: panel steps across the batch.  Step p runs three launches over the
// instances still active (n > 64 p), which are a prefix of the order (n descending):
//   k_bp_diag    one workgroup per instance: block p's factor;
//   k_bp_trsm    eight rows per workgroup: trsm of the rows below block p;
//   k_bp_update  one trailing 64 x 64 tile per workgroup (256 threads);
// then the block-step solves below.  Per-step prefix sums of row groups and tiles
// map a workgroup to its (instance, piece).  Each element receives exactly the operations
// of schedule 1, in the same order.  Step 0 reads S and writes L, and copies S's upper
// triangle (the factor leaves it as it is), each piece the mirror of what it factors.
constexpr int TRSM_ROWS = 128;  // rows per k_bp_trsm workgroup, one thread each

__device__ __forceinline__ int piece_owner(const long long* pre, int A, long long g) {  // pre[j] <= g < pre[j+1]
  int lo = 0, hi = A;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (pre[mid] <= g) lo = mid;
    else hi = mid;
  }
  return lo;
}

__global__ void k_bp_init(int* status, int count) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) status[i] = 0;
}

template <bool FIRST>
__global__ __launch_bounds__(NT) void k_bp_diag(const BatchDesc* __restrict__ desc, const int* __restrict__ order,
                                                int* __restrict__ status, const double* __restrict__ sh,
                                                const double* __restrict__ sl, double* __restrict__ lh,
                                                double* __restrict__ ll, long jb0) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  DiagLds& d = *reinterpret_cast<DiagLds*>(smem_raw);
  __shared__ int s_fail;
  __shared__ double s_inv[2];
  const int inst = order[blockIdx.x];
  if (status[inst]) return;
  const long n = desc[inst].n, mo = desc[inst].mat_off;
  const int jb = (int)min((long)NB, n - jb0);
  const double* Srch = (FIRST ? sh : lh) + mo;
  const double* Srcl = (FIRST ? sl : ll) + mo;
  if (FIRST)  // the block's strict upper triangle, as in S
    for (int idx = threadIdx.x; idx < jb * jb; idx += NT) {
      const int i = idx / jb, c = idx % jb;
      if (c > i) {
        const long o = (jb0 + i) * n + jb0 + c;
        lh[mo + o] = sh[mo + o];
        ll[mo + o] = sl[mo + o];
      }
    }
  bt_diag(Srch, Srcl, lh + mo, ll + mo, n, jb0, jb, d, s_inv, &s_fail, status, inst);
}

// trsm_right_lt (linalg.rs:224-235) of TRSM_ROWS rows below the step's diagonal block, one
// thread per row, the block's lower triangle (packed) and the pivots' reciprocals in LDS,
// read at uniform addresses.  Entry j of a row is t_j = a_j - sum_{k<j} b_k l_jk (k
// ascending), b_j = t_j * recip(l_jj), as in bt_trsm_rows (same operations and order, so
// bit-identical).  Left-looking in four phases of 16 columns: phase P holds its 16 entries in
// registers, subtracts the earlier phases' finished values b_k (k < 16P, read back from
// the row it stored) in k order, then finishes its own 16 columns as a fully unrolled
// triangle.  Exactly the 2016 useful multiply-subtracts per row are issued, and a row
// holds 16 entries in registers instead of 64.
template <bool FIRST>
__global__ __launch_bounds__(TRSM_ROWS) void k_bp_trsm(const BatchDesc* __restrict__ desc, const int* __restrict__ order,
                                                       const int* __restrict__ status, const long long* __restrict__ pre,
                                                       int active, const double* __restrict__ sh,
                                                       const double* __restrict__ sl, double* lh, double* ll, long jb0) {
  __shared__ double2 Lp[NB * (NB + 1) / 2];  // (j, k), k <= j, at j (j + 1) / 2 + k
  __shared__ double2 Inv[NB];
  const int j = piece_owner(pre, active, blockIdx.x);
  const int inst = order[j];
  if (status[inst]) return;
  const long n = desc[inst].n, mo = desc[inst].mat_off;
  const int tid = threadIdx.x;
  // rows below a full block p exist, so the block is 64 wide
  for (int r = 0; r < NB; ++r)
    if (tid <= r) {
      const long o = mo + (jb0 + r) * n + jb0 + tid;
      Lp[r * (r + 1) / 2 + tid] = make_double2(lh[o], ll[o]);
    }
  __syncthreads();
  if (tid < NB) {
    const double2 v = Lp[tid * (tid + 1) / 2 + tid];
    const dd r = dd_recip(mkdd(v.x, v.y));
    Inv[tid] = make_double2(r.hi, r.lo);
  }
  __syncthreads();
  const long row = jb0 + NB + (long)(blockIdx.x - pre[j]) * TRSM_ROWS + tid;
  if (row >= n) return;
  const double* Srch = (FIRST ? sh : lh) + mo;
  const double* Srcl = (FIRST ? sl : ll) + mo;
  double* Rh = lh + mo + row * n + jb0;  // this row's entries of the panel (the output)
  double* Rl = ll + mo + row * n + jb0;
  constexpr int W = 16;
#pragma unroll 1
  for (int P = 0; P < NB / W; ++P) {
    dd t[W];
#pragma unroll
    for (int q = 0; q < W; ++q) t[q] = ldm(Srch, Srcl, row * n + jb0 + W * P + q);
    // the earlier phases' finished values, k ascending, eight at a time
#pragma unroll 1
    for (int k0 = 0; k0 < W * P; k0 += 8) {
      dd bk[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) bk[u] = mkdd(Rh[k0 + u], Rl[k0 + u]);
#pragma unroll
      for (int u = 0; u < 8; ++u) {
#pragma unroll
        for (int q = 0; q < W; ++q) {
          const int c = W * P + q;
          const double2 l = Lp[c * (c + 1) / 2 + k0 + u];
          t[q] = dd_sub(t[q], dd_mul(bk[u], mkdd(l.x, l.y)));
        }
        __builtin_amdgcn_sched_barrier(0);  // keeps the LDS operands of one k live at a time
      }
    }
    // the phase's own triangle
#pragma unroll
    for (int kk = 0; kk < W; ++kk) {
      const int k = W * P + kk;
      const double2 iv = Inv[k];
      const dd b = dd_mul(t[kk], mkdd(iv.x, iv.y));
      t[kk] = b;
#pragma unroll
      for (int q = kk + 1; q < W; ++q) {
        const int c = W * P + q;
        const double2 l = Lp[c * (c + 1) / 2 + k];
        t[q] = dd_sub(t[q], dd_mul(b, mkdd(l.x, l.y)));
      }
      __builtin_amdgcn_sched_barrier(0);
    }
#pragma unroll
    for (int q = 0; q < W; ++q) {
      Rh[W * P + q] = t[q].hi;
      Rl[W * P + q] = t[q].lo;
    }
    __builtin_amdgcn_sched_barrier(0);  // the next phase's loads stay in the next phase
  }
  if (FIRST)  // the mirror of this row's panel entries: S[jb0 + k][row], k < 64
    for (int k = 0; k < NB; ++k) {
      const long o = mo + (jb0 + k) * n + row;
      lh[o] = sh[o];
      ll[o] = sl[o];
    }
}

// PART 1: the tiles of the next panel's column block (tc = 0); PART 2: the others
template <bool FIRST, int PART>
__global__ __launch_bounds__(256) void k_bp_update(const BatchDesc* __restrict__ desc, const int* __restrict__ order,
                                                   const int* __restrict__ status, const long long* __restrict__ pre,
                                                   int active, const double* __restrict__ sh,
                                                   const double* __restrict__ sl, double* __restrict__ lh,
                                                   double* __restrict__ ll, long jb0) {
  __shared__ double As[2][16][64], Bs[2][16][64];
  const int j = piece_owner(pre, active, blockIdx.x);
  const int inst = order[j];
  if (status[inst]) return;
  const long n = desc[inst].n, mo = desc[inst].mat_off;
  long ti, tc;
  if (PART == 1) {
    ti = (long)blockIdx.x - pre[j];
    tc = 0;
  } else {  // the lower triangle of tiles from (1, 1)
    tile_of((long)blockIdx.x - pre[j], ti, tc);
    ++ti;
    ++tc;
  }
  const long t0 = jb0 + NB, i0 = t0 + ti * 64, c0 = t0 + tc * 64;
  const double* Srch = (FIRST ? sh : lh) + mo;
  const double* Srcl = (FIRST ? sl : ll) + mo;
  bt_tile(Srch, Srcl, lh + mo, ll + mo, n, jb0, NB, i0, c0, true, threadIdx.x, As, Bs);
  if (FIRST) {  // the mirror tile of S's upper triangle (its strict upper part on the diagonal)
    for (int idx = threadIdx.x; idx < 64 * 64; idx += 256) {
      const long r = c0 + idx / 64, c = i0 + idx % 64;
      if (r < n && c < n && c > r) {
        const long o = mo + r * n + c;
        lh[o] = sh[o];
        ll[o] = sl[o];
      }
    }
  }
}


// ---- the panel-step solves: L L^T X = B for every instance in block steps across the
// batch, the same operations in the same order as bt_solve (results identical to schedule
// 1's).  X = B first; then per block b ascending: k_bs_diag<true> (one workgroup per active
// instance: the block's forward substitution chains, one wavefront per right-hand side,
// lane = row) and k_bs_update<true> (64 rows below the block per workgroup: each element
// takes the block's 64 contributions in column order); then per block b descending the
// same with the backward chains and the 64-row groups above the block.  A 64 x 64 tile of
// L is staged in LDS with coalesced row loads (rows of the block's columns forward, rows of
// the block's rows backward), so no lane walks a row of L.
struct SolveLds {
  double Th[NB][NB + 1], Tl[NB][NB + 1];
  double Yh[MAXRHS][NB], Yl[MAXRHS][NB];
};

__global__ void k_bs_copy(const BatchDesc* __restrict__ desc, const int* __restrict__ status,
                          const double* __restrict__ bh, const double* __restrict__ bl, double* __restrict__ xh,
                          double* __restrict__ xl, int nrhs) {
  const int inst = blockIdx.x;
  if (status[inst]) return;
  const long n = desc[inst].n, ro = desc[inst].rhs_off;
  for (long idx = threadIdx.x; idx < (long)nrhs * n; idx += blockDim.x) {
    xh[ro + idx] = bh[ro + idx];
    xl[ro + idx] = bl[ro + idx];
  }
}

template <bool FWD>
__global__ __launch_bounds__(NT) void k_bs_diag(const BatchDesc* __restrict__ desc, const int* __restrict__ order,
                                                const int* __restrict__ status, const double* __restrict__ lh,
                                                const double* __restrict__ ll, double* __restrict__ xh,
                                                double* __restrict__ xl, int nrhs, long b0) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  DiagLds& d = *reinterpret_cast<DiagLds*>(smem_raw);  // Lh / Ll only: the staged diagonal block
  double (*Dh)[NB + 1] = d.Lh;
  double (*Dl)[NB + 1] = d.Ll;
  const int inst = order[blockIdx.x];
  if (status[inst]) return;
  const long n = desc[inst].n, mo = desc[inst].mat_off, ro = desc[inst].rhs_off;
  const int bw = (int)min((long)NB, n - b0);
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  for (int idx = tid; idx < bw * bw; idx += NT) {
    const int i = idx / bw, c = idx % bw;
    if (c <= i) {
      const long o = mo + (b0 + i) * n + b0 + c;
      Dh[i][c] = lh[o];
      Dl[i][c] = ll[o];
    }
  }
  __syncthreads();
  for (int r = wv; r < nrhs; r += NT / 64) {
    double* Xh = xh + ro + (long)r * n;
    double* Xl = xl + ro + (long)r * n;
    dd y = lane < bw ? ldm(Xh, Xl, b0 + lane) : mkdd(0.0);
    const dd rjj = dd_recip(lane < bw ? mkdd(Dh[lane][lane], Dl[lane][lane]) : mkdd(Dh[0][0], Dl[0][0]));
    if (FWD) {
      for (int j = 0; j < bw; ++j) {
        dd yj = dd_mul(y, rjj);
        if (lane == j) y = yj;
        dd bj = rlane(yj, j);
        if (lane > j && lane < bw) y = dd_sub(y, dd_mul(mkdd(Dh[lane][j], Dl[lane][j]), bj));
      }
    } else {
      for (int j = bw - 1; j >= 0; --j) {
        dd xj = dd_mul(y, rjj);
        if (lane == j) y = xj;
        dd bj = rlane(xj, j);
        if (lane < j) y = dd_sub(y, dd_mul(mkdd(Dh[j][lane], Dl[j][lane]), bj));
      }
    }
    if (lane < bw) stm(Xh, Xl, b0 + lane, y);
  }
}

// 64 rows (below the block forward, above it backward) of one instance take the block's
// contributions: forward x_i -= sum_j l[i][b0+j] y_j (j ascending), backward
// x_i -= sum_j l[b0+j][i] x_{b0+j} (j descending); thread (row, right-hand-side group)
template <bool FWD>
__global__ __launch_bounds__(256) void k_bs_update(const BatchDesc* __restrict__ desc, const int* __restrict__ order,
                                                   const int* __restrict__ status, const long long* __restrict__ pre,
                                                   int active, const double* __restrict__ lh,
                                                   const double* __restrict__ ll, double* __restrict__ xh,
                                                   double* __restrict__ xl, int nrhs, long b0) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  SolveLds& sl = *reinterpret_cast<SolveLds*>(smem_raw);
  double (*Th)[NB + 1] = sl.Th;  // forward [row][col], backward [col][row]
  double (*Tl)[NB + 1] = sl.Tl;
  double (*Yh)[NB] = sl.Yh;
  double (*Yl)[NB] = sl.Yl;
  const int j = piece_owner(pre, active, blockIdx.x);
  const int inst = order[j];
  if (status[inst]) return;
  const long n = desc[inst].n, mo = desc[inst].mat_off, ro = desc[inst].rhs_off;
  const int bw = (int)min((long)NB, n - b0);
  const long c = (long)blockIdx.x - pre[j];
  const long i0 = FWD ? b0 + bw + c * NB : c * NB;
  const int rows = (int)min((long)NB, (FWD ? n : b0) - i0);
  const int tid = threadIdx.x;
  for (int idx = tid; idx < NB * NB; idx += 256) {
    const int a = idx / NB, e = idx % NB;  // forward: row a, column e; backward: block row a, row e
    const bool ok = FWD ? (a < rows && e < bw) : (a < bw && e < rows);
    const long o = mo + (FWD ? (i0 + a) * n + b0 + e : (b0 + a) * n + i0 + e);
    Th[a][e] = ok ? lh[o] : 0.0;
    Tl[a][e] = ok ? ll[o] : 0.0;
  }
  for (int idx = tid; idx < nrhs * NB; idx += 256) {
    const int r = idx / NB, e = idx % NB;
    const long o = ro + (long)r * n + b0 + e;
    Yh[r][e] = e < bw ? xh[o] : 0.0;
    Yl[r][e] = e < bw ? xl[o] : 0.0;
  }
  __syncthreads();
  const int row = tid & 63;
  if (row >= rows) return;
  for (int r = tid >> 6; r < nrhs; r += 4) {
    const long o = ro + (long)r * n + i0 + row;
    dd y = ldm(xh, xl, o);
    if (FWD) {
#pragma unroll 4
      for (int q = 0; q < bw; ++q) y = dd_sub(y, dd_mul(mkdd(Th[row][q], Tl[row][q]), mkdd(Yh[r][q], Yl[r][q])));
    } else {
#pragma unroll 4
      for (int q = bw - 1; q >= 0; --q) y = dd_sub(y, dd_mul(mkdd(Th[q][row], Tl[q][row]), mkdd(Yh[r][q], Yl[r][q])));
    }
    stm(xh, xl, o, y);
  }
}

}  // namespace
}  // namespace hp

using namespace hp;

struct hp_batch {
  int count = 0, nrhs = 0;
  int schedule = 0;  // HP_BATCH_PER_INSTANCE or HP_BATCH_PANEL_STEPS
  DevBuf desc, order, counter, pre;
  // panel steps: per step p, the active instances (a prefix of `order`), and the offsets of
  // its row-group and tile prefix sums in `pre` with their totals
  std::vector<int> active;
  std::vector<long long> trsm_off, trsm_tot, next_off, next_tot, rest_off, rest_tot;
  // the block-step solves: per block b, the 64-row groups below (forward) and above
  // (backward) the block over the instances active at b
  std::vector<long long> fs_off, fs_tot, bs_off, bs_tot;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  // panel steps: the second stream (the tiles beyond the next panel) and its events
  hipStream_t s2 = nullptr;
  std::vector<hipEvent_t> ev;
  ~hp_batch() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (s2) {
      (void)hipStreamSynchronize(s2);
      (void)hipStreamDestroy(s2);
    }
    for (hipEvent_t e : ev) (void)hipEventDestroy(e);
  }
};

namespace {
std::atomic<int> g_batch_schedule{0};  // 0: the autotuner chooses; else pinned (hp_batch_set_schedule)
std::mutex g_batch_mu;
std::map<uint64_t, int> g_batch_tuned;  // schedule per batch class (under g_batch_mu)

uint64_t batch_class(const int* sizes, size_t count, int nrhs) {  // count, max n, n^3 total: powers of two
  long long mx = 0;
  double macs = 0;
  for (size_t i = 0; i < count; ++i) {
    mx = std::max<long long>(mx, sizes[i]);
    macs += (double)sizes[i] * sizes[i] * sizes[i];
  }
  auto lg = [](double v) { return (uint64_t)std::max(0, (int)std::ceil(std::log2(std::max(v, 1.0)))); };
  return (lg((double)count) << 24) | (lg((double)mx) << 16) | (lg(macs) << 4) | (uint64_t)nrhs;
}

int batch_prepare(hp_batch* b, const int* sizes, size_t count, int nrhs) {
  if (nrhs < 1 || nrhs > MAXRHS) {
    set_error("batch: nrhs must be in [1, %d]", MAXRHS);
    return HP_EINVAL;
  }
  std::vector<BatchDesc> d(count);
  long long mo = 0, ro = 0;
  long long nmax = 0;
  for (size_t i = 0; i < count; ++i) {
    if (sizes[i] <= 0) {
      set_error("batch: instance %zu has size %d", i, sizes[i]);
      return HP_ESHAPE;
    }
    d[i].n = sizes[i];
    d[i].mat_off = mo;
    d[i].rhs_off = ro;
    mo += (long long)sizes[i] * sizes[i];
    ro += (long long)nrhs * sizes[i];
    nmax = std::max<long long>(nmax, sizes[i]);
  }
  std::vector<int> ord(count);
  std::iota(ord.begin(), ord.end(), 0);
  std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) { return sizes[x] > sizes[y]; });
  // panel steps
  const int steps = (int)((nmax + NB - 1) / NB);
  std::vector<long long> pre;
  b->active.assign(steps, 0);
  b->trsm_off.assign(steps, 0);
  b->trsm_tot.assign(steps, 0);
  b->next_off.assign(steps, 0);
  b->next_tot.assign(steps, 0);
  b->rest_off.assign(steps, 0);
  b->rest_tot.assign(steps, 0);
  b->fs_off.assign(steps, 0);
  b->fs_tot.assign(steps, 0);
  b->bs_off.assign(steps, 0);
  b->bs_tot.assign(steps, 0);
  for (int p = 0; p < steps; ++p) {
    const long long jb0 = (long long)p * NB;
    int A = 0;
    while (A < (int)count && sizes[ord[A]] > jb0) ++A;
    b->active[p] = A;
    b->trsm_off[p] = (long long)pre.size();
    long long acc = 0;
    for (int j = 0; j < A; ++j) {
      pre.push_back(acc);
      const long long rows = std::max<long long>(0, sizes[ord[j]] - jb0 - NB);
      acc += (rows + TRSM_ROWS - 1) / TRSM_ROWS;
    }
    pre.push_back(acc);
    b->trsm_tot[p] = acc;
    b->next_off[p] = (long long)pre.size();
    acc = 0;
    for (int j = 0; j < A; ++j) {
      pre.push_back(acc);
      const long long rows = std::max<long long>(0, sizes[ord[j]] - jb0 - NB), T = (rows + 63) / 64;
      acc += T;
    }
    pre.push_back(acc);
    b->next_tot[p] = acc;
    b->rest_off[p] = (long long)pre.size();
    acc = 0;
    for (int j = 0; j < A; ++j) {
      pre.push_back(acc);
      const long long rows = std::max<long long>(0, sizes[ord[j]] - jb0 - NB), T = (rows + 63) / 64;
      acc += T * (T - 1) / 2;
    }
    pre.push_back(acc);
    b->rest_tot[p] = acc;
    b->fs_off[p] = (long long)pre.size();
    acc = 0;
    for (int j = 0; j < A; ++j) {
      pre.push_back(acc);
      acc += (std::max<long long>(0, sizes[ord[j]] - jb0 - NB) + NB - 1) / NB;
    }
    pre.push_back(acc);
    b->fs_tot[p] = acc;
    b->bs_off[p] = (long long)pre.size();
    acc = 0;
    for (int j = 0; j < A; ++j) {
      pre.push_back(acc);
      acc += jb0 / NB;
    }
    pre.push_back(acc);
    b->bs_tot[p] = acc;
  }
  int rc;
  if ((rc = b->desc.alloc(count * sizeof(BatchDesc))) || (rc = b->order.alloc(count * sizeof(int))) ||
      (rc = b->counter.alloc(sizeof(int))) || (rc = b->pre.alloc(std::max<size_t>(1, pre.size()) * sizeof(long long))))
    return rc;
  HP_HIP(hipMemcpy(b->desc.p, d.data(), count * sizeof(BatchDesc), hipMemcpyHostToDevice));
  HP_HIP(hipMemcpy(b->order.p, ord.data(), count * sizeof(int), hipMemcpyHostToDevice));
  if (!pre.empty()) HP_HIP(hipMemcpy(b->pre.p, pre.data(), pre.size() * sizeof(long long), hipMemcpyHostToDevice));
  b->count = (int)count;
  b->nrhs = nrhs;
  static std::once_flag attr_once;
  static hipError_t attr_err = hipSuccess;
  std::call_once(attr_once, [] {
    hipError_t e[7] = {
        hipFuncSetAttribute((const void*)k_batch_chol_solve, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)sizeof(BatchLds)),
        hipFuncSetAttribute((const void*)k_bp_diag<true>, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)sizeof(DiagLds)),
        hipFuncSetAttribute((const void*)k_bp_diag<false>, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)sizeof(DiagLds)),
        hipFuncSetAttribute((const void*)k_bs_diag<true>, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)offsetof(DiagLds, Sh)),
        hipFuncSetAttribute((const void*)k_bs_diag<false>, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)offsetof(DiagLds, Sh)),
        hipFuncSetAttribute((const void*)k_bs_update<true>, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)sizeof(SolveLds)),
        hipFuncSetAttribute((const void*)k_bs_update<false>, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)sizeof(SolveLds))};
    for (hipError_t x : e)
      if (x != hipSuccess) attr_err = x;
  });
  HP_HIP(attr_err);
  return HP_OK;
}

int batch_enqueue(hp_batch* b, int schedule, const double* s_hi, const double* s_lo, double* l_hi, double* l_lo,
                  const double* b_hi, const double* b_lo, double* x_hi, double* x_lo, int* status, hipStream_t s) {
  const BatchDesc* desc = (const BatchDesc*)b->desc.p;
  const int* order = (const int*)b->order.p;
  if (schedule == HP_BATCH_PER_INSTANCE) {
    int dev = 0, cus = 0;
    HP_HIP(hipGetDevice(&dev));
    HP_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    const int grid = std::max(1, std::min(b->count, cus));  // one resident workgroup per CU
    hipLaunchKernelGGL(k_batch_reset, dim3(1), dim3(64), 0, s, (int*)b->counter.p);
    HP_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_batch_chol_solve, dim3(grid), dim3(NT), sizeof(BatchLds), s, desc, order, b->count,
                       (int*)b->counter.p, s_hi, s_lo, l_hi, l_lo, b_hi, b_lo, x_hi, x_lo, b->nrhs, status);
    HP_LAUNCH_CHECK();
    return HP_OK;
  }
  const int steps = (int)b->active.size();
  // lookahead on two streams: per step p, `s` runs the diagonal block, the row solves and
  // the next panel's column block of the trailing update (tiles (i, p+1)), which is all
  // step p+1 waits for; `s2` runs the rest of the trailing update beside step p+1's
  // diagonal block and row solves (they touch column block p+1 only).  Every tile still
  // takes its panels' updates in order: the next column block's tiles wait for the
  // previous step's rest, and the rest of step p waits for step p's row solves.
  if (!b->s2) {
    HP_HIP(hipStreamCreateWithFlags(&b->s2, hipStreamNonBlocking));
    b->ev.assign(2 * steps + 2, nullptr);
    for (hipEvent_t& e : b->ev) HP_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  }
  hipStream_t s2 = b->s2;
  hipEvent_t ev_fork = b->ev[2 * steps], ev_join = b->ev[2 * steps + 1];
  hipLaunchKernelGGL(k_bp_init, dim3((unsigned)((b->count + 255) / 256)), dim3(256), 0, s, status, b->count);
  HP_LAUNCH_CHECK();
  HP_HIP(hipEventRecord(ev_fork, s));
  HP_HIP(hipStreamWaitEvent(s2, ev_fork, 0));
  const long long* pre = (const long long*)b->pre.p;
  for (int p = 0; p < steps; ++p) {
    const long jb0 = (long)p * NB;
    const bool first = p == 0;
    const int A = b->active[p];
    hipEvent_t ev_trsm = b->ev[2 * p], ev_rest = b->ev[2 * p + 1];
    if (first)
      hipLaunchKernelGGL(k_bp_diag<true>, dim3(A), dim3(NT), offsetof(DiagLds, Sh), s, desc, order, status, s_hi, s_lo,
                         l_hi, l_lo, jb0);
    else
      hipLaunchKernelGGL(k_bp_diag<false>, dim3(A), dim3(NT), offsetof(DiagLds, Sh), s, desc, order, status, s_hi,
                         s_lo, l_hi, l_lo, jb0);
    HP_LAUNCH_CHECK();
    if (b->trsm_tot[p] > 0) {
      const unsigned g = (unsigned)b->trsm_tot[p];
      if (first)
        hipLaunchKernelGGL(k_bp_trsm<true>, dim3(g), dim3(TRSM_ROWS), 0, s, desc, order, status, pre + b->trsm_off[p],
                           A, s_hi, s_lo, l_hi, l_lo, jb0);
      else
        hipLaunchKernelGGL(k_bp_trsm<false>, dim3(g), dim3(TRSM_ROWS), 0, s, desc, order, status, pre + b->trsm_off[p],
                           A, s_hi, s_lo, l_hi, l_lo, jb0);
      HP_LAUNCH_CHECK();
    }
    HP_HIP(hipEventRecord(ev_trsm, s));
    if (p > 0) HP_HIP(hipStreamWaitEvent(s, b->ev[2 * p - 1], 0));  // the previous step's rest
    if (b->next_tot[p] > 0) {
      const unsigned g = (unsigned)b->next_tot[p];
      if (first)
        hipLaunchKernelGGL((k_bp_update<true, 1>), dim3(g), dim3(256), 0, s, desc, order, status,
                           pre + b->next_off[p], A, s_hi, s_lo, l_hi, l_lo, jb0);
      else
        hipLaunchKernelGGL((k_bp_update<false, 1>), dim3(g), dim3(256), 0, s, desc, order, status,
                           pre + b->next_off[p], A, s_hi, s_lo, l_hi, l_lo, jb0);
      HP_LAUNCH_CHECK();
    }
    HP_HIP(hipStreamWaitEvent(s2, ev_trsm, 0));
    if (b->rest_tot[p] > 0) {
      const unsigned g = (unsigned)b->rest_tot[p];
      if (first)
        hipLaunchKernelGGL((k_bp_update<true, 2>), dim3(g), dim3(256), 0, s2, desc, order, status,
                           pre + b->rest_off[p], A, s_hi, s_lo, l_hi, l_lo, jb0);
      else
        hipLaunchKernelGGL((k_bp_update<false, 2>), dim3(g), dim3(256), 0, s2, desc, order, status,
                           pre + b->rest_off[p], A, s_hi, s_lo, l_hi, l_lo, jb0);
      HP_LAUNCH_CHECK();
    }
    HP_HIP(hipEventRecord(ev_rest, s2));
  }
  HP_HIP(hipEventRecord(ev_join, s2));
  HP_HIP(hipStreamWaitEvent(s, ev_join, 0));
  // the solves in block steps across the batch (one workgroup per instance took up to 3 ms
  // for an n = 1024 instance with four right-hand sides: a single CU's FP64 rate)
  hipLaunchKernelGGL(k_bs_copy, dim3(b->count), dim3(256), 0, s, desc, static_cast<const int*>(status), b_hi, b_lo,
                     x_hi, x_lo, b->nrhs);
  HP_LAUNCH_CHECK();
  for (int fwd = 1; fwd >= 0; --fwd)
    for (int q = 0; q < steps; ++q) {
      const int p = fwd ? q : steps - 1 - q;
      const long b0 = (long)p * NB;
      const int A = b->active[p];
      const long long off = fwd ? b->fs_off[p] : b->bs_off[p], tot = fwd ? b->fs_tot[p] : b->bs_tot[p];
      if (fwd) {
        hipLaunchKernelGGL(k_bs_diag<true>, dim3(A), dim3(NT), offsetof(DiagLds, Sh), s, desc, order, static_cast<const int*>(status),
                           static_cast<const double*>(l_hi), static_cast<const double*>(l_lo), x_hi, x_lo, b->nrhs,
                           b0);
        if (tot > 0)
          hipLaunchKernelGGL(k_bs_update<true>, dim3((unsigned)tot), dim3(256), sizeof(SolveLds), s, desc, order,
                             static_cast<const int*>(status), pre + off, A, static_cast<const double*>(l_hi),
                             static_cast<const double*>(l_lo), x_hi, x_lo, b->nrhs, b0);
      } else {
        hipLaunchKernelGGL(k_bs_diag<false>, dim3(A), dim3(NT), offsetof(DiagLds, Sh), s, desc, order, static_cast<const int*>(status),
                           static_cast<const double*>(l_hi), static_cast<const double*>(l_lo), x_hi, x_lo, b->nrhs,
                           b0);
        if (tot > 0)
          hipLaunchKernelGGL(k_bs_update<false>, dim3((unsigned)tot), dim3(256), sizeof(SolveLds), s, desc, order,
                             static_cast<const int*>(status), pre + off, A, static_cast<const double*>(l_hi),
                             static_cast<const double*>(l_lo), x_hi, x_lo, b->nrhs, b0);
      }
      HP_LAUNCH_CHECK();
    }
  return HP_OK;
}

// The schedule for this batch: pinned, tuned before for its class, or timed now (each
// schedule once on the caller's buffers, HIP events on `s`; the results are overwritten by
// the run that follows)
int batch_choose(hp_batch* b, const int* sizes, const double* s_hi, const double* s_lo, double* l_hi, double* l_lo,
                 const double* b_hi, const double* b_lo, double* x_hi, double* x_lo, int* status, hipStream_t s) {
  const int pinned = g_batch_schedule.load();
  if (pinned) {
    b->schedule = pinned;
    return HP_OK;
  }
  const uint64_t key = batch_class(sizes, (size_t)b->count, b->nrhs);
  {
    std::lock_guard<std::mutex> lock(g_batch_mu);
    auto it = g_batch_tuned.find(key);
    if (it != g_batch_tuned.end()) {
      b->schedule = it->second;
      return HP_OK;
    }
  }
  hipEvent_t e[3];
  for (hipEvent_t& x : e) HP_HIP(hipEventCreate(&x));
  int rc;
  HP_HIP(hipEventRecord(e[0], s));
  if ((rc = batch_enqueue(b, HP_BATCH_PER_INSTANCE, s_hi, s_lo, l_hi, l_lo, b_hi, b_lo, x_hi, x_lo, status, s)))
    return rc;
  HP_HIP(hipEventRecord(e[1], s));
  if ((rc = batch_enqueue(b, HP_BATCH_PANEL_STEPS, s_hi, s_lo, l_hi, l_lo, b_hi, b_lo, x_hi, x_lo, status, s)))
    return rc;
  HP_HIP(hipEventRecord(e[2], s));
  HP_HIP(hipEventSynchronize(e[2]));
  float t1 = 0.f, t2 = 0.f;
  HP_HIP(hipEventElapsedTime(&t1, e[0], e[1]));
  HP_HIP(hipEventElapsedTime(&t2, e[1], e[2]));
  for (hipEvent_t& x : e) (void)hipEventDestroy(x);
  b->schedule = t2 < t1 ? HP_BATCH_PANEL_STEPS : HP_BATCH_PER_INSTANCE;
  std::lock_guard<std::mutex> lock(g_batch_mu);
  g_batch_tuned[key] = b->schedule;
  return HP_OK;
}
}  // namespace

extern "C" {

int hp_batch_set_schedule(int schedule) {
  if (schedule != 0 && schedule != HP_BATCH_PER_INSTANCE && schedule != HP_BATCH_PANEL_STEPS) {
    set_error("batch schedule must be 0 (autotune), HP_BATCH_PER_INSTANCE or HP_BATCH_PANEL_STEPS");
    return HP_EINVAL;
  }
  g_batch_schedule.store(schedule);
  return HP_OK;
}

int hp_dev_batch_cholesky_solve(const int* sizes, size_t count, int nrhs, const double* s_hi, const double* s_lo,
                                double* l_hi, double* l_lo, const double* b_hi, const double* b_lo, double* x_hi,
                                double* x_lo, int* status, void* stream) {
  if (count == 0) return HP_OK;
  hp_batch b;
  int rc = batch_prepare(&b, sizes, count, nrhs);
  if (rc) return rc;
  hipStream_t s = (hipStream_t)stream;
  if ((rc = batch_choose(&b, sizes, s_hi, s_lo, l_hi, l_lo, b_hi, b_lo, x_hi, x_lo, status, s))) return rc;
  if ((rc = batch_enqueue(&b, b.schedule, s_hi, s_lo, l_hi, l_lo, b_hi, b_lo, x_hi, x_lo, status, s))) return rc;
  HP_HIP(hipStreamSynchronize(s));  // the descriptors are freed on return
  return HP_OK;
}

hp_batch* hp_batch_graph_create(const int* sizes, size_t count, int nrhs, const double* s_hi, const double* s_lo,
                                double* l_hi, double* l_lo, const double* b_hi, const double* b_lo, double* x_hi,
                                double* x_lo, int* status, void* stream) {
  hp_batch* b = new hp_batch();
  hipStream_t s = (hipStream_t)stream;
  hipStream_t cap = nullptr;
  if (count == 0 || batch_prepare(b, sizes, count, nrhs) != HP_OK ||
      batch_choose(b, sizes, s_hi, s_lo, l_hi, l_lo, b_hi, b_lo, x_hi, x_lo, status, s) != HP_OK ||
      hipStreamSynchronize(s) != hipSuccess) {
    delete b;
    return nullptr;
  }
  if (hipStreamCreateWithFlags(&cap, hipStreamNonBlocking) != hipSuccess) { delete b; return nullptr; }
  bool ok = hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal) == hipSuccess;
  ok = ok && batch_enqueue(b, b->schedule, s_hi, s_lo, l_hi, l_lo, b_hi, b_lo, x_hi, x_lo, status, cap) == HP_OK;
  ok = (hipStreamEndCapture(cap, &b->graph) == hipSuccess) && ok;
  ok = ok && hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0) == hipSuccess;
  (void)hipStreamDestroy(cap);
  if (!ok) {
    set_error("batch graph capture failed");
    delete b;
    return nullptr;
  }
  return b;
}

int hp_batch_graph_launch(hp_batch* b, void* stream) {
  if (!b || !b->exec) {
    set_error("batch graph: null handle");
    return HP_EINVAL;
  }
  HP_HIP(hipGraphLaunch(b->exec, (hipStream_t)stream));
  return HP_OK;
}

int hp_batch_graph_schedule(const hp_batch* b) { return b ? b->schedule : 0; }

void hp_batch_graph_destroy(hp_batch* b) { delete b; }

}  // extern "C"
