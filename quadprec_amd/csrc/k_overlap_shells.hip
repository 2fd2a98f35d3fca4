// k_overlap_shells.hip -- assemble_overlap_mixed over contracted s/p/d/f shells
// (crates/overlap/src/assemble.rs:35-85, screen.rs:57-101, integral.rs:65-346).
//
// Three launches: every shell pair (i >= j) is screened once and listed by (tier, l_i, l_j);
// one thread per (listed pair, row u of its block) forms the row of v components in the
// tier's arithmetic, templated on l_i, l_j and the tier so that no wavefront diverges; the
// upper triangle of distinct-shell blocks is then copied from the lower through LDS tiles.
// Per component the operations are the reference's, in its order: primitive pairs p outer,
// q inner, the reference's set-up and 1D Obara-Saika recurrences, the same expression for
// each term.  libm exp (glibc on the host, the device library here) is the one operation
// whose last bit may differ.
#include <cmath>
#include <vector>

#include "hp_arith.hpp"
#include "hp_internal.hpp"

namespace hp {
namespace {

constexpr double PI_F64 = 3.141592653589793;

struct ShellRec {
  double c[3];
  double amin, amax;  // alpha_min / alpha_max over the contraction (basis.rs:40-48)
  long p0, p1;        // primitives
  int l, off;         // angular momentum, first AO
};

__device__ __forceinline__ double dist2(const double* a, const double* b) {
  const double dx = a[0] - b[0], dy = a[1] - b[1], dz = a[2] - b[2];
  return dx * dx + dy * dy + dz * dz;
}

// screen_shell_pair: screen.rs:57-101 (0 F64, 1 F128, 2 Skip); prm = {diffuse, ratio,
// -ln(eps_skip), safe_distance}
__device__ int screen(const ShellRec& a, const ShellRec& b, const double* prm) {
  const double r2 = dist2(a.c, b.c);
  const double alpha_sum = a.amin + b.amin;
  const double gamma = a.amin * b.amin / alpha_sum;
  if (gamma * r2 > prm[2]) return 2;
  const bool ad = a.amin < prm[0], bd = b.amin < prm[0];
  if (!ad && !bd) return 0;
  const double ratio = fmax(a.amax / b.amin, b.amax / a.amin);
  const double r = sqrt(r2);
  if (ratio > prm[1] && r < prm[3]) return 1;
  if (ad && bd && r < prm[3]) return 1;
  return 0;
}

// cart_components (integral.rs:220-231): the u-th (lx, ly, lz) of angular momentum l
__device__ __forceinline__ void cart(int l, int u, int& lx, int& ly, int& lz) {
  int n = 0;
  lx = ly = lz = 0;
#pragma unroll
  for (int x = 3; x >= 0; --x)
#pragma unroll
    for (int y = 3; y >= 0; --y) {
      if (x > l || y > l - x) continue;
      if (n == u) {
        lx = x;
        ly = y;
        lz = l - x - y;
      }
      ++n;
    }
}

// overlap_primitive_f64 (integral.rs:15-25)
__device__ double prim_f64(double ai, const double* ci, double aj, const double* cj) {
  const double gamma = ai + aj;
  const double mu = ai * aj / gamma;
  const double r2 = dist2(ci, cj);
  const double sq = sqrt(PI_F64 / gamma);
  const double prefactor = (sq * sq) * sq;  // powi(3)
  return prefactor * exp(-mu * r2);
}
// overlap_primitive_f128 (integral.rs:29-56): f64 pi, f64 centre differences, f64 exp
__device__ dd prim_dd(double ai, const double* ci, double aj, const double* cj) {
  const dd a = mkdd(ai), b = mkdd(aj);
  const dd gamma = dd_add(a, b);
  const dd mu = dd_div(dd_mul(a, b), gamma);
  const dd dx = mkdd(ci[0] - cj[0]), dy = mkdd(ci[1] - cj[1]), dz = mkdd(ci[2] - cj[2]);
  const dd r2 = dd_add(dd_add(dd_mul(dx, dx), dd_mul(dy, dy)), dd_mul(dz, dz));
  const dd ratio = dd_div(mkdd(PI_F64), gamma);
  const dd prefactor = dd_mul(ratio, dd_sqrt(ratio));
  const dd nmr2 = dd_mul(dd_sub(mkdd(0.0), mu), r2);
  const dd exp_val = dd_mul(mkdd(exp(nmr2.hi)), dd_add(mkdd(1.0), mkdd(nmr2.lo)));
  return dd_mul(prefactor, exp_val);
}

// ---- assembly in three launches:
//   k_classify: every shell pair (i >= j) screened once (screen.rs:57-101), counted by tier
//     and listed by (tier, l_i, l_j) (counts gathered in LDS per block of 4096 pairs), so
//     that the next launches run without divergence;
//   k_rows<LA, LB, TIER>: one thread per (pair, row u of the block): per primitive pair the
//     reference's set-up and its three 1D Obara-Saika tables, then the whole row of v
//     components (integral.rs:65-346); every component accumulates over the primitive
//     pairs p outer, q inner, with the reference's expression;
//   k_mirror: the upper triangle of distinct-shell blocks copied from the lower through
//     64 x 64 LDS tiles (assemble.rs:74-78 writes the mirror; a shell's own block is
//     computed whole, as in the reference).
constexpr int NBUCKET = 32;  // tier (0 f64, 1 dd) x l of the row shell x l of the column shell

__device__ __forceinline__ void decode_pair(long k, long& i, long& j) {
  i = (long)((sqrt(8.0 * (double)k + 1.0) - 1.0) * 0.5);
  while (i * (i + 1) / 2 > k) --i;
  while ((i + 1) * (i + 2) / 2 <= k) ++i;
  j = k - i * (i + 1) / 2;
}

// A block takes 4096 pairs (4 per thread): counts per tier and per bucket gather in LDS,
// then one global atomic per nonempty bucket (and tier) per block; in the fill pass each
// pair is listed at its bucket's start + the block's base + its rank within the block.
constexpr int CL_PER = 4, CL_THREADS = 1024;
template <bool FILL>
__global__ __launch_bounds__(CL_THREADS) void k_classify(const ShellRec* __restrict__ sh, long npairs,
                                                         const double* __restrict__ prm,
                                                         unsigned long long* __restrict__ tiers,
                                                         unsigned int* __restrict__ bcount,
                                                         const unsigned int* __restrict__ bstart,
                                                         int2* __restrict__ list) {
  __shared__ unsigned int cnt[NBUCKET], base[NBUCKET];
  __shared__ unsigned int tcnt[3];
  const int t = threadIdx.x;
  if (t < NBUCKET) cnt[t] = 0;
  if (t < 3) tcnt[t] = 0;
  __syncthreads();
  const long k0 = (long)blockIdx.x * CL_PER * CL_THREADS;
  int bk[CL_PER];
  unsigned int rank[CL_PER];
  int pi[CL_PER], pj[CL_PER];
#pragma unroll
  for (int e = 0; e < CL_PER; ++e) {
    bk[e] = -1;
    rank[e] = 0;
    pi[e] = pj[e] = 0;
    const long k = k0 + (long)e * CL_THREADS + t;
    if (k < npairs) {
      long i, j;
      decode_pair(k, i, j);
      const ShellRec A = sh[i], B = sh[j];
      const int tier = screen(A, B, prm);
      if (!FILL) atomicAdd(&tcnt[tier], 1u);
      if (tier < 2) {
        bk[e] = tier * 16 + A.l * 4 + B.l;
        rank[e] = atomicAdd(&cnt[bk[e]], 1u);
        pi[e] = (int)i;
        pj[e] = (int)j;
      }
    }
  }
  __syncthreads();
  if (t < NBUCKET && cnt[t]) base[t] = atomicAdd(&bcount[t], cnt[t]);
  if (!FILL && t < 3 && tcnt[t]) atomicAdd(&tiers[t], (unsigned long long)tcnt[t]);
  if (!FILL) return;
  __syncthreads();
#pragma unroll
  for (int e = 0; e < CL_PER; ++e)
    if (bk[e] >= 0) list[bstart[bk[e]] + base[bk[e]] + rank[e]] = make_int2(pi[e], pj[e]);
}

// Row ei (runtime, <= LA) of the 1D Obara-Saika table s[i][j], i <= LA, j <= LB, of
// obara_saika_1d_f64 / _f128 (integral.rs:158-216), with their expressions.  Entry (i, j)
// uses only (i, j - 1), (i - 1, j) and, in column 0, (i - 1, 0), (i - 2, 0), so building
// the table row by row gives every entry the value the reference's column order gives;
// only the previous row is kept, and the wanted row is picked as a whole.
template <int LA, int LB>
__device__ __forceinline__ void row_f64(double pa, double pb, double gamma, int ei, double (&out)[LB + 1]) {
  const double oo2g = 0.5 / gamma;
  double pm[LB + 1], cur[LB + 1], pm2_0 = 0.0;
#pragma unroll
  for (int j = 0; j <= LB; ++j) pm[j] = 0.0;
#pragma unroll
  for (int i = 0; i <= LA; ++i) {
    cur[0] = i == 0 ? 1.0 : pa * pm[0] + oo2g * (double)(i - 1) * (i - 1 > 0 ? pm2_0 : 0.0);
#pragma unroll
    for (int j = 0; j < LB; ++j) {
      const double tim1 = oo2g * (double)i * (i > 0 ? pm[j] : 0.0);
      const double tjm1 = oo2g * (double)j * (j > 0 ? cur[j > 0 ? j - 1 : 0] : 0.0);
      cur[j + 1] = pb * cur[j] + tim1 + tjm1;
    }
#pragma unroll
    for (int j = 0; j <= LB; ++j) out[j] = (i == 0 || ei == i) ? cur[j] : out[j];
    pm2_0 = pm[0];
#pragma unroll
    for (int j = 0; j <= LB; ++j) pm[j] = cur[j];
  }
}
template <int LA, int LB>
__device__ __forceinline__ void row_dd(dd pa, dd pb, dd gamma, int ei, dd (&out)[LB + 1]) {
  const dd oo2g = dd_div(mkdd(0.5), gamma);
  dd pm[LB + 1], cur[LB + 1], pm2_0 = mkdd(0.0);
#pragma unroll
  for (int j = 0; j <= LB; ++j) pm[j] = mkdd(0.0);
#pragma unroll
  for (int i = 0; i <= LA; ++i) {
    if (i == 0) {
      cur[0] = mkdd(1.0);
    } else {
      const dd prev = i - 1 > 0 ? pm2_0 : mkdd(0.0);
      cur[0] = dd_add(dd_mul(pa, pm[0]), dd_mul(dd_mul(oo2g, mkdd((double)(i - 1))), prev));
    }
    const dd fi = mkdd((double)i);
#pragma unroll
    for (int j = 0; j < LB; ++j) {
      const dd fj = mkdd((double)j);
      const dd im1 = i > 0 ? pm[j] : mkdd(0.0);
      const dd jm1 = j > 0 ? cur[j > 0 ? j - 1 : 0] : mkdd(0.0);
      cur[j + 1] = dd_add(dd_add(dd_mul(pb, cur[j]), dd_mul(dd_mul(oo2g, fi), im1)), dd_mul(dd_mul(oo2g, fj), jm1));
    }
    const bool take = i == 0 || ei == i;
#pragma unroll
    for (int j = 0; j <= LB; ++j) {
      out[j].hi = take ? cur[j].hi : out[j].hi;
      out[j].lo = take ? cur[j].lo : out[j].lo;
    }
    pm2_0 = pm[0];
#pragma unroll
    for (int j = 0; j <= LB; ++j) pm[j] = cur[j];
  }
}
// the lx of component v of angular momentum L, in cart_components order (compile time)
__host__ __device__ constexpr int cart_x(int L, int v) {
  int n = 0;
  for (int x = L; x >= 0; --x)
    for (int y = L - x; y >= 0; --y) {
      if (n == v) return x;
      ++n;
    }
  return 0;
}
__host__ __device__ constexpr int cart_y(int L, int v) {
  int n = 0;
  for (int x = L; x >= 0; --x)
    for (int y = L - x; y >= 0; --y) {
      if (n == v) return y;
      ++n;
    }
  return 0;
}

// thread t of bucket (LA, LB, TIER): row u = t mod NU of listed pair t / NU
template <int LA, int LB, int TIER>
__device__ __forceinline__ void rows_body(const ShellRec* __restrict__ sh, const int2* __restrict__ list, long cnt,
                                          const double* __restrict__ alpha, const double* __restrict__ coeff,
                                          long n_ao, double* __restrict__ s, long t) {
  constexpr int NU = (LA + 1) * (LA + 2) / 2, NV = (LB + 1) * (LB + 2) / 2;
  if (t >= cnt * NU) return;
  const long k = t / NU;
  const int u = (int)(t - k * NU);
  const int2 pr = list[k];
  const ShellRec A = sh[pr.x], B = sh[pr.y];
  double* row = s + (long)(A.off + u) * n_ao + B.off;
  if (TIER == 0) {
    if (LA == 0 && LB == 0) {  // integral.rs:99-110
      double sum = 0.0;
      for (long p = A.p0; p < A.p1; ++p)
        for (long q = B.p0; q < B.p1; ++q) sum += coeff[p] * coeff[q] * prim_f64(alpha[p], A.c, alpha[q], B.c);
      row[0] = sum;
      return;
    }
    int ux, uy, uz;
    cart(LA, u, ux, uy, uz);
    double acc[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) acc[v] = 0.0;
    for (long p = A.p0; p < A.p1; ++p)  // integral.rs:233-282
      for (long q = B.p0; q < B.p1; ++q) {
        const double ai = alpha[p], aj = alpha[q];
        const double gamma = ai + aj;
        const double mu = ai * aj / gamma;
        const double r2 = dist2(A.c, B.c);
        const double sq = sqrt(PI_F64 / gamma);
        const double prefactor = (sq * sq) * sq * exp(-mu * r2);
        double pa[3], pb[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          const double pk = (ai * A.c[d] + aj * B.c[d]) / gamma;
          pa[d] = pk - A.c[d];
          pb[d] = pk - B.c[d];
        }
        double rx[LB + 1], ry[LB + 1], rz[LB + 1];
        row_f64<LA, LB>(pa[0], pb[0], gamma, ux, rx);
        row_f64<LA, LB>(pa[1], pb[1], gamma, uy, ry);
        row_f64<LA, LB>(pa[2], pb[2], gamma, uz, rz);
        const double cp = coeff[p], cq = coeff[q];
#pragma unroll
        for (int v = 0; v < NV; ++v) {
          const int vx = cart_x(LB, v), vy = cart_y(LB, v), vz = LB - vx - vy;
          acc[v] += rx[vx] * ry[vy] * rz[vz] * prefactor * cp * cq;
        }
      }
#pragma unroll
    for (int v = 0; v < NV; ++v) row[v] = acc[v];
    return;
  }
  if (LA == 0 && LB == 0) {  // integral.rs:119-136
    dd sum = mkdd(0.0);
    for (long p = A.p0; p < A.p1; ++p)
      for (long q = B.p0; q < B.p1; ++q)
        sum = dd_add(sum, dd_mul(dd_mul(mkdd(coeff[p]), mkdd(coeff[q])), prim_dd(alpha[p], A.c, alpha[q], B.c)));
    row[0] = sum.hi + sum.lo;
    return;
  }
  int ux, uy, uz;
  cart(LA, u, ux, uy, uz);
  dd acc[NV];
#pragma unroll
  for (int v = 0; v < NV; ++v) acc[v] = mkdd(0.0);
  for (long p = A.p0; p < A.p1; ++p)  // integral.rs:284-346
    for (long q = B.p0; q < B.p1; ++q) {
      const dd ai = mkdd(alpha[p]), aj = mkdd(alpha[q]);
      const dd gamma = dd_add(ai, aj);
      const dd mu = dd_div(dd_mul(ai, aj), gamma);
      const dd dx = mkdd(A.c[0] - B.c[0]), dy = mkdd(A.c[1] - B.c[1]), dz = mkdd(A.c[2] - B.c[2]);
      const dd r2 = dd_add(dd_add(dd_mul(dx, dx), dd_mul(dy, dy)), dd_mul(dz, dz));
      const dd ratio = dd_div(mkdd(PI_F64), gamma);
      const dd prefactor = dd_mul(ratio, dd_sqrt(ratio));
      const dd nmr2 = dd_mul(dd_sub(mkdd(0.0), mu), r2);
      const dd exp_val = dd_mul(mkdd(exp(nmr2.hi)), dd_add(mkdd(1.0), mkdd(nmr2.lo)));
      const dd total_pre = dd_mul(prefactor, exp_val);
      dd pa[3], pb[3];
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        const dd cik = mkdd(A.c[d]), cjk = mkdd(B.c[d]);
        const dd pk = dd_div(dd_add(dd_mul(ai, cik), dd_mul(aj, cjk)), gamma);
        pa[d] = dd_sub(pk, cik);
        pb[d] = dd_sub(pk, cjk);
      }
      dd rx[LB + 1], ry[LB + 1], rz[LB + 1];
      row_dd<LA, LB>(pa[0], pb[0], gamma, ux, rx);
      row_dd<LA, LB>(pa[1], pb[1], gamma, uy, ry);
      row_dd<LA, LB>(pa[2], pb[2], gamma, uz, rz);
      const dd cf = dd_mul(dd_mul(mkdd(coeff[p]), mkdd(coeff[q])), total_pre);
#pragma unroll
      for (int v = 0; v < NV; ++v) {
        const int vx = cart_x(LB, v), vy = cart_y(LB, v), vz = LB - vx - vy;
        acc[v] = dd_add(acc[v], dd_mul(dd_mul(dd_mul(rx[vx], ry[vy]), rz[vz]), cf));
      }
    }
#pragma unroll
  for (int v = 0; v < NV; ++v) row[v] = acc[v].hi + acc[v].lo;
}

// tile pair p of the T(T+1)/2 pairs (ta <= tb) of a T x T tiling, row by row of ta
__device__ __forceinline__ void tile_pair(long p, long T, long& ta, long& tb) {
  // row ta holds the T - ta pairs from index ta T - ta (ta - 1) / 2
  long r = (long)((2.0 * T + 1.0 - sqrt((2.0 * T + 1.0) * (2.0 * T + 1.0) - 8.0 * (double)p)) * 0.5);
  auto first = [&](long x) { return x * T - x * (x - 1) / 2; };
  while (r > 0 && first(r) > p) --r;
  while (r + 1 < T && first(r + 1) <= p) ++r;
  ta = r;
  tb = r + (p - first(r));
}

// S[a][b] = S[b][a] for a < b in distinct shells; block p takes the tile pair (ta, tb),
// ta <= tb: it reads the lower tile (tb, ta) into LDS and writes the upper tile (ta, tb),
// so each tile is read or written by one block only
__global__ __launch_bounds__(256) void k_mirror(double* __restrict__ s, const int* __restrict__ ao_shell, long n) {
  __shared__ double tile[64][65];
  __shared__ int sha[64], shb[64];  // shells of the upper tile's rows / columns
  const long T = (n + 63) / 64;
  long ta, tb;
  tile_pair(blockIdx.x, T, ta, tb);
  if (threadIdx.x < 128) {
    const long g = (threadIdx.x < 64 ? ta : tb) * 64 + (threadIdx.x & 63);
    (threadIdx.x < 64 ? sha : shb)[threadIdx.x & 63] = g < n ? ao_shell[g] : -1;
  }
#pragma unroll 4
  for (int idx = threadIdx.x; idx < 64 * 64; idx += 256) {
    const int r = idx >> 6, c = idx & 63;
    const long row = tb * 64 + r, col = ta * 64 + c;
    tile[r][c] = (row < n && col < n) ? s[row * n + col] : 0.0;
  }
  __syncthreads();
#pragma unroll 4
  for (int idx = threadIdx.x; idx < 64 * 64; idx += 256) {
    const int r = idx >> 6, c = idx & 63;
    const long a = ta * 64 + r, b = tb * 64 + c;
    if (a < n && b < n && a < b && sha[r] != shb[c]) s[a * n + b] = tile[c][r];
  }
}

// zeroes the tiles on and below the diagonal, where the pairs the screen skips must read
// zero (k_mirror then writes every upper entry outside the shells' own blocks, which
// k_rows computes whole): block p takes the tile pair (ta, tb), ta <= tb, and zeroes
// tile (tb, ta)
__global__ __launch_bounds__(256) void k_zero_lower(double* __restrict__ s, long n) {
  const long T = (n + 63) / 64;
  long ta, tb;
  tile_pair(blockIdx.x, T, ta, tb);
#pragma unroll 4
  for (int idx = threadIdx.x; idx < 64 * 64; idx += 256) {
    const long row = tb * 64 + (idx >> 6), col = ta * 64 + (idx & 63);
    if (row < n && col < n) s[row * n + col] = 0.0;
  }
}

// One launch per tier covers all its (l_i, l_j) buckets: blocks [first[e], first[e + 1])
// run bucket code[e] (4 l_i + l_j) over its count[e] pairs listed from start[e].  The host
// orders the buckets by descending per-thread work so that the longest blocks start first.
struct RowsTab {
  unsigned int first[17], start[16], count[16];
  int code[16], n;
};
constexpr int ROWS_THREADS = 128;

template <int TIER>
__global__ __launch_bounds__(ROWS_THREADS) void k_rows(const ShellRec* __restrict__ sh, const int2* __restrict__ list,
                                                       RowsTab tab, const double* __restrict__ alpha,
                                                       const double* __restrict__ coeff, long n_ao,
                                                       double* __restrict__ s) {
  const unsigned int blk = blockIdx.x;
  int e = 0;
  while (e + 1 < tab.n && blk >= tab.first[e + 1]) ++e;
  const long t = (long)(blk - tab.first[e]) * ROWS_THREADS + threadIdx.x;
  const int2* lst = list + tab.start[e];
  const long cnt = tab.count[e];
#define HP_ROWS_CASE(LA, LB) \
  case LA * 4 + LB: rows_body<LA, LB, TIER>(sh, lst, cnt, alpha, coeff, n_ao, s, t); break;
  switch (tab.code[e]) {
    HP_ROWS_CASE(0, 0) HP_ROWS_CASE(0, 1) HP_ROWS_CASE(0, 2) HP_ROWS_CASE(0, 3)
    HP_ROWS_CASE(1, 0) HP_ROWS_CASE(1, 1) HP_ROWS_CASE(1, 2) HP_ROWS_CASE(1, 3)
    HP_ROWS_CASE(2, 0) HP_ROWS_CASE(2, 1) HP_ROWS_CASE(2, 2) HP_ROWS_CASE(2, 3)
    HP_ROWS_CASE(3, 0) HP_ROWS_CASE(3, 1) HP_ROWS_CASE(3, 2) HP_ROWS_CASE(3, 3)
    default: break;
  }
#undef HP_ROWS_CASE
}

// the nonempty buckets of one tier, heaviest per thread first (larger l_i + l_j, then l_i)
RowsTab rows_table(int tier, const unsigned int* hcount, const unsigned int* hstart) {
  RowsTab tab{};
  int order[16];
  int n = 0;
  for (int c = 0; c < 16; ++c)
    if (hcount[tier * 16 + c]) order[n++] = c;
  for (int i = 1; i < n; ++i)
    for (int j = i; j > 0; --j) {
      const int x = order[j - 1], y = order[j];
      const int kx = ((x >> 2) + (x & 3)) * 4 + (x >> 2), ky = ((y >> 2) + (y & 3)) * 4 + (y >> 2);
      if (ky <= kx) break;
      order[j - 1] = y;
      order[j] = x;
    }
  unsigned int blocks = 0;
  for (int e = 0; e < n; ++e) {
    const int c = order[e], la = c >> 2;
    const unsigned int cnt = hcount[tier * 16 + c];
    const unsigned long long threads = (unsigned long long)cnt * ((la + 1) * (la + 2) / 2);
    tab.first[e] = blocks;
    tab.start[e] = hstart[tier * 16 + c];
    tab.count[e] = cnt;
    tab.code[e] = c;
    blocks += (unsigned int)((threads + ROWS_THREADS - 1) / ROWS_THREADS);
  }
  tab.first[n] = blocks;
  tab.n = n;
  return tab;
}

}  // namespace
}  // namespace hp

using namespace hp;

namespace {
// s_dev: the caller's device matrix (n_ao x n_ao), or null to assemble into a scratch
// buffer and copy the result to s_host
int assemble_mixed(const double* centers, const int* l, const int64_t* prim_off, const double* alpha,
                   const double* coeff, size_t n_shells, const double* params4, double* s_host, double* s_dev,
                   size_t n_ao, size_t* counts, hipStream_t st) {
  if (!counts || !params4) {
    set_error("null pointer argument");
    return HP_EINVAL;
  }
  counts[0] = counts[1] = counts[2] = 0;
  if (n_shells == 0) return n_ao == 0 ? HP_OK : HP_ESHAPE;
  if (!centers || !l || !prim_off || !alpha || !coeff || (!s_host && !s_dev)) {
    set_error("null pointer argument");
    return HP_EINVAL;
  }
  std::vector<ShellRec> rec(n_shells);
  std::vector<int> ao_shell;
  long off = 0;
  for (size_t i = 0; i < n_shells; ++i) {
    if (l[i] < 0 || l[i] > 3 || prim_off[i + 1] < prim_off[i]) {
      set_error("assemble_overlap_mixed: shell %zu: l must be 0..3 (S, P, D, F) and primitives non-negative", i);
      return HP_EINVAL;
    }
    ShellRec& r = rec[i];
    for (int k = 0; k < 3; ++k) r.c[k] = centers[3 * i + k];
    r.amin = INFINITY;
    r.amax = 0.0;
    for (int64_t p = prim_off[i]; p < prim_off[i + 1]; ++p) {
      r.amin = std::fmin(r.amin, alpha[p]);
      r.amax = std::fmax(r.amax, alpha[p]);
    }
    r.p0 = prim_off[i];
    r.p1 = prim_off[i + 1];
    r.l = l[i];
    r.off = (int)off;
    const int nf = (l[i] + 1) * (l[i] + 2) / 2;
    for (int k = 0; k < nf; ++k) ao_shell.push_back((int)i);
    off += nf;
  }
  if ((size_t)off != n_ao) {
    set_error("assemble_overlap_mixed: n_ao = %zu, the shells give %ld", n_ao, off);
    return HP_ESHAPE;
  }
  const long np = (long)prim_off[n_shells];
  const double prm[4] = {params4[0], params4[1], -std::log(params4[2]), params4[3]};
  DevBuf dsh, dao, dal, dco, dprm, ds, dcnt;
  int rc;
  if ((rc = dsh.alloc(n_shells * sizeof(ShellRec))) || (rc = dao.alloc(n_ao * sizeof(int))) ||
      (rc = dal.alloc(np * sizeof(double))) || (rc = dco.alloc(np * sizeof(double))) || (rc = dprm.alloc(32)) ||
      (rc = dcnt.alloc(3 * sizeof(unsigned long long))))
    return rc;
  if (!s_dev) {
    if ((rc = ds.alloc(n_ao * n_ao * sizeof(double)))) return rc;
    s_dev = ds.d();
  }
  HP_HIP(hipMemcpyAsync(dsh.p, rec.data(), n_shells * sizeof(ShellRec), hipMemcpyHostToDevice, st));
  HP_HIP(hipMemcpyAsync(dao.p, ao_shell.data(), n_ao * sizeof(int), hipMemcpyHostToDevice, st));
  if (np > 0) {
    HP_HIP(hipMemcpyAsync(dal.p, alpha, np * sizeof(double), hipMemcpyHostToDevice, st));
    HP_HIP(hipMemcpyAsync(dco.p, coeff, np * sizeof(double), hipMemcpyHostToDevice, st));
  }
  HP_HIP(hipMemcpyAsync(dprm.p, prm, sizeof(prm), hipMemcpyHostToDevice, st));
  const long tiles = (long)((n_ao + 63) / 64), tpairs = tiles * (tiles + 1) / 2;
  hipLaunchKernelGGL(k_zero_lower, dim3((unsigned)tpairs), dim3(256), 0, st, s_dev, (long)n_ao);
  HP_LAUNCH_CHECK();
  const long npairs = (long)n_shells * ((long)n_shells + 1) / 2;
  DevBuf dbk, dlist;
  if ((rc = dbk.alloc(2 * NBUCKET * sizeof(unsigned int))) || (rc = dlist.alloc(npairs * sizeof(int2)))) return rc;
  unsigned int* bcount = static_cast<unsigned int*>(dbk.p);
  unsigned int* bstart = bcount + NBUCKET;
  auto* tiers = static_cast<unsigned long long*>(dcnt.p);
  HP_HIP(hipMemsetAsync(tiers, 0, 3 * sizeof(unsigned long long), st));
  HP_HIP(hipMemsetAsync(bcount, 0, NBUCKET * sizeof(unsigned int), st));
  const unsigned cgrid = (unsigned)((npairs + CL_PER * CL_THREADS - 1) / (CL_PER * CL_THREADS));
  const ShellRec* dshp = static_cast<const ShellRec*>(dsh.p);
  int2* list = static_cast<int2*>(dlist.p);
  hipLaunchKernelGGL(k_classify<false>, dim3(cgrid), dim3(CL_THREADS), 0, st, dshp, npairs, dprm.d(), tiers, bcount,
                     (const unsigned int*)nullptr, list);
  HP_LAUNCH_CHECK();
  unsigned int hcount[NBUCKET], hstart[NBUCKET];
  HP_HIP(hipMemcpyAsync(hcount, bcount, sizeof(hcount), hipMemcpyDeviceToHost, st));
  HP_HIP(hipStreamSynchronize(st));
  unsigned int acc = 0;
  for (int b = 0; b < NBUCKET; ++b) {
    hstart[b] = acc;
    acc += hcount[b];
  }
  HP_HIP(hipMemcpyAsync(bstart, hstart, sizeof(hstart), hipMemcpyHostToDevice, st));
  HP_HIP(hipMemsetAsync(bcount, 0, NBUCKET * sizeof(unsigned int), st));  // cursors of the fill pass
  hipLaunchKernelGGL(k_classify<true>, dim3(cgrid), dim3(CL_THREADS), 0, st, dshp, npairs, dprm.d(),
                     (unsigned long long*)nullptr, bcount, (const unsigned int*)bstart, list);
  HP_LAUNCH_CHECK();
  for (int tier = 1; tier >= 0; --tier) {
    const RowsTab tab = rows_table(tier, hcount, hstart);
    if (!tab.n) continue;
    if (tier)
      hipLaunchKernelGGL(k_rows<1>, dim3(tab.first[tab.n]), dim3(ROWS_THREADS), 0, st, dshp, list, tab, dal.d(),
                         dco.d(), (long)n_ao, s_dev);
    else
      hipLaunchKernelGGL(k_rows<0>, dim3(tab.first[tab.n]), dim3(ROWS_THREADS), 0, st, dshp, list, tab, dal.d(),
                         dco.d(), (long)n_ao, s_dev);
    HP_LAUNCH_CHECK();
  }
  hipLaunchKernelGGL(k_mirror, dim3((unsigned)tpairs), dim3(256), 0, st, s_dev, static_cast<const int*>(dao.p),
                     (long)n_ao);
  HP_LAUNCH_CHECK();
  unsigned long long hc[3];
  if (s_host) HP_HIP(hipMemcpyAsync(s_host, s_dev, n_ao * n_ao * sizeof(double), hipMemcpyDeviceToHost, st));
  HP_HIP(hipMemcpyAsync(hc, tiers, sizeof(hc), hipMemcpyDeviceToHost, st));
  HP_HIP(hipStreamSynchronize(st));
  for (int k = 0; k < 3; ++k) counts[k] = (size_t)hc[k];
  return HP_OK;
}
}  // namespace

extern "C" {

int hp_assemble_overlap_mixed(const double* centers, const int* l, const int64_t* prim_off, const double* alpha,
                              const double* coeff, size_t n_shells, const double* params4, double* s, size_t n_ao,
                              size_t* counts) {
  if (n_shells && !s) {
    set_error("null pointer argument");
    return HP_EINVAL;
  }
  return assemble_mixed(centers, l, prim_off, alpha, coeff, n_shells, params4, s, nullptr, n_ao, counts,
                        hipStreamPerThread);
}

int hp_dev_assemble_overlap_mixed(const double* centers, const int* l, const int64_t* prim_off, const double* alpha,
                                  const double* coeff, size_t n_shells, const double* params4, double* s_dev,
                                  size_t n_ao, size_t* counts, void* stream) {
  if (n_shells && !s_dev) {
    set_error("null pointer argument");
    return HP_EINVAL;
  }
  return assemble_mixed(centers, l, prim_off, alpha, coeff, n_shells, params4, nullptr, s_dev, n_ao, counts,
                        (hipStream_t)stream);
}

}  // extern "C"
