// k_gemm_dd.hip -- plain double-double GEMM on FP64 VALU, bit-identical to the reference.
//
// Reference: crates/quad/src/linalg.rs:150-165 (gemm) and 168-182 (gemm_atb).  Each
// c[i,j] is the sequential sum  s = 0; for p in 0..k: s = s + a[i,p]*b[p,j]  with the dd
// add and mul of dd.rs; this kernel keeps exactly that per-element order (no split-K,
// no reassociation), so every output bit matches the reference.
//
// Layout: structure-of-arrays hi / lo planes, row-major.  A is m x k (or k x m for
// gemm_atb), B is k x n, C is m x n, each with its own leading dimension so that blocks
// of larger matrices can be operands.  Options used by the adjoint's triangular inverse
// and S^-1 = L^-T L^-1 products: C = -op(A) B (negate), a batch of equally strided
// products in one launch (grid z), and `tri`: A^T B with A, B lower triangular, where
// tiles above the diagonal are skipped and each tile's sum starts at p = its first row
// (the skipped terms are exact zeros, so per-element results are unchanged).
//
// Work decomposition (gfx950): a 256-thread workgroup owns a 64 x 64 tile of C; each
// thread owns a 4 x 4 register block (16 independent dd accumulation chains, which hides
// the ~13-deep FP64 dependency of a dd add).  K advances in 16-deep slabs staged through
// LDS as [k][m] and [k][n] planes so that a thread's 4 A values and 4 B values for one p
// are two 32-byte LDS reads per plane.  The next slab is fetched from HBM into registers
// while the current one is consumed.  Per p a thread issues 16 dd MACs = 464 FP64 VALU
// operations against 8 LDS reads, so the kernel is bound by FP64 VALU issue.
#include "hp_arith.hpp"
#include "hp_internal.hpp"

namespace hp {
namespace {

constexpr int BM = 64, BN = 64, BK = 16, NT = 256;

template <bool TRANS_A>
__device__ __forceinline__ void load_a_regs(const double* __restrict__ ahi, const double* __restrict__ alo,
                                            int m, int k, long lda, int i0, int p0, int t, double (&rh)[4],
                                            double (&rl)[4]) {
  if (!TRANS_A) {
    // A row-major m x k: thread t reads row i0 + t/4, columns p0 + (t%4)*4 .. +3
    int r = t >> 2, c = (t & 3) * 4;
    int gi = i0 + r;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      int gp = p0 + c + q;
      bool ok = gi < m && gp < k;
      size_t off = (size_t)gi * lda + gp;
      rh[q] = ok ? ahi[off] : 0.0;
      rl[q] = ok ? alo[off] : 0.0;
    }
  } else {
    // A stored k x m: thread t reads k-row p0 + t/16, columns i0 + (t%16)*4 .. +3
    int pr = t >> 4, c = (t & 15) * 4;
    int gp = p0 + pr;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      int gi = i0 + c + q;
      bool ok = gi < m && gp < k;
      size_t off = (size_t)gp * lda + gi;
      rh[q] = ok ? ahi[off] : 0.0;
      rl[q] = ok ? alo[off] : 0.0;
    }
  }
}

template <bool TRANS_A>
__device__ __forceinline__ void store_a_lds(double (*As)[BK][BM], int t, const double (&rh)[4], const double (&rl)[4]) {
  if (!TRANS_A) {
    int r = t >> 2, c = (t & 3) * 4;
#pragma unroll
    for (int q = 0; q < 4; ++q) { As[0][c + q][r] = rh[q]; As[1][c + q][r] = rl[q]; }
  } else {
    int pr = t >> 4, c = (t & 15) * 4;
#pragma unroll
    for (int q = 0; q < 4; ++q) { As[0][pr][c + q] = rh[q]; As[1][pr][c + q] = rl[q]; }
  }
}

__device__ __forceinline__ void load_b_regs(const double* __restrict__ bhi, const double* __restrict__ blo,
                                            int n, int k, long ldb, int j0, int p0, int t, double (&rh)[4],
                                            double (&rl)[4]) {
  int pr = t >> 4, c = (t & 15) * 4;
  int gp = p0 + pr;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    int gj = j0 + c + q;
    bool ok = gj < n && gp < k;
    size_t off = (size_t)gp * ldb + gj;
    rh[q] = ok ? bhi[off] : 0.0;
    rl[q] = ok ? blo[off] : 0.0;
  }
}

__device__ __forceinline__ void store_b_lds(double (*Bs)[BK][BN], int t, const double (&rh)[4], const double (&rl)[4]) {
  int pr = t >> 4, c = (t & 15) * 4;
#pragma unroll
  for (int q = 0; q < 4; ++q) { Bs[0][pr][c + q] = rh[q]; Bs[1][pr][c + q] = rl[q]; }
}

__device__ __forceinline__ void mac_step(double (*As)[BK][BM], double (*Bs)[BK][BN], int kk, int ty, int tx,
                                         dd (&acc)[4][4]) {
  dd a[4], b[4];
#pragma unroll
  for (int r = 0; r < 4; ++r) a[r] = mkdd(As[0][kk][ty * 4 + r], As[1][kk][ty * 4 + r]);
#pragma unroll
  for (int c = 0; c < 4; ++c) b[c] = mkdd(Bs[0][kk][tx * 4 + c], Bs[1][kk][tx * 4 + c]);
#pragma unroll
  for (int r = 0; r < 4; ++r)
#pragma unroll
    for (int c = 0; c < 4; ++c) acc[r][c] = dd_add(acc[r][c], dd_mul(a[r], b[c]));
}

struct GemmArgs {
  const double *ahi, *alo, *bhi, *blo;
  double *chi, *clo;
  int m, n, k;
  long lda, ldb, ldc;
  long sa, sb, sc;  // batch strides (elements) along grid z
  bool neg;         // C = -op(A) B
  bool tri;         // A^T B, A and B lower triangular: lower tiles only, p from the tile's first row
};

template <bool TRANS_A>
__global__ __launch_bounds__(NT) void k_gemm_dd_plain(GemmArgs g) {
  __shared__ double As[2][BK][BM];
  __shared__ double Bs[2][BK][BN];
  const int t = threadIdx.x;
  const int tx = t & 15, ty = t >> 4;
  const int i0 = blockIdx.y * BM, j0 = blockIdx.x * BN;
  if (g.tri && j0 >= i0 + BM) return;  // tile entirely above the diagonal
  const int m = g.m, n = g.n, k = g.k;
  const long z = blockIdx.z;
  const double* __restrict__ ahi = g.ahi + z * g.sa;
  const double* __restrict__ alo = g.alo + z * g.sa;
  const double* __restrict__ bhi = g.bhi + z * g.sb;
  const double* __restrict__ blo = g.blo + z * g.sb;

  dd acc[4][4];
#pragma unroll
  for (int r = 0; r < 4; ++r)
#pragma unroll
    for (int c = 0; c < 4; ++c) acc[r][c] = mkdd(0.0, 0.0);

  double ra_h[4], ra_l[4], rb_h[4], rb_l[4];
  // tri: A(p, i) = 0 for p < i, so every row of this tile has p >= i0 (BK divides BM)
  const int sl0 = g.tri ? i0 / BK : 0;
  const int nslab = (k + BK - 1) / BK;
  if (sl0 < nslab) {
    load_a_regs<TRANS_A>(ahi, alo, m, k, g.lda, i0, sl0 * BK, t, ra_h, ra_l);
    load_b_regs(bhi, blo, n, k, g.ldb, j0, sl0 * BK, t, rb_h, rb_l);
  }
  for (int sl = sl0; sl < nslab; ++sl) {
    store_a_lds<TRANS_A>(As, t, ra_h, ra_l);
    store_b_lds(Bs, t, rb_h, rb_l);
    __syncthreads();
    if (sl + 1 < nslab) {  // prefetch the next slab while this one is consumed
      load_a_regs<TRANS_A>(ahi, alo, m, k, g.lda, i0, (sl + 1) * BK, t, ra_h, ra_l);
      load_b_regs(bhi, blo, n, k, g.ldb, j0, (sl + 1) * BK, t, rb_h, rb_l);
    }
    const int kmax = min(BK, k - sl * BK);
    if (kmax == BK) {
#pragma unroll 2
      for (int kk = 0; kk < BK; ++kk) mac_step(As, Bs, kk, ty, tx, acc);
    } else {
      for (int kk = 0; kk < kmax; ++kk) mac_step(As, Bs, kk, ty, tx, acc);
    }
    __syncthreads();
  }
  double* __restrict__ chi = g.chi + z * g.sc;
  double* __restrict__ clo = g.clo + z * g.sc;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    int gi = i0 + ty * 4 + r;
    if (gi >= m) continue;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      int gj = j0 + tx * 4 + c;
      if (gj < n) {
        size_t off = (size_t)gi * g.ldc + gj;
        chi[off] = g.neg ? -acc[r][c].hi : acc[r][c].hi;
        clo[off] = g.neg ? -acc[r][c].lo : acc[r][c].lo;
      }
    }
  }
}

}  // namespace

int launch_gemm_dd_plain_ex(int trans_a, const double* ahi, const double* alo, long lda, const double* bhi,
                            const double* blo, long ldb, double* chi, double* clo, long ldc, size_t m, size_t n,
                            size_t k, bool neg, bool tri, int batch, long sa, long sb, long sc, hipStream_t s) {
  if (m == 0 || n == 0 || batch <= 0) return HP_OK;
  if (m > (1u << 30) || n > (1u << 30) || k > (1u << 30) || batch > 65535) {
    set_error("gemm: dimension too large");
    return HP_ESHAPE;
  }
  if (tri && (!trans_a || m != n)) {
    set_error("gemm: the triangular option needs A^T B with m == n");
    return HP_EINVAL;
  }
  GemmArgs g{ahi, alo, bhi, blo, chi, clo, (int)m, (int)n, (int)k, lda, ldb, ldc, sa, sb, sc, neg, tri};
  dim3 grid((unsigned)((n + BN - 1) / BN), (unsigned)((m + BM - 1) / BM), (unsigned)batch);
  if (trans_a)
    hipLaunchKernelGGL(k_gemm_dd_plain<true>, grid, dim3(NT), 0, s, g);
  else
    hipLaunchKernelGGL(k_gemm_dd_plain<false>, grid, dim3(NT), 0, s, g);
  HP_LAUNCH_CHECK();
  return HP_OK;
}

int launch_gemm_dd_plain(int trans_a, const double* ahi, const double* alo, const double* bhi, const double* blo,
                         double* chi, double* clo, size_t m, size_t n, size_t k, hipStream_t s) {
  return launch_gemm_dd_plain_ex(trans_a, ahi, alo, trans_a ? (long)m : (long)k, bhi, blo, (long)n, chi, clo,
                                 (long)n, m, n, k, false, false, 1, 0, 0, 0, s);
}

}  // namespace hp
