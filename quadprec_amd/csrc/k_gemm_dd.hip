// k_gemm_dd.hip -- plain double-double GEMM on FP64 VALU, bit-identical to the reference.
//
// Reference: crates/quad/src/linalg.rs:150-165 (gemm) and 168-182 (gemm_atb).  Each
// c[i,j] is the sequential sum  s = 0; for p in 0..k: s = s + a[i,p]*b[p,j]  with the dd
// add and mul of dd.rs; this kernel keeps exactly that per-element order (no split-K,
// no reassociation), so every output bit matches the reference.
//
// Layout: structure-of-arrays hi / lo planes, row-major.  A is m x k (or k x m for
// gemm_atb), B is k x n, C is m x n.
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
                                            int m, int k, int i0, int p0, int t, double (&rh)[4], double (&rl)[4]) {
  if (!TRANS_A) {
    // A row-major m x k: thread t reads row i0 + t/4, columns p0 + (t%4)*4 .. +3
    int r = t >> 2, c = (t & 3) * 4;
    int gi = i0 + r;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      int gp = p0 + c + q;
      bool ok = gi < m && gp < k;
      size_t off = (size_t)gi * k + gp;
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
      size_t off = (size_t)gp * m + gi;
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
                                            int n, int k, int j0, int p0, int t, double (&rh)[4], double (&rl)[4]) {
  int pr = t >> 4, c = (t & 15) * 4;
  int gp = p0 + pr;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    int gj = j0 + c + q;
    bool ok = gj < n && gp < k;
    size_t off = (size_t)gp * n + gj;
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

template <bool TRANS_A>
__global__ __launch_bounds__(NT) void k_gemm_dd_plain(const double* __restrict__ ahi, const double* __restrict__ alo,
                                                      const double* __restrict__ bhi, const double* __restrict__ blo,
                                                      double* __restrict__ chi, double* __restrict__ clo,
                                                      int m, int n, int k) {
  __shared__ double As[2][BK][BM];
  __shared__ double Bs[2][BK][BN];
  const int t = threadIdx.x;
  const int tx = t & 15, ty = t >> 4;
  const int i0 = blockIdx.y * BM, j0 = blockIdx.x * BN;

  dd acc[4][4];
#pragma unroll
  for (int r = 0; r < 4; ++r)
#pragma unroll
    for (int c = 0; c < 4; ++c) acc[r][c] = mkdd(0.0, 0.0);

  double ra_h[4], ra_l[4], rb_h[4], rb_l[4];
  const int nslab = (k + BK - 1) / BK;
  if (nslab > 0) {
    load_a_regs<TRANS_A>(ahi, alo, m, k, i0, 0, t, ra_h, ra_l);
    load_b_regs(bhi, blo, n, k, j0, 0, t, rb_h, rb_l);
  }
  for (int sl = 0; sl < nslab; ++sl) {
    store_a_lds<TRANS_A>(As, t, ra_h, ra_l);
    store_b_lds(Bs, t, rb_h, rb_l);
    __syncthreads();
    if (sl + 1 < nslab) {  // prefetch the next slab while this one is consumed
      load_a_regs<TRANS_A>(ahi, alo, m, k, i0, (sl + 1) * BK, t, ra_h, ra_l);
      load_b_regs(bhi, blo, n, k, j0, (sl + 1) * BK, t, rb_h, rb_l);
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
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    int gi = i0 + ty * 4 + r;
    if (gi >= m) continue;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      int gj = j0 + tx * 4 + c;
      if (gj < n) {
        size_t off = (size_t)gi * n + gj;
        chi[off] = acc[r][c].hi;
        clo[off] = acc[r][c].lo;
      }
    }
  }
}

}  // namespace

int launch_gemm_dd_plain(int trans_a, const double* ahi, const double* alo, const double* bhi, const double* blo,
                         double* chi, double* clo, size_t m, size_t n, size_t k, hipStream_t s) {
  if (m == 0 || n == 0) return HP_OK;
  if (m > (1u << 30) || n > (1u << 30) || k > (1u << 30)) {
    set_error("gemm: dimension too large");
    return HP_ESHAPE;
  }
  dim3 grid((unsigned)((n + BN - 1) / BN), (unsigned)((m + BM - 1) / BM));
  if (trans_a)
    hipLaunchKernelGGL(k_gemm_dd_plain<true>, grid, dim3(NT), 0, s, ahi, alo, bhi, blo, chi, clo, (int)m, (int)n, (int)k);
  else
    hipLaunchKernelGGL(k_gemm_dd_plain<false>, grid, dim3(NT), 0, s, ahi, alo, bhi, blo, chi, clo, (int)m, (int)n, (int)k);
  HP_LAUNCH_CHECK();
  return HP_OK;
}

}  // namespace hp
