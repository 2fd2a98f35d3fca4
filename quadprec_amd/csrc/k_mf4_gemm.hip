// k_mf4_gemm.hip -- MultiFloat<4> (f256) GEMM, bit-identical to the reference.
//
// Reference: crates/quad/src/multifloat/linalg.rs:39-61 (mf_gemm) and 64-85
// (mf_gemm_atb): per element s = 0; for p in 0..k: s = s + a[i,p]*b[p,j] with the MF4
// Add (multifloat/add.rs:19-106) and Mul (multifloat/mul.rs:12-111).  The per-element
// order is kept, so results match bit for bit.
//
// One MF4 MAC is ~600-900 FP64 operations with data-dependent compaction, so the
// kernel is VALU-bound by a wide margin; operand traffic is negligible.  A 256-thread
// workgroup owns a 16 x 16 tile, one element per thread, and stages 16-deep k slabs of
// the four limb planes through LDS.  Layout: limb-major planes, plane l of an r x c
// matrix at offset l*r*c.
#include "hp_arith.hpp"
#include "hp_internal.hpp"

namespace hp {
namespace {

constexpr int TB = 16, KB = 16;

template <bool TRANS_A>
__global__ __launch_bounds__(256) void k_mf4_gemm(const double* __restrict__ a, const double* __restrict__ b,
                                                  double* __restrict__ c, int m, int n, int k) {
  __shared__ double As[4][KB][TB + 1];
  __shared__ double Bs[4][KB][TB + 1];
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
  const int i = blockIdx.y * TB + ty, j = blockIdx.x * TB + tx;
  const size_t pa = (size_t)m * k, pb = (size_t)k * n, pc = (size_t)m * n;
  mf4 s = mkmf(0.0);
  for (int p0 = 0; p0 < k; p0 += KB) {
    // A slab: rows blockIdx.y*TB + ty, k p0 + tx ; B slab: k p0 + ty, cols j
    {
      int gi = blockIdx.y * TB + ty, gp = p0 + tx;
      bool ok = gi < m && gp < k;
      size_t off = TRANS_A ? (size_t)gp * m + gi : (size_t)gi * k + gp;
#pragma unroll
      for (int l = 0; l < 4; ++l) As[l][tx][ty] = ok ? a[l * pa + off] : 0.0;
      int gp2 = p0 + ty;
      bool ok2 = j < n && gp2 < k;
      size_t off2 = (size_t)gp2 * n + j;
#pragma unroll
      for (int l = 0; l < 4; ++l) Bs[l][ty][tx] = ok2 ? b[l * pb + off2] : 0.0;
    }
    __syncthreads();
    const int kmax = min(KB, k - p0);
    for (int kk = 0; kk < kmax; ++kk) {
      mf4 av = mkmf(As[0][kk][ty], As[1][kk][ty], As[2][kk][ty], As[3][kk][ty]);
      mf4 bv = mkmf(Bs[0][kk][tx], Bs[1][kk][tx], Bs[2][kk][tx], Bs[3][kk][tx]);
      s = mf_add(s, mf_mul(av, bv));
    }
    __syncthreads();
  }
  if (i < m && j < n) {
    size_t off = (size_t)i * n + j;
    c[off] = s.l0;
    c[pc + off] = s.l1;
    c[2 * pc + off] = s.l2;
    c[3 * pc + off] = s.l3;
  }
}

}  // namespace

int launch_mf4_gemm(int trans_a, const double* a, const double* b, double* c, size_t m, size_t n, size_t k,
                    hipStream_t s) {
  if (m == 0 || n == 0) return HP_OK;
  if ((m + TB - 1) / TB > 65535 || k > (1u << 30) || n > (1u << 30)) {
    set_error("mf_gemm: dimension too large");
    return HP_ESHAPE;
  }
  dim3 grid((unsigned)((n + TB - 1) / TB), (unsigned)((m + TB - 1) / TB));
  if (trans_a)
    hipLaunchKernelGGL(k_mf4_gemm<true>, grid, dim3(256), 0, s, a, b, c, (int)m, (int)n, (int)k);
  else
    hipLaunchKernelGGL(k_mf4_gemm<false>, grid, dim3(256), 0, s, a, b, c, (int)m, (int)n, (int)k);
  HP_LAUNCH_CHECK();
  return HP_OK;
}

}  // namespace hp
