// k_overlap.hip -- normalised s-type Gaussian overlap matrix in double-double.
//
//   S_uv = (4 a_u a_v / (a_u + a_v)^2)^{3/4} * exp(-(a_u a_v / (a_u + a_v)) |R_u - R_v|^2)
//
// the single-primitive, coeff = (2a/pi)^{3/4} case of the reference's s-s overlap
// (crates/overlap/src/integral.rs:15-56, 93-142; assembly crates/overlap/src/assemble.rs:35-85)
// in the closed form of crates/demo/src/main.rs:104-124 with d = 3.  Every step is a dd
// operation of crates/quad/src/dd.rs (add, mul, div, sqrt, exp); centre differences are
// exact (TwoSum), the ^{3/4} is sqrt(t)*sqrt(sqrt(t)).  The reference's dd exp builds 2^k
// from exponent bits and is meaningful for arguments >= about -708 (dd.rs:119); below
// that the entry is 0 (the true value is < 1e-307, far under a dd ulp of the unit
// diagonal).  No screening: every pair is computed (the reference's ScreenParams skip
// below 1e-14 would destroy a kappa ~ 1e20 matrix).
//
// Per entry ~1.5k FP64 VALU operations (the exp's ~20 Taylor terms, each a dd mul and a dd
// div) against 16 bytes written, so the kernel is FP64-VALU-bound.  One thread per (u, v)
// of a 16x16 tile of the lower triangle; each entry is mirrored into the upper triangle.
#include "hp_arith.hpp"
#include "hp_internal.hpp"
#include "hp_overlap.hpp"

namespace hp {

namespace {
__global__ __launch_bounds__(256) void k_overlap_s(const double* __restrict__ centers, const double* __restrict__ alpha,
                                                   long n, double* __restrict__ shi, double* __restrict__ slo) {
  const int ti = blockIdx.y, tj = blockIdx.x;
  if (tj > ti) return;
  const long u = (long)ti * 16 + (threadIdx.x >> 4), v = (long)tj * 16 + (threadIdx.x & 15);
  if (u >= n || v > u) return;
  const double cu[3] = {centers[3 * u], centers[3 * u + 1], centers[3 * u + 2]};
  const double cv[3] = {centers[3 * v], centers[3 * v + 1], centers[3 * v + 2]};
  dd s = overlap_s_entry(alpha[u], cu, alpha[v], cv);
  shi[u * n + v] = s.hi;
  slo[u * n + v] = s.lo;
  shi[v * n + u] = s.hi;
  slo[v * n + u] = s.lo;
}
}  // namespace

int launch_overlap_s(const double* centers, const double* alpha, size_t n, double* shi, double* slo, hipStream_t s) {
  if (n == 0) return HP_OK;
  size_t t = (n + 15) / 16;
  if (t > 65535) {
    set_error("overlap: n too large");
    return HP_ESHAPE;
  }
  hipLaunchKernelGGL(k_overlap_s, dim3((unsigned)t, (unsigned)t), dim3(256), 0, s, centers, alpha, (long)n, shi, slo);
  HP_LAUNCH_CHECK();
  return HP_OK;
}

}  // namespace hp
