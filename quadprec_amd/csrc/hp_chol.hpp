// hp_chol.hpp -- arithmetic policies shared by the Cholesky / triangular-solve kernels.
//
// The exact (reference-order) factorisations are written once and instantiated for
//   ArDD  : double-double, as crates/quad/src/linalg.rs:42-126, 224-293
//   ArF64 : binary64, as crates/solver/src/lib.rs:169-213 (cholesky_f64 and its solves)
// A matrix is a pair of planes (hi, lo); the f64 policy ignores lo.
#pragma once
#include "hp_arith.hpp"

namespace hp {

struct Mat {
  double* hi;
  double* lo;
  long ld;
};

struct ErrRec {       // first failure of a factorisation (device memory)
  int flag;           // 0 ok, 1 not positive definite
  int pad;
  long long pivot;    // global pivot index
  double value;       // f64 value of the failing diagonal
};

struct ArDD {
  using T = dd;
  static constexpr int LIMBS = 2;
  __device__ static T zero() { return mkdd(0.0, 0.0); }
  __device__ static T one() { return mkdd(1.0, 0.0); }
  __device__ static T ld(const Mat& m, long i, long j) { long o = i * m.ld + j; return mkdd(m.hi[o], m.lo[o]); }
  __device__ static void st(const Mat& m, long i, long j, T v) { long o = i * m.ld + j; m.hi[o] = v.hi; m.lo[o] = v.lo; }
  __device__ static T ldv(const double* hi, const double* lo, long i) { return mkdd(hi[i], lo[i]); }
  __device__ static void stv(double* hi, double* lo, long i, T v) { hi[i] = v.hi; lo[i] = v.lo; }
  __device__ static T add(T a, T b) { return dd_add(a, b); }
  __device__ static T sub(T a, T b) { return dd_sub(a, b); }
  __device__ static T mul(T a, T b) { return dd_mul(a, b); }
  __device__ static T div(T a, T b) { return dd_div(a, b); }
  __device__ static T sqrt(T a) { return dd_sqrt(a); }
  __device__ static T recip(T a) { return dd_recip(a); }          // ljj.recip() (linalg.rs:60)
  __device__ static bool nonpos(T d) { return d.hi <= 0.0; }      // linalg.rs:55
  __device__ static double f64(T d) { return d.hi + d.lo; }       // to_f64 (dd.rs:53-55)
  __device__ static T readlane(T v, int lane) {
    return mkdd(__shfl(v.hi, lane, 64), __shfl(v.lo, lane, 64));
  }
  // sqrt(d) and 1/sqrt(d) by one dd Newton step, y = y0 + y0 (1 - d y0^2) / 2, sqrt = d y,
  // from y0 = the hardware reciprocal square root of d.hi refined by two binary64 Newton
  // steps: each at least doubles the correct bits, so any seed better than 2^-14 reaches
  // binary64 rounding (~2^-52) before the dd step and ~2^-104 after.  No division (the
  // reference takes sqrt, then recip: dd.rs:61-83); used where reference order is not
  // required.
  __device__ static void sqrt_recip_fast(T d, T& root, T& inv) {
    double r = __builtin_amdgcn_rsq(d.hi);
    r = r * __builtin_fma(-(0.5 * d.hi) * r, r, 1.5);
    r = r * __builtin_fma(-(0.5 * d.hi) * r, r, 1.5);
    const dd y0 = mkdd(r);
    const dd e = dd_sub(mkdd(1.0), dd_mul(d, dd_mul(y0, y0)));
    inv = dd_add(y0, dd_mul_f(dd_mul(y0, e), 0.5));
    root = dd_mul(d, inv);
  }
};

struct ArF64 {
  using T = double;
  static constexpr int LIMBS = 1;
  __device__ static T zero() { return 0.0; }
  __device__ static T one() { return 1.0; }
  __device__ static T ld(const Mat& m, long i, long j) { return m.hi[i * m.ld + j]; }
  __device__ static void st(const Mat& m, long i, long j, T v) { m.hi[i * m.ld + j] = v; }
  __device__ static T ldv(const double* hi, const double*, long i) { return hi[i]; }
  __device__ static void stv(double* hi, double*, long i, T v) { hi[i] = v; }
  __device__ static T add(T a, T b) { return a + b; }
  __device__ static T sub(T a, T b) { return a - b; }
  __device__ static T mul(T a, T b) { return a * b; }
  __device__ static T div(T a, T b) { return a / b; }
  __device__ static T sqrt(T a) { return __builtin_sqrt(a); }
  __device__ static T recip(T a) { return 1.0 / a; }              // lib.rs:182
  __device__ static bool nonpos(T d) { return d <= 0.0; }         // lib.rs:177
  __device__ static double f64(T d) { return d; }
  __device__ static T readlane(T v, int lane) { return __shfl(v, lane, 64); }
};

}  // namespace hp
