// hp_overlap.hpp -- the normalised s-type Gaussian overlap entry in double-double,
//   S_uv = (4 a_u a_v / (a_u + a_v)^2)^{3/4} exp(-(a_u a_v / (a_u + a_v)) |R_u - R_v|^2),
// shared by the overlap build and its adjoint (which recomputes entries instead of
// storing them).  Every step is a dd operation of crates/quad/src/dd.rs; centre
// differences are exact (TwoSum); arguments below -708 give 0 (dd.rs:119 builds 2^k from
// exponent bits and is meaningful only above that).
#pragma once
#include "hp_arith.hpp"

namespace hp {

__host__ __device__ __forceinline__ dd overlap_s_entry(double au, const double* cu, double av, const double* cv) {
  dd a = mkdd(au), b = mkdd(av);
  dd g = dd_add(a, b);
  dd ab = dd_mul(a, b);
  dd p = dd_div(ab, g);
  dd ratio = dd_div(dd_mul_f(ab, 4.0), dd_mul(g, g));
  dd t = dd_sqrt(ratio);
  dd pref = dd_mul(t, dd_sqrt(t));
  dd d[3];
  for (int c = 0; c < 3; ++c) two_sum(cu[c], -cv[c], d[c].hi, d[c].lo);
  dd r2 = dd_add(dd_add(dd_mul(d[0], d[0]), dd_mul(d[1], d[1])), dd_mul(d[2], d[2]));
  dd arg = dd_neg(dd_mul(p, r2));
  if (arg.hi < -708.0) return mkdd(0.0, 0.0);
  return dd_mul(pref, dd_exp(arg));
}

}  // namespace hp
