// hp_arith.hpp -- double-double and MultiFloat<4> arithmetic for gfx950 kernels and the
// host side of the library.
//
// The operation sequences are the reference's, so results are bit-identical to it:
//   error-free transforms    crates/quad/src/dd.rs:15-35
//   dd add / mul / div       dd.rs:147-195, mixed f64 ops dd.rs:228-260
//   dd sqrt / recip / exp    dd.rs:61-83, 106-121
//   MF4 renormalisation      crates/quad/src/multifloat/core.rs:57-168
//   MF4 add / mul            multifloat/add.rs:19-106, multifloat/mul.rs:12-111
// Every source including this header must be compiled with -ffp-contract=off: the only
// fused multiply-add allowed is the explicit one in two_prod (dd.rs:33).
//
// The MF4 routines index their small work arrays only with compile-time indices (fully
// unrolled loops, data-dependent positions resolved with selects) so that the arrays
// live in VGPRs rather than in scratch memory.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define HP_HD __host__ __device__ __forceinline__

namespace hp {

struct dd { double hi, lo; };
struct mf4 { double l0, l1, l2, l3; };

// ------------------------------------------------------------------ error-free transforms
HP_HD void two_sum(double a, double b, double& s, double& e) {
  double ss = a + b;
  double v = ss - a;
  e = (a - (ss - v)) + (b - v);
  s = ss;
}
HP_HD void quick_two_sum(double a, double b, double& s, double& e) {
  double ss = a + b;
  e = b - (ss - a);
  s = ss;
}
HP_HD void two_prod(double a, double b, double& p, double& e) {
  double pp = a * b;
  e = __builtin_fma(a, b, -pp);
  p = pp;
}

// ------------------------------------------------------------------ double-double
HP_HD dd mkdd(double hi, double lo = 0.0) { dd r; r.hi = hi; r.lo = lo; return r; }
HP_HD dd renorm(double hi, double lo) { dd r; quick_two_sum(hi, lo, r.hi, r.lo); return r; }
HP_HD dd dd_neg(dd a) { return mkdd(-a.hi, -a.lo); }
HP_HD dd dd_add(dd a, dd b) {
  double s1, e1, s2, e2, s, e;
  two_sum(a.hi, b.hi, s1, e1);
  two_sum(a.lo, b.lo, s2, e2);
  e = e1 + s2;
  quick_two_sum(s1, e, s, e);
  return renorm(s, e + e2);
}
HP_HD dd dd_sub(dd a, dd b) { return dd_add(a, dd_neg(b)); }
HP_HD dd dd_mul(dd a, dd b) {
  double p1, e1;
  two_prod(a.hi, b.hi, p1, e1);
  double e = e1 + (a.hi * b.lo + a.lo * b.hi);
  return renorm(p1, e);
}
HP_HD dd dd_div(dd a, dd b) {
  double q1 = a.hi / b.hi;
  dd r = dd_sub(a, dd_mul(mkdd(q1), b));
  double q2 = r.hi / b.hi;
  return renorm(q1, q2);
}
HP_HD dd dd_add_f(dd a, double b) {
  double s1, e1;
  two_sum(a.hi, b, s1, e1);
  return renorm(s1, e1 + a.lo);
}
HP_HD dd dd_mul_f(dd a, double b) {
  double p1, e1;
  two_prod(a.hi, b, p1, e1);
  return renorm(p1, e1 + a.lo * b);
}
HP_HD dd dd_abs(dd a) { return __builtin_signbit(a.hi) ? dd_neg(a) : a; }
HP_HD dd dd_sqrt(dd a) {
  if (a.hi < 0.0) return mkdd(__builtin_nan(""), 0.0);
  if (a.hi == 0.0) return mkdd(0.0, 0.0);
  if (!__builtin_isfinite(a.hi)) return mkdd(a.hi, 0.0);
  dd x = mkdd(1.0 / __builtin_sqrt(a.hi));
  dd ax = dd_mul(a, x);
  dd hx = dd_mul(x, mkdd(0.5));
  dd r = dd_add(ax, dd_mul(ax, dd_sub(mkdd(0.5), dd_mul(ax, hx))));
  return dd_mul(dd_add(r, dd_div(a, r)), mkdd(0.5));
}
HP_HD dd dd_recip(dd a) {
  dd x0 = mkdd(1.0 / a.hi);
  return dd_add(x0, dd_mul(x0, dd_sub(mkdd(1.0), dd_mul(a, x0))));
}
// exp(a) by reduction k = round(a.hi / ln2.hi) and a Taylor series to |term| < 1e-35.
// The scale 2^k is built from exponent bits exactly as the reference does, so like the
// reference it is meaningful only for k >= -1022 (a.hi >= about -708).
HP_HD dd dd_exp(dd a) {
  const dd LN2 = mkdd(0.6931471805599453, 2.3190468138462996e-17);
  if (a.hi == 0.0 && a.lo == 0.0) return mkdd(1.0, 0.0);
  double k = __builtin_round(a.hi / LN2.hi);
  dd r = dd_sub(a, dd_mul(mkdd(k), LN2));
  dd term = mkdd(1.0), sum = mkdd(1.0);
  for (int i = 1; i <= 25; ++i) {
    term = dd_div(dd_mul(term, r), mkdd((double)i));
    sum = dd_add(sum, term);
    if (__builtin_fabs(term.hi) < 1e-35) break;
  }
  long long ki = (k >= 9.2233720368547758e18) ? INT64_MAX : (k <= -9.2233720368547758e18 ? INT64_MIN : (long long)k);
  uint64_t bits = ((uint64_t)1023 + (uint64_t)ki) << 52;  // wraps like the reference's `as u64`
  double scale = __builtin_bit_cast(double, bits);
  return dd_mul(sum, mkdd(scale));
}
// ln(a): binary64 seed log(a.hi), then three Newton steps x += a exp(-x) - 1 in dd
// (dd.rs:124-137).  NaN for a.hi <= 0, exactly 0 for a == 1.  The seed comes from the
// platform's log; after the Newton steps the result is fixed by the dd operations.
HP_HD dd dd_ln(dd a) {
  if (a.hi <= 0.0) return mkdd(__builtin_nan(""), __builtin_nan(""));
  if (a.hi == 1.0 && a.lo == 0.0) return mkdd(0.0, 0.0);
  dd x = mkdd(log(a.hi));
  for (int i = 0; i < 3; ++i) x = dd_add(x, dd_sub(dd_mul(a, dd_exp(dd_neg(x))), mkdd(1.0)));
  return x;
}

// ------------------------------------------------------------------ MultiFloat<4>
HP_HD mf4 mkmf(double a, double b = 0.0, double c = 0.0, double d = 0.0) {
  mf4 r; r.l0 = a; r.l1 = b; r.l2 = c; r.l3 = d; return r;
}
// select-based store into slot k of a 4-array held in named registers
#define HP_PUT4(o, k, v) do { double _v = (v); int _k = (k);                  \
    o[0] = (_k == 0) ? _v : o[0]; o[1] = (_k == 1) ? _v : o[1];               \
    o[2] = (_k == 2) ? _v : o[2]; o[3] = (_k == 3) ? _v : o[3]; } while (0)

HP_HD void mf_renorm_pass(double (&l)[4]) {
  double s, e, ns;
  two_sum(l[2], l[3], s, e);
  l[3] = e;
  two_sum(l[1], s, ns, e); s = ns; l[2] = e;
  two_sum(l[0], s, ns, e); s = ns; l[1] = e;
  l[0] = s;
  double out[4] = {0.0, 0.0, 0.0, 0.0};
  s = l[0];
  int k = 0;
#pragma unroll
  for (int i = 1; i < 4; ++i) {
    quick_two_sum(s, l[i], ns, e);
    s = ns;
    if (e != 0.0) { HP_PUT4(out, k, s); s = e; ++k; }
  }
  HP_PUT4(out, k, s);
#pragma unroll
  for (int i = 0; i < 4; ++i) l[i] = out[i];
}
HP_HD void mf_renormalize(double (&l)[4]) { mf_renorm_pass(l); mf_renorm_pass(l); }

HP_HD void mf_renormalize_from8(const double (&in)[8], double (&out)[4]) {
  double t[8];
#pragma unroll
  for (int i = 0; i < 8; ++i) t[i] = in[i];
  double s, e, ns;
  // pass 1: bottom-up two_sum, then top-down compaction into 8 slots
  two_sum(t[6], t[7], s, e);
  t[7] = e;
#pragma unroll
  for (int i = 5; i >= 0; --i) { two_sum(t[i], s, ns, e); s = ns; t[i + 1] = e; }
  t[0] = s;
  {
    double buf[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    s = t[0];
    int k = 0;
#pragma unroll
    for (int i = 1; i < 8; ++i) {
      quick_two_sum(s, t[i], ns, e);
      s = ns;
      if (e != 0.0) {
#pragma unroll
        for (int q = 0; q < 8; ++q) buf[q] = (k == q) ? s : buf[q];
        s = e; ++k;
      }
    }
#pragma unroll
    for (int q = 0; q < 8; ++q) buf[q] = (k == q) ? s : buf[q];
#pragma unroll
    for (int i = 0; i < 8; ++i) t[i] = buf[i];
  }
  // pass 2: same sweep, emitting at most 4 limbs
  two_sum(t[6], t[7], s, e);
  t[7] = e;
#pragma unroll
  for (int i = 5; i >= 0; --i) { two_sum(t[i], s, ns, e); s = ns; t[i + 1] = e; }
  t[0] = s;
  out[0] = out[1] = out[2] = out[3] = 0.0;
  s = t[0];
  int k = 0;
  bool done = false;
#pragma unroll
  for (int i = 1; i < 8; ++i) {
    quick_two_sum(s, t[i], ns, e);
    if (!done) {
      s = ns;
      if (e != 0.0) {
        HP_PUT4(out, k, s);
        s = e; ++k;
        if (k >= 4) done = true;
      }
    }
  }
  if (!done) HP_PUT4(out, k, s);
}

// The general grow-expansion of the sorted terms w into ex (zeros past its length), as in
// mf_add_t<false>.  Before term i the expansion holds at most i limbs (each term adds at
// most one), so the loops stop at i: the same operations as running them to 8 under the
// len guard, which skips the rest.
HP_HD void mf_add_expand(const double (&w)[8], double (&ex)[8]) {
#pragma unroll
  for (int i = 0; i < 8; ++i) ex[i] = 0.0;
  int len = 0;
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    double term = w[i];
    if (term == 0.0) continue;
    if (len == 0) { ex[0] = term; len = 1; continue; }
    double carry = term;
#pragma unroll
    for (int j = i - 1; j >= 0; --j) {
      if (j < len) {
        double s, e;
        two_sum(ex[j], carry, s, e);
        ex[j] = e;
        carry = s;
      }
    }
    if (carry != 0.0) {
      int nl = len + 1;
#pragma unroll
      for (int q = i; q >= 1; --q)
        if (q < nl) ex[q] = ex[q - 1];
      ex[0] = carry;
      len = nl;
    }
  }
}
// FIXED: the expansion takes the fixed sequence below when its conditions hold (fewer
// dependent operations per add: for serial chains), else the general loop only (fewer
// registers: for throughput-bound kernels such as the GEMM); both give the same bits.
template <bool FIXED>
HP_HD mf4 mf_add_t(mf4 a, mf4 b) {
  double w[8];
  two_sum(a.l0, b.l0, w[0], w[1]);
  two_sum(a.l1, b.l1, w[2], w[3]);
  two_sum(a.l2, b.l2, w[4], w[5]);
  two_sum(a.l3, b.l3, w[6], w[7]);
  // stable insertion sort by descending magnitude (add.rs:42-54)
#pragma unroll
  for (int i = 1; i < 8; ++i) {
    double v = w[i];
    double mag = __builtin_fabs(v);
    bool moving = true;
#pragma unroll
    for (int j = i; j >= 1; --j) {
      bool c = moving && (__builtin_fabs(w[j - 1]) < mag);
      double nw = c ? w[j - 1] : (moving ? v : w[j]);
      moving = c;
      w[j] = nw;
    }
    if (moving) w[0] = v;
  }
  // grow-expansion, most significant first (add.rs:60-95).  When no term is zero and no
  // cascade's carry is zero (the usual case), term i meets exactly i limbs: the fixed
  // sequence below performs the reference's operations without its length bookkeeping,
  // and a lane where either condition fails redoes the expansion with the general loop
  if constexpr (!FIXED) {
    double ex[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    int len = 0;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      double term = w[i];
      if (term == 0.0) continue;
      if (len == 0) { ex[0] = term; len = 1; continue; }
      double carry = term;
#pragma unroll
      for (int j = i - 1; j >= 0; --j) {
        if (j < len) {
          double s, e;
          two_sum(ex[j], carry, s, e);
          ex[j] = e;
          carry = s;
        }
      }
      if (carry != 0.0) {
        int nl = len + 1;
#pragma unroll
        for (int q = i; q >= 1; --q)
          if (q < nl) ex[q] = ex[q - 1];
        ex[0] = carry;
        len = nl;
      }
    }
    double l[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) l[i] = (i < len) ? ex[i] : 0.0;
    mf_renormalize(l);
    return mkmf(l[0], l[1], l[2], l[3]);
  }
  double ex[8];
  bool irregular = w[0] == 0.0;
  ex[0] = w[0];
#pragma unroll
  for (int i = 1; i < 8; ++i) {
    double carry = w[i];
    irregular |= carry == 0.0;
#pragma unroll
    for (int j = i - 1; j >= 0; --j) {
      double s, e;
      two_sum(ex[j], carry, s, e);
      ex[j] = e;
      carry = s;
    }
    irregular |= carry == 0.0;
#pragma unroll
    for (int q = i; q >= 1; --q) ex[q] = ex[q - 1];
    ex[0] = carry;
  }
  if (irregular) mf_add_expand(w, ex);
  double l[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) l[i] = ex[i];
  mf_renormalize(l);
  return mkmf(l[0], l[1], l[2], l[3]);
}
HP_HD mf4 mf_add(mf4 a, mf4 b) { return mf_add_t<false>(a, b); }
HP_HD mf4 mf_neg(mf4 a) { return mkmf(-a.l0, -a.l1, -a.l2, -a.l3); }
HP_HD mf4 mf_sub(mf4 a, mf4 b) { return mf_add(a, mf_neg(b)); }
// the same subtraction with the fixed expansion sequence, for serial chains
HP_HD mf4 mf_sub_chain(mf4 a, mf4 b) { return mf_add_t<true>(a, mf_neg(b)); }

// error-free accumulate of v into w[idx] with the cascade of mul.rs:31-42 (limit 8)
template <int IDX>
HP_HD void mf_accumulate(double (&w)[8], double v) {
  double s, e;
  two_sum(w[IDX], v, s, e);
  w[IDX] = s;
#pragma unroll
  for (int k = IDX + 1; k < 8; ++k) {
    if (e != 0.0) {
      double s2, e2;
      two_sum(w[k], e, s2, e2);
      w[k] = s2;
      e = e2;
    }
  }
}
// mf_accumulate while the slots from F on are still +0.0 (no cascade has reached them)
// and every value is finite: two_sum(+0.0, e) is (e + 0.0, +0.0) for a finite e, so a
// cascade that reaches slot F stops there and the later levels are never taken; slot F
// then holds 0.0 + e, which is also the +0.0 it keeps when the cascade stopped earlier
// (e = +-0.0).  Slot IDX >= F takes 0.0 + v the same way.
template <int IDX, int F>
HP_HD void mf_accumulate_front(double (&w)[8], double v) {
  if constexpr (IDX >= F) {
    w[IDX] = 0.0 + v;
  } else {
    double s, e;
    two_sum(w[IDX], v, s, e);
    w[IDX] = s;
#pragma unroll
    for (int k = IDX + 1; k < F; ++k) {  // unguarded: the same bits (mf_accumulate_finite)
      double s2, e2;
      two_sum(w[k], e, s2, e2);
      w[k] = s2;
      e = e2;
    }
    w[F] = 0.0 + e;
  }
}
// mf_accumulate without the e != 0 guard, for finite values in slots that never hold -0.0.
// Slots start at +0.0 and a two_sum's sum is -0.0 only when both addends are -0.0, so no
// slot becomes -0.0; for such a finite slot w, two_sum(w, +-0.0) returns (w, +0.0) (w + 0
// is w, (w - w) + (+-0 - +0) is +0), so the levels the guard would skip leave every slot
// and the carried error unchanged: the same bits as mf_accumulate, without the compare
// and selects per level.
template <int IDX>
HP_HD void mf_accumulate_finite(double (&w)[8], double v) {
  double s, e;
  two_sum(w[IDX], v, s, e);
  w[IDX] = s;
#pragma unroll
  for (int k = IDX + 1; k < 8; ++k) {
    double s2, e2;
    two_sum(w[k], e, s2, e2);
    w[k] = s2;
    e = e2;
  }
}
// (i, j) pairs with i + j < 4 in the reference's loop order (mul.rs:212-224); FAST: every
// limb is below 2^500 in magnitude, so no product, sum or error overflows and the first
// four products' cascades stop at the first untouched slot (slots 0..7 are reached in
// order: the frontier before their eight accumulations is 0, 1, ..., 7)
template <bool FAST>
HP_HD mf4 mf_mul_impl(mf4 a, mf4 b) {
  double w[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  const double A[4] = {a.l0, a.l1, a.l2, a.l3};
  const double B[4] = {b.l0, b.l1, b.l2, b.l3};
  double p, e;
#define HP_MFMUL_TERM(I, J)            \
  two_prod(A[I], B[J], p, e);          \
  mf_accumulate<(I) + (J)>(w, p);      \
  mf_accumulate<(I) + (J) + 1>(w, e);
#define HP_MFMUL_FRONT(I, J, F)                        \
  two_prod(A[I], B[J], p, e);                          \
  mf_accumulate_front<(I) + (J), (F)>(w, p);           \
  mf_accumulate_front<(I) + (J) + 1, (F) + 1>(w, e);
  if constexpr (FAST) {
    HP_MFMUL_FRONT(0, 0, 0) HP_MFMUL_FRONT(0, 1, 2) HP_MFMUL_FRONT(0, 2, 4) HP_MFMUL_FRONT(0, 3, 6)
  } else {
    HP_MFMUL_TERM(0, 0) HP_MFMUL_TERM(0, 1) HP_MFMUL_TERM(0, 2) HP_MFMUL_TERM(0, 3)
  }
#define HP_MFMUL_FIN(I, J)             \
  two_prod(A[I], B[J], p, e);          \
  mf_accumulate_finite<(I) + (J)>(w, p); \
  mf_accumulate_finite<(I) + (J) + 1>(w, e);
  if constexpr (FAST) {  // all values finite, the slots never -0.0 (see mf_accumulate_finite)
    HP_MFMUL_FIN(1, 0) HP_MFMUL_FIN(1, 1) HP_MFMUL_FIN(1, 2)
    HP_MFMUL_FIN(2, 0) HP_MFMUL_FIN(2, 1)
    HP_MFMUL_FIN(3, 0)
  } else {
    HP_MFMUL_TERM(1, 0) HP_MFMUL_TERM(1, 1) HP_MFMUL_TERM(1, 2)
    HP_MFMUL_TERM(2, 0) HP_MFMUL_TERM(2, 1)
    HP_MFMUL_TERM(3, 0)
  }
#undef HP_MFMUL_FIN
#undef HP_MFMUL_FRONT
#undef HP_MFMUL_TERM
  double o[4];
  mf_renormalize_from8(w, o);
  return mkmf(o[0], o[1], o[2], o[3]);
}
HP_HD mf4 mf_mul(mf4 a, mf4 b) {
  constexpr double LIM = 0x1p500;
  const bool fast = (__builtin_fabs(a.l0) < LIM) & (__builtin_fabs(a.l1) < LIM) & (__builtin_fabs(a.l2) < LIM) &
                    (__builtin_fabs(a.l3) < LIM) & (__builtin_fabs(b.l0) < LIM) & (__builtin_fabs(b.l1) < LIM) &
                    (__builtin_fabs(b.l2) < LIM) & (__builtin_fabs(b.l3) < LIM);
  if (fast) return mf_mul_impl<true>(a, b);
  return mf_mul_impl<false>(a, b);
}

// MultiFloat<4> reciprocal (multifloat/div.rs:14-41): Newton from the f64 reciprocal,
// ceil(log2 4) + 2 = 4 steps of x + x (1 - a x); 1/0 is the signed infinity, 1/non-finite 0
HP_HD mf4 mf_recip(mf4 a) {
  if (a.l0 == 0.0) return mkmf(signbit(a.l0) ? -(double)INFINITY : (double)INFINITY);
  if (!isfinite(a.l0)) return mkmf(0.0);
  mf4 x = mkmf(1.0 / a.l0);
  const mf4 one = mkmf(1.0);
  for (int s = 0; s < 4; ++s) {
    const mf4 residual = mf_sub(one, mf_mul(a, x));
    x = mf_add(x, mf_mul(x, residual));
  }
  return x;
}
// a / b (div.rs:44-77): the special cases, then a * recip(b)
HP_HD mf4 mf_div(mf4 a, mf4 b) {
  if (b.l0 == 0.0) {
    const double sign = (!signbit(a.l0) == !signbit(b.l0)) ? 1.0 : -1.0;
    return mkmf(a.l0 == 0.0 ? (double)NAN : sign * (double)INFINITY);
  }
  if (a.l0 == 0.0) return mkmf(0.0);
  if (!isfinite(a.l0) || !isfinite(b.l0)) return mkmf(a.l0 / b.l0);
  return mf_mul(a, mf_recip(b));
}
// sqrt (multifloat/math.rs:22-40): NaN below zero, 0 at 0, non-finite passed through, else
// ceil(log2 4) + 1 = 3 steps of (x + a / x) / 2 from the f64 root
HP_HD mf4 mf_sqrt(mf4 a) {
  if (a.l0 < 0.0) return mkmf((double)NAN);
  if (a.l0 == 0.0) return mkmf(0.0);
  if (!isfinite(a.l0)) return mkmf(a.l0);
  mf4 x = mkmf(sqrt(a.l0));
  const mf4 half = mkmf(0.5);
  for (int s = 0; s < 3; ++s) x = mf_mul(mf_add(x, mf_div(a, x)), half);
  return x;
}

}  // namespace hp
