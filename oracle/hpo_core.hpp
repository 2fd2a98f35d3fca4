// hpo_core.hpp -- templated restatement of the reference arithmetic (oracle only).
//
// TEST INFRASTRUCTURE: see hpo.h.  Every function is templated on the binary64 scalar
// type F so that the same code runs on plain double (the oracle) and on a counting
// double (op counts for the roofline's algorithmic work).
//
// Each function names the reference lines it restates.  The order of every floating
// point operation is the reference's; nothing is reassociated or fused.
#pragma once
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>

namespace hpo {

// ---- scalar helpers so F may be double or a counting wrapper ----
inline double fma_(double a, double b, double c) { return std::fma(a, b, c); }
inline double sqrt_(double a) { return std::sqrt(a); }
inline double val(double a) { return a; }

// ---- error-free transforms: crates/quad/src/dd.rs:15-35, multifloat/core.rs:19-42 ----
template <class F> inline void two_sum(F a, F b, F& s, F& e) {
  F ss = a + b;
  F v = ss - a;
  e = (a - (ss - v)) + (b - v);
  s = ss;
}
template <class F> inline void quick_two_sum(F a, F b, F& s, F& e) {
  F ss = a + b;
  e = b - (ss - a);
  s = ss;
}
template <class F> inline void two_prod(F a, F b, F& p, F& e) {
  F pp = a * b;
  e = fma_(a, b, -pp);
  p = pp;
}

// ---- f128 double-double: crates/quad/src/dd.rs ----
template <class F> struct DD { F hi, lo; };

template <class F> inline DD<F> dd(F hi, F lo) { return DD<F>{hi, lo}; }
template <class F> inline DD<F> dd_from(F x) { return DD<F>{x, F(0.0)}; }
template <class F> inline DD<F> renorm(F hi, F lo) {  // dd.rs:86-89
  DD<F> r; quick_two_sum(hi, lo, r.hi, r.lo); return r;
}
template <class F> inline DD<F> neg(DD<F> a) { return DD<F>{-a.hi, -a.lo}; }  // dd.rs:167-173
template <class F> inline DD<F> add(DD<F> a, DD<F> b) {  // dd.rs:147-157
  F s1, e1, s2, e2;
  two_sum(a.hi, b.hi, s1, e1);
  two_sum(a.lo, b.lo, s2, e2);
  F e = e1 + s2;
  F s;
  quick_two_sum(s1, e, s, e);
  return renorm(s, e + e2);
}
template <class F> inline DD<F> sub(DD<F> a, DD<F> b) { return add(a, neg(b)); }  // dd.rs:159-165
template <class F> inline DD<F> mul(DD<F> a, DD<F> b) {  // dd.rs:175-183
  F p1, e1;
  two_prod(a.hi, b.hi, p1, e1);
  F e = e1 + (a.hi * b.lo + a.lo * b.hi);
  return renorm(p1, e);
}
template <class F> inline DD<F> div(DD<F> a, DD<F> b) {  // dd.rs:185-195
  F q1 = a.hi / b.hi;
  DD<F> r = sub(a, mul(dd_from(q1), b));
  F q2 = r.hi / b.hi;
  return renorm(q1, q2);
}
template <class F> inline DD<F> add_f(DD<F> a, F b) {  // dd.rs:228-235
  F s1, e1;
  two_sum(a.hi, b, s1, e1);
  return renorm(s1, e1 + a.lo);
}
template <class F> inline DD<F> mul_f(DD<F> a, F b) {  // dd.rs:243-250
  F p1, e1;
  two_prod(a.hi, b, p1, e1);
  return renorm(p1, e1 + a.lo * b);
}
template <class F> inline DD<F> div_f(DD<F> a, F b) {  // dd.rs:252-260
  F q1 = a.hi / b;
  DD<F> r = sub(a, mul(dd_from(q1), dd_from(b)));
  return renorm(q1, r.hi / b);
}
template <class F> inline DD<F> dabs(DD<F> a) {  // dd.rs:57-59
  return std::signbit(val(a.hi)) ? neg(a) : a;
}
template <class F> inline DD<F> dsqrt(DD<F> a) {  // dd.rs:61-78
  if (val(a.hi) < 0.0) return DD<F>{F(std::numeric_limits<double>::quiet_NaN()), F(0.0)};
  if (val(a.hi) == 0.0) return DD<F>{F(0.0), F(0.0)};
  if (!std::isfinite(val(a.hi))) return DD<F>{a.hi, F(0.0)};
  DD<F> x = dd_from(F(1.0) / sqrt_(a.hi));
  DD<F> ax = mul(a, x);
  DD<F> hx = mul(x, dd_from(F(0.5)));
  DD<F> r = add(ax, mul(ax, sub(dd_from(F(0.5)), mul(ax, hx))));
  return mul(add(r, div(a, r)), dd_from(F(0.5)));
}
template <class F> inline DD<F> recip(DD<F> a) {  // dd.rs:80-83
  DD<F> x0 = dd_from(F(1.0) / a.hi);
  return add(x0, mul(x0, sub(dd_from(F(1.0)), mul(a, x0))));
}

// Rust `f as i64` saturates; restated here without C's undefined behaviour.
inline int64_t sat_i64(double k) {
  if (std::isnan(k)) return 0;
  if (k >= 9.2233720368547758e18) return INT64_MAX;
  if (k <= -9.2233720368547758e18) return INT64_MIN;
  return (int64_t)k;
}
inline double pow2_bits(int64_t e) {  // f64::from_bits(((1023 + k) as u64) << 52), wrapping
  uint64_t bits = ((uint64_t)(1023 + e)) << 52;
  double d; std::memcpy(&d, &bits, 8); return d;
}

template <class F> inline DD<F> dexp(DD<F> a) {  // dd.rs:106-121
  const DD<F> LN2{F(0.6931471805599453), F(2.3190468138462996e-17)};  // dd.rs:40
  if (val(a.hi) == 0.0 && val(a.lo) == 0.0) return DD<F>{F(1.0), F(0.0)};
  F kq = a.hi / LN2.hi;
  double k = std::round(val(kq));
  DD<F> r = sub(a, mul(dd_from(F(k)), LN2));
  DD<F> term{F(1.0), F(0.0)}, sum{F(1.0), F(0.0)};
  for (int i = 1; i <= 25; ++i) {
    term = div(mul(term, r), dd_from(F((double)i)));
    sum = add(sum, term);
    if (val(dabs(term).hi) < 1e-35) break;
  }
  F scale = F(pow2_bits(sat_i64(k)));
  return mul(sum, dd_from(scale));
}
template <class F> inline DD<F> dln(DD<F> a) {  // dd.rs:124-137
  if (val(a.hi) <= 0.0) return DD<F>{F(NAN), F(NAN)};
  if (val(a.hi) == 1.0 && val(a.lo) == 0.0) return DD<F>{F(0.0), F(0.0)};
  DD<F> x = dd_from(F(std::log(val(a.hi))));
  for (int i = 0; i < 3; ++i) {
    DD<F> e = dexp(neg(x));
    x = add(x, sub(mul(a, e), dd_from(F(1.0))));
  }
  return x;
}

// ---- MultiFloat<4>: crates/quad/src/multifloat ----
template <class F> struct MF { F l[4]; };

template <class F> inline MF<F> mf_from(F x) { return MF<F>{{x, F(0.0), F(0.0), F(0.0)}}; }
template <class F> inline MF<F> mf_zero() { return mf_from(F(0.0)); }

template <class F> inline void renormalize(F* limbs) {  // core.rs:57-93, N = 4
  for (int pass = 0; pass < 2; ++pass) {
    F s, e;
    two_sum(limbs[2], limbs[3], s, e);
    limbs[3] = e;
    for (int i = 1; i >= 0; --i) {
      F ns;
      two_sum(limbs[i], s, ns, e);
      s = ns;
      limbs[i + 1] = e;
    }
    limbs[0] = s;
    F out[4] = {F(0.0), F(0.0), F(0.0), F(0.0)};
    s = limbs[0];
    int k = 0;
    for (int i = 1; i < 4; ++i) {
      F ns;
      quick_two_sum(s, limbs[i], ns, e);
      s = ns;
      if (val(e) != 0.0) { out[k] = s; s = e; ++k; }
    }
    out[k] = s;
    for (int i = 0; i < 4; ++i) limbs[i] = out[i];
  }
}

template <class F> inline void renormalize_from8(const F* in, F* out4) {  // core.rs:99-168, N=4, M=8
  const int M = 8, N = 4;
  F t[8];
  for (int i = 0; i < M; ++i) t[i] = in[i];
  {
    F s, e;
    two_sum(t[M - 2], t[M - 1], s, e);
    t[M - 1] = e;
    for (int i = M - 3; i >= 0; --i) { F ns; two_sum(t[i], s, ns, e); s = ns; t[i + 1] = e; }
    t[0] = s;
    F buf[8];
    for (int i = 0; i < M; ++i) buf[i] = F(0.0);
    s = t[0];
    int k = 0;
    for (int i = 1; i < M; ++i) {
      F ns; quick_two_sum(s, t[i], ns, e); s = ns;
      if (val(e) != 0.0) { buf[k] = s; s = e; ++k; }
    }
    buf[k] = s;
    for (int i = 0; i < M; ++i) t[i] = buf[i];
  }
  {
    F s, e;
    two_sum(t[M - 2], t[M - 1], s, e);
    t[M - 1] = e;
    for (int i = M - 3; i >= 0; --i) { F ns; two_sum(t[i], s, ns, e); s = ns; t[i + 1] = e; }
    t[0] = s;
    for (int i = 0; i < N; ++i) out4[i] = F(0.0);
    s = t[0];
    int k = 0;
    for (int i = 1; i < M; ++i) {
      F ns; quick_two_sum(s, t[i], ns, e); s = ns;
      if (val(e) != 0.0) {
        out4[k] = s; s = e; ++k;
        if (k >= N) return;
      }
    }
    if (k < N) out4[k] = s;
  }
}

template <class F> inline MF<F> mf_neg(MF<F> a) {  // add.rs:4-17
  return MF<F>{{-a.l[0], -a.l[1], -a.l[2], -a.l[3]}};
}

template <class F> inline MF<F> mf_add(MF<F> a, MF<F> b) {  // add.rs:19-106
  const int N = 4, WL = 8, MAXW = 32;
  F work[8];
  for (int i = 0; i < N; ++i) two_sum(a.l[i], b.l[i], work[2 * i], work[2 * i + 1]);
  for (int i = 1; i < WL; ++i) {  // insertion sort, descending |.|
    F v = work[i];
    double mag = std::fabs(val(v));
    int j = i;
    while (j > 0 && std::fabs(val(work[j - 1])) < mag) { work[j] = work[j - 1]; --j; }
    work[j] = v;
  }
  F ex[MAXW];
  int exp_len = 0;
  for (int i = 0; i < WL; ++i) {  // grow-expansion
    F term = work[i];
    if (val(term) == 0.0) continue;
    if (exp_len == 0) { ex[0] = term; exp_len = 1; continue; }
    F carry = term;
    int j = exp_len;
    while (j > 0) {
      --j;
      F s, e;
      two_sum(ex[j], carry, s, e);
      ex[j] = e;
      carry = s;
    }
    if (val(carry) != 0.0) {
      int new_len = exp_len < MAXW ? exp_len + 1 : MAXW;
      for (int k = new_len - 1; k > 0; --k) ex[k] = ex[k - 1];
      ex[0] = carry;
      exp_len = new_len;
    }
  }
  MF<F> r = mf_zero<F>();
  for (int i = 0; i < N && i < exp_len; ++i) r.l[i] = ex[i];
  renormalize(r.l);
  return r;
}
template <class F> inline MF<F> mf_sub(MF<F> a, MF<F> b) { return mf_add(a, mf_neg(b)); }  // add.rs:108-114

template <class F> inline void mf_accumulate(F* work, int idx, F v, int limit) {  // mul.rs:31-42
  F s, e;
  two_sum(work[idx], v, s, e);
  work[idx] = s;
  int k = idx + 1;
  while (val(e) != 0.0 && k < limit) {
    F s2, e2;
    two_sum(work[k], e, s2, e2);
    work[k] = s2;
    e = e2;
    ++k;
  }
}
template <class F> inline MF<F> mf_mul(MF<F> a, MF<F> b) {  // mul.rs:12-111, N = 4
  const int N = 4, limit = 8;
  F work[16];
  for (int i = 0; i < 16; ++i) work[i] = F(0.0);
  for (int i = 0; i < N; ++i)
    for (int j = 0; j < N; ++j) {
      if (i + j >= N) break;
      F p, e;
      two_prod(a.l[i], b.l[j], p, e);
      mf_accumulate(work, i + j, p, limit);
      if (i + j + 1 < limit) mf_accumulate(work, i + j + 1, e, limit);
    }
  MF<F> r;
  renormalize_from8(work, r.l);
  return r;
}
template <class F> inline MF<F> mf_mul_f(MF<F> a, F rhs) {  // mul.rs:124-202, N = 4
  const int N = 4;
  F work[16];
  for (int i = 0; i < 16; ++i) work[i] = F(0.0);
  for (int i = 0; i < N; ++i) {
    F p, e, s, err;
    two_prod(a.l[i], rhs, p, e);
    two_sum(work[i], p, s, err);
    work[i] = s;
    if (i + 1 < 2 * N) {
      F s2, err2, s3, err3;
      two_sum(work[i + 1], e, s2, err2);
      work[i + 1] = s2;
      if (i + 2 < 2 * N) work[i + 2] = work[i + 2] + err2;
      two_sum(work[i + 1], err, s3, err3);
      work[i + 1] = s3;
      if (i + 2 < 2 * N) work[i + 2] = work[i + 2] + err3;
    }
  }
  MF<F> r;
  renormalize_from8(work, r.l);
  return r;
}
template <class F> inline MF<F> mf_recip(MF<F> a) {  // div.rs:14-41
  if (val(a.l[0]) == 0.0) {
    MF<F> r = mf_zero<F>();
    r.l[0] = F(std::signbit(val(a.l[0])) ? -INFINITY : INFINITY);
    return r;
  }
  if (!std::isfinite(val(a.l[0]))) return mf_from(F(0.0));
  MF<F> x = mf_from(F(1.0) / a.l[0]);
  MF<F> one = mf_from(F(1.0));
  const int steps = 4;  // ceil(log2 4) + 2
  for (int s = 0; s < steps; ++s) {
    MF<F> residual = mf_sub(one, mf_mul(a, x));
    x = mf_add(x, mf_mul(x, residual));
  }
  return x;
}
template <class F> inline MF<F> mf_div(MF<F> a, MF<F> b) {  // div.rs:44-77
  if (val(b.l[0]) == 0.0) {
    double sign = (!std::signbit(val(a.l[0])) == !std::signbit(val(b.l[0]))) ? 1.0 : -1.0;
    MF<F> r = mf_zero<F>();
    r.l[0] = F(val(a.l[0]) == 0.0 ? NAN : sign * INFINITY);
    return r;
  }
  if (val(a.l[0]) == 0.0) return mf_zero<F>();
  if (!std::isfinite(val(a.l[0])) || !std::isfinite(val(b.l[0]))) return mf_from(a.l[0] / b.l[0]);
  return mf_mul(a, mf_recip(b));
}
template <class F> inline MF<F> mf_sqrt(MF<F> a) {  // math.rs:22-40
  if (val(a.l[0]) < 0.0) return mf_from(F(NAN));
  if (val(a.l[0]) == 0.0) return mf_zero<F>();
  if (!std::isfinite(val(a.l[0]))) return mf_from(a.l[0]);
  MF<F> x = mf_from(sqrt_(a.l[0]));
  MF<F> half = mf_from(F(0.5));
  const int steps = 3;  // ceil(log2 4) + 1
  for (int s = 0; s < steps; ++s) x = mf_mul(mf_add(x, mf_div(a, x)), half);
  return x;
}
template <class F> inline bool mf_negligible(const MF<F>& t, double eps) {  // transcendental.rs:47-62
  double mag = std::fabs(val(t.l[0]));
  if (mag > eps) return false;
  if (mag == 0.0)
    for (int i = 1; i < 4; ++i)
      if (std::fabs(val(t.l[i])) > eps) return false;
  return true;
}
template <class F> inline MF<F> mf_exp(MF<F> a) {  // transcendental.rs:179-243, N = 4
  if (std::isnan(val(a.l[0]))) return a;
  if (val(a.l[0]) == 0.0) return mf_from(F(1.0));
  if (!std::isfinite(val(a.l[0]))) return val(a.l[0]) > 0.0 ? mf_from(F(INFINITY)) : mf_zero<F>();
  if (val(a.l[0]) > 709.0) return mf_from(F(INFINITY));
  if (val(a.l[0]) < -745.0) return mf_zero<F>();
  MF<F> ln2{{F(6.931471805599452862e-01), F(2.319046813846299558e-17),
             F(5.707708438416212066e-34), F(-3.582432210601811423e-50)}};  // transcendental.rs:4-9
  renormalize(ln2.l);  // from_limb_array -> from_limbs (transcendental.rs:27-32)
  F kq = a.l[0] / ln2.l[0];
  double k = std::round(val(kq));
  MF<F> r = mf_sub(a, mf_mul(mf_from(F(k)), ln2));
  const int m = 8;  // N * 2
  MF<F> rs = mf_div(r, mf_from(F(256.0)));
  MF<F> term = mf_from(F(1.0)), sum = mf_from(F(1.0));
  double eps = 1.0;
  for (int i = 0; i < 4; ++i) eps *= 2.220446049250313e-16;
  for (int i = 1; i <= 80; ++i) {
    term = mf_div(mf_mul(term, rs), mf_from(F((double)i)));
    sum = mf_add(sum, term);
    if (mf_negligible(term, eps)) break;
  }
  for (int i = 0; i < m; ++i) sum = mf_mul(sum, sum);
  int64_t ki = sat_i64(k);
  if (ki >= -1022 && ki <= 1023) return mf_mul(sum, mf_from(F(pow2_bits(ki))));
  int64_t half = ki / 2, other = ki - half;
  return mf_mul(mf_mul(sum, mf_from(F(pow2_bits(half)))), mf_from(F(pow2_bits(other))));
}

}  // namespace hpo
