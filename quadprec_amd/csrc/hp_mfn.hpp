// hp_mfn.hpp -- MultiFloat<N> (N = 1..8) arithmetic for gfx950 kernels: the reference's
// generic expansion algorithms with its operation sequences, so results are bit-identical.
//   renormalize<N>          crates/quad/src/multifloat/core.rs:57-93
//   renormalize_from<N, M>  core.rs:95-168
//   Add / Sub / Neg         multifloat/add.rs:4-114
//   Mul                     multifloat/mul.rs:12-111
//   recip / Div             multifloat/div.rs:14-77
//   sqrt                    multifloat/math.rs:22-40
// Every work array is indexed with compile-time indices only: the loops have constant trip
// counts and are unrolled (the multiply's (i, j) pairs and cascade levels by template
// recursion), and the positions the data decide (the insertion point of a sorted term, the
// slot an emitted limb goes to, the length of the growing expansion) are resolved with
// selects, so the arrays stay in VGPRs.  MultiFloat<4> keeps its own tuned routines in
// hp_arith.hpp (same results: the oracle checks the generic restatement against them).
// Compile with -ffp-contract=off, as hp_arith.hpp.
#pragma once
#include <type_traits>

#include "hp_arith.hpp"

namespace hp {

template <int N>
struct mfn {
  double l[N];
};

template <int I, int END, class F>
HP_HD void static_for(F&& f) {
  if constexpr (I < END) {
    f(std::integral_constant<int, I>{});
    static_for<I + 1, END>(f);
  }
}

// out[k] = v for a data-dependent k < CAP
template <int CAP>
HP_HD void put_at(double (&out)[CAP], int k, double v) {
#pragma unroll
  for (int q = 0; q < CAP; ++q) out[q] = (k == q) ? v : out[q];
}

// One bottom-up two_sum sweep over t[0..M) and the top-down quick_two_sum compaction into
// out[0..CAP) (zeros past the emitted limbs), stopping once CAP limbs are out (the
// reference's early return in renormalize_from's second pass; with CAP = M it never does).
template <int M, int CAP>
HP_HD void mfn_sweep(double (&t)[M], double (&out)[CAP]) {
  double s, e, ns;
  two_sum(t[M - 2], t[M - 1], s, e);
  t[M - 1] = e;
#pragma unroll
  for (int i = M - 3; i >= 0; --i) {
    two_sum(t[i], s, ns, e);
    s = ns;
    t[i + 1] = e;
  }
  t[0] = s;
#pragma unroll
  for (int q = 0; q < CAP; ++q) out[q] = 0.0;
  s = t[0];
  int k = 0;
  bool done = false;
#pragma unroll
  for (int i = 1; i < M; ++i) {
    quick_two_sum(s, t[i], ns, e);
    if (!done) {
      s = ns;
      if (e != 0.0) {
        put_at<CAP>(out, k, s);
        s = e;
        ++k;
        if (k >= CAP) done = true;
      }
    }
  }
  if (!done) put_at<CAP>(out, k, s);
}

template <int N>
HP_HD void mfn_renormalize(double (&l)[N]) {  // two full passes
  if constexpr (N > 1) {
#pragma unroll
    for (int pass = 0; pass < 2; ++pass) {
      double out[N];
      mfn_sweep<N, N>(l, out);
#pragma unroll
      for (int i = 0; i < N; ++i) l[i] = out[i];
    }
  }
}

template <int N, int M>
HP_HD void mfn_renormalize_from(const double (&in)[M], double (&out)[N]) {
  static_assert(M > N, "renormalize_from: M > N");
  double t[M], buf[M];
#pragma unroll
  for (int i = 0; i < M; ++i) t[i] = in[i];
  mfn_sweep<M, M>(t, buf);
#pragma unroll
  for (int i = 0; i < M; ++i) t[i] = buf[i];
  mfn_sweep<M, N>(t, out);
}

template <int N>
HP_HD mfn<N> mfn_from(double x) {
  mfn<N> r;
#pragma unroll
  for (int i = 0; i < N; ++i) r.l[i] = 0.0;
  r.l[0] = x;
  return r;
}
template <int N>
HP_HD mfn<N> mfn_neg(mfn<N> a) {
#pragma unroll
  for (int i = 0; i < N; ++i) a.l[i] = -a.l[i];
  return a;
}

// add.rs:19-106: pairwise two_sum into 2N terms, stable insertion sort by descending
// magnitude, grow-expansion (a term touches only the limbs already present: before term i
// there are at most i), the top N limbs renormalised
template <int N>
HP_HD mfn<N> mfn_add(const mfn<N>& a, const mfn<N>& b) {
  constexpr int W = 2 * N;
  double w[W];
#pragma unroll
  for (int i = 0; i < N; ++i) two_sum(a.l[i], b.l[i], w[2 * i], w[2 * i + 1]);
#pragma unroll
  for (int i = 1; i < W; ++i) {
    const double v = w[i], mag = __builtin_fabs(v);
    bool moving = true;
#pragma unroll
    for (int j = i; j >= 1; --j) {
      const bool c = moving && (__builtin_fabs(w[j - 1]) < mag);
      const double nw = c ? w[j - 1] : (moving ? v : w[j]);
      moving = c;
      w[j] = nw;
    }
    if (moving) w[0] = v;
  }
  double ex[W];
#pragma unroll
  for (int i = 0; i < W; ++i) ex[i] = 0.0;
  int len = 0;
#pragma unroll
  for (int i = 0; i < W; ++i) {
    const double term = w[i];
    if (term == 0.0) continue;
    if (len == 0) {
      ex[0] = term;
      len = 1;
      continue;
    }
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
      const int nl = len + 1;
#pragma unroll
      for (int q = i; q >= 1; --q)
        if (q < nl) ex[q] = ex[q - 1];
      ex[0] = carry;
      len = nl;
    }
  }
  mfn<N> r;
#pragma unroll
  for (int i = 0; i < N; ++i) r.l[i] = (i < len) ? ex[i] : 0.0;
  mfn_renormalize<N>(r.l);
  return r;
}
template <int N>
HP_HD mfn<N> mfn_sub(const mfn<N>& a, const mfn<N>& b) {
  return mfn_add<N>(a, mfn_neg<N>(b));
}

// mul.rs:31-42: error-free accumulate of v into w[IDX], the error cascading upward while
// non-zero, up to the 2N-slot limit.  GUARD = false drops the e != 0 test, which gives the
// same bits when every value is finite and no slot holds -0.0: the slots start at +0.0 and
// a two_sum's sum is -0.0 only when both addends are, so none becomes -0.0, and for such a
// finite slot w, two_sum(w, +-0.0) returns (w, +0.0), leaving every later level unchanged.
template <int IDX, int LIMIT, bool GUARD = true>
HP_HD void mfn_accumulate(double (&w)[LIMIT], double v) {
  double s, e;
  two_sum(w[IDX], v, s, e);
  w[IDX] = s;
#pragma unroll
  for (int k = IDX + 1; k < LIMIT; ++k) {
    if (!GUARD || e != 0.0) {
      double s2, e2;
      two_sum(w[k], e, s2, e2);
      w[k] = s2;
      e = e2;
    }
  }
}
// mul.rs:12-111: the products a_i b_j with i + j < N in the reference's loop order (i
// outer, j inner), each two_prod's product into slot i + j and its error into i + j + 1,
// then the 2N slots renormalised to N limbs.  FAST: every limb below 2^500 in magnitude, so
// no product, sum or error overflows and the cascades run unguarded (mfn_accumulate).
template <int N, bool FAST>
HP_HD mfn<N> mfn_mul_impl(const mfn<N>& a, const mfn<N>& b) {
  constexpr int LIMIT = 2 * N;
  double w[LIMIT];
#pragma unroll
  for (int i = 0; i < LIMIT; ++i) w[i] = 0.0;
  static_for<0, N>([&](auto ic) {
    constexpr int I = decltype(ic)::value;
    static_for<0, N - I>([&](auto jc) {
      constexpr int J = decltype(jc)::value;
      double p, e;
      two_prod(a.l[I], b.l[J], p, e);
      mfn_accumulate<I + J, LIMIT, !FAST>(w, p);
      mfn_accumulate<I + J + 1, LIMIT, !FAST>(w, e);  // I + J + 1 <= N < 2N
    });
  });
  mfn<N> r;
  mfn_renormalize_from<N, LIMIT>(w, r.l);
  return r;
}
template <int N>
HP_HD mfn<N> mfn_mul(const mfn<N>& a, const mfn<N>& b) {
  constexpr double LIM = 0x1p500;
  bool fast = true;
#pragma unroll
  for (int i = 0; i < N; ++i) fast &= (__builtin_fabs(a.l[i]) < LIM) & (__builtin_fabs(b.l[i]) < LIM);
  if (fast) return mfn_mul_impl<N, true>(a, b);
  return mfn_mul_impl<N, false>(a, b);
}

constexpr int mfn_ceil_log2(int n) {  // (n as f64).log2().ceil() for n = 1..8
  int c = 0;
  while ((1 << c) < n) ++c;
  return c;
}

// div.rs:14-41: Newton from the f64 reciprocal, ceil(log2 N) + 2 steps (1 for N = 1)
template <int N>
HP_HD mfn<N> mfn_recip(const mfn<N>& a) {
  if (a.l[0] == 0.0) return mfn_from<N>(signbit(a.l[0]) ? -(double)INFINITY : (double)INFINITY);
  if (!isfinite(a.l[0])) return mfn_from<N>(0.0);
  mfn<N> x = mfn_from<N>(1.0 / a.l[0]);
  const mfn<N> one = mfn_from<N>(1.0);
  constexpr int STEPS = N <= 1 ? 1 : mfn_ceil_log2(N) + 2;
#pragma unroll 1
  for (int s = 0; s < STEPS; ++s) {
    const mfn<N> residual = mfn_sub<N>(one, mfn_mul<N>(a, x));
    x = mfn_add<N>(x, mfn_mul<N>(x, residual));
  }
  return x;
}
// div.rs:44-77: the special cases of a zero, zero-over or non-finite operand, then a * recip(b)
template <int N>
HP_HD bool mfn_div_special(const mfn<N>& a, const mfn<N>& b, mfn<N>& r) {
  if (b.l[0] == 0.0) {
    const double sign = (!signbit(a.l[0]) == !signbit(b.l[0])) ? 1.0 : -1.0;
    r = mfn_from<N>(a.l[0] == 0.0 ? (double)NAN : sign * (double)INFINITY);
    return true;
  }
  if (a.l[0] == 0.0) {
    r = mfn_from<N>(0.0);
    return true;
  }
  if (!isfinite(a.l[0]) || !isfinite(b.l[0])) {
    r = mfn_from<N>(a.l[0] / b.l[0]);
    return true;
  }
  return false;
}
template <int N>
HP_HD mfn<N> mfn_div(const mfn<N>& a, const mfn<N>& b) {
  mfn<N> r;
  if (mfn_div_special<N>(a, b, r)) return r;
  return mfn_mul<N>(a, mfn_recip<N>(b));
}
// math.rs:22-40: NaN below zero, 0 at 0, non-finite passed through, else ceil(log2 N) + 1
// steps of (x + a / x) * 0.5 from the f64 root
template <int N>
HP_HD mfn<N> mfn_sqrt(const mfn<N>& a) {
  if (a.l[0] < 0.0) return mfn_from<N>((double)NAN);
  if (a.l[0] == 0.0) return mfn_from<N>(0.0);
  if (!isfinite(a.l[0])) return mfn_from<N>(a.l[0]);
  mfn<N> x = mfn_from<N>(sqrt(a.l[0]));
  const mfn<N> half = mfn_from<N>(0.5);
  constexpr int STEPS = mfn_ceil_log2(N) + 1;
#pragma unroll 1
  for (int s = 0; s < STEPS; ++s) x = mfn_mul<N>(mfn_add<N>(x, mfn_div<N>(a, x)), half);
  return x;
}

}  // namespace hp
