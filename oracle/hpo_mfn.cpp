// MultiFloat<N> for N = 1..8: the CPU restatement of the reference's generic expansion
// arithmetic and linear algebra (test infrastructure: the checker of the device's
// hp_mfn_* entry points, never part of the product path).
//
//   renormalize<N>            crates/quad/src/multifloat/core.rs:57-93
//   renormalize_from<N, M>    core.rs:95-168
//   Add / Sub / Neg           multifloat/add.rs:4-114
//   Mul                       multifloat/mul.rs:12-111
//   recip / Div               multifloat/div.rs:14-77
//   sqrt                      multifloat/math.rs:22-40
//   mf_dot / mf_gemv / mf_gemm / mf_gemm_atb / mf_cholesky / mf_cholesky_solve
//                             multifloat/linalg.rs:7-172
//
// The loops follow the reference's control flow (dynamic lengths, early exits) rather
// than the device code's select-based unrolled forms, so the two are independent
// restatements.  The N = 4 instantiation is also checked bit for bit against the
// oracle's MultiFloat<4> (hpo_core.hpp), written separately.
#include <cmath>
#include <cstddef>
#include <cstring>
#include <vector>

namespace {

template <int N>
struct MFN {
  double l[N];
};

// FP64 operations performed by the error-free transforms (the only arithmetic of add and
// mul), counted for the work figure of the MultiFloat<N> GEMM
thread_local long t_ops = 0;

inline void tsum(double a, double b, double& s, double& e) {  // core.rs:19-25
  t_ops += 6;
  s = a + b;
  double v = s - a;
  e = (a - (s - v)) + (b - v);
}
inline void qtsum(double a, double b, double& s, double& e) {  // core.rs:29-33
  t_ops += 3;
  s = a + b;
  e = b - (s - a);
}
inline void tprod(double a, double b, double& p, double& e) {  // core.rs:37-41
  t_ops += 2;
  p = a * b;
  e = std::fma(a, b, -p);
}

// one bottom-up two_sum sweep over t[0..M) followed by the top-down quick_two_sum
// compaction into out (zeros past the emitted limbs); at most `cap` limbs are emitted
// when the reference returns early (renormalize_from's second pass)
template <int M>
void sweep(double (&t)[M], double* out, int cap) {
  double s, e;
  tsum(t[M - 2], t[M - 1], s, e);
  t[M - 1] = e;
  for (int i = M - 3; i >= 0; --i) {
    double ns;
    tsum(t[i], s, ns, e);
    s = ns;
    t[i + 1] = e;
  }
  t[0] = s;
  for (int i = 0; i < cap; ++i) out[i] = 0.0;
  s = t[0];
  int k = 0;
  for (int i = 1; i < M; ++i) {
    double ns;
    qtsum(s, t[i], ns, e);
    s = ns;
    if (e != 0.0) {
      out[k] = s;
      s = e;
      ++k;
      if (k >= cap) return;
    }
  }
  out[k] = s;
}

template <int N>
void renormalize(double (&l)[N]) {  // core.rs:57-93: two full passes
  if constexpr (N > 1) {
    for (int pass = 0; pass < 2; ++pass) {
      double out[N];
      sweep<N>(l, out, N);
      for (int i = 0; i < N; ++i) l[i] = out[i];
    }
  }
}

template <int N, int M>
void renormalize_from(const double (&in)[M], double (&out)[N]) {  // core.rs:95-168
  static_assert(M > N, "renormalize_from: M > N");
  double t[M];
  for (int i = 0; i < M; ++i) t[i] = in[i];
  double buf[M];
  sweep<M>(t, buf, M);  // first pass: compaction within the M slots
  for (int i = 0; i < M; ++i) t[i] = buf[i];
  sweep<M>(t, out, N);  // second pass: at most N limbs out
}

template <int N>
MFN<N> zero() {
  MFN<N> r;
  for (int i = 0; i < N; ++i) r.l[i] = 0.0;
  return r;
}
template <int N>
MFN<N> from_f64(double x) {
  MFN<N> r = zero<N>();
  r.l[0] = x;
  return r;
}
template <int N>
MFN<N> neg(MFN<N> a) {
  for (int i = 0; i < N; ++i) a.l[i] = -a.l[i];
  return a;
}

template <int N>
MFN<N> add(const MFN<N>& a, const MFN<N>& b) {  // add.rs:19-106
  constexpr int W = 2 * N, MAX_WORK = 32;
  double work[W];
  for (int i = 0; i < N; ++i) tsum(a.l[i], b.l[i], work[2 * i], work[2 * i + 1]);
  for (int i = 1; i < W; ++i) {  // insertion sort, descending magnitude, stable
    double val = work[i], mag = std::fabs(val);
    int j = i;
    while (j > 0 && std::fabs(work[j - 1]) < mag) {
      work[j] = work[j - 1];
      --j;
    }
    work[j] = val;
  }
  double ex[MAX_WORK] = {};
  int len = 0;
  for (int i = 0; i < W; ++i) {  // grow-expansion
    double term = work[i];
    if (term == 0.0) continue;
    if (len == 0) {
      ex[0] = term;
      len = 1;
      continue;
    }
    double carry = term;
    for (int j = len - 1; j >= 0; --j) {
      double s, e;
      tsum(ex[j], carry, s, e);
      ex[j] = e;
      carry = s;
    }
    if (carry != 0.0) {
      int nl = len < MAX_WORK ? len + 1 : MAX_WORK;
      for (int k = nl - 1; k > 0; --k) ex[k] = ex[k - 1];
      ex[0] = carry;
      len = nl;
    }
  }
  MFN<N> r = zero<N>();
  for (int i = 0; i < N && i < len; ++i) r.l[i] = ex[i];
  renormalize<N>(r.l);
  return r;
}
template <int N>
MFN<N> sub(const MFN<N>& a, const MFN<N>& b) {  // add.rs:108-114
  return add<N>(a, neg<N>(b));
}

template <int N>
MFN<N> mul(const MFN<N>& a, const MFN<N>& b) {  // mul.rs:12-111
  constexpr int LIMIT = 2 * N;
  double work[LIMIT] = {};
  auto accumulate = [&](int idx, double val) {
    double s, e;
    tsum(work[idx], val, s, e);
    work[idx] = s;
    int k = idx + 1;
    while (e != 0.0 && k < LIMIT) {
      double s2, e2;
      tsum(work[k], e, s2, e2);
      work[k] = s2;
      e = e2;
      ++k;
    }
  };
  for (int i = 0; i < N; ++i)
    for (int j = 0; j < N; ++j) {
      if (i + j >= N) break;
      double p, e;
      tprod(a.l[i], b.l[j], p, e);
      accumulate(i + j, p);
      if (i + j + 1 < LIMIT) accumulate(i + j + 1, e);
    }
  MFN<N> r;
  renormalize_from<N, LIMIT>(work, r.l);
  return r;
}

constexpr int ceil_log2(int n) {  // (n as f64).log2().ceil() for n = 1..8
  int c = 0;
  while ((1 << c) < n) ++c;
  return c;
}

template <int N>
MFN<N> recip(const MFN<N>& a) {  // div.rs:14-41
  if (a.l[0] == 0.0) return from_f64<N>(std::signbit(a.l[0]) ? -INFINITY : INFINITY);
  if (!std::isfinite(a.l[0])) return from_f64<N>(0.0);
  MFN<N> x = from_f64<N>(1.0 / a.l[0]);
  const MFN<N> one = from_f64<N>(1.0);
  const int steps = N <= 1 ? 1 : ceil_log2(N) + 2;
  for (int s = 0; s < steps; ++s) {
    MFN<N> residual = sub<N>(one, mul<N>(a, x));
    x = add<N>(x, mul<N>(x, residual));
  }
  return x;
}
template <int N>
MFN<N> div(const MFN<N>& a, const MFN<N>& b) {  // div.rs:44-77
  if (b.l[0] == 0.0) {
    const double sign = (!std::signbit(a.l[0]) == !std::signbit(b.l[0])) ? 1.0 : -1.0;
    return from_f64<N>(a.l[0] == 0.0 ? NAN : sign * INFINITY);
  }
  if (a.l[0] == 0.0) return zero<N>();
  if (!std::isfinite(a.l[0]) || !std::isfinite(b.l[0])) return from_f64<N>(a.l[0] / b.l[0]);
  return mul<N>(a, recip<N>(b));
}
template <int N>
MFN<N> sqrt_(const MFN<N>& a) {  // math.rs:22-40
  if (a.l[0] < 0.0) return from_f64<N>(NAN);
  if (a.l[0] == 0.0) return zero<N>();
  if (!std::isfinite(a.l[0])) return from_f64<N>(a.l[0]);
  MFN<N> x = from_f64<N>(std::sqrt(a.l[0]));
  const MFN<N> half = from_f64<N>(0.5);
  const int steps = ceil_log2(N) + 1;
  for (int s = 0; s < steps; ++s) x = mul<N>(add<N>(x, div<N>(a, x)), half);
  return x;
}

template <int N>
MFN<N> ld(const double* p, size_t i) {
  MFN<N> r;
  std::memcpy(r.l, p + N * i, sizeof(r.l));
  return r;
}
template <int N>
void st(double* p, size_t i, const MFN<N>& v) {
  std::memcpy(p + N * i, v.l, sizeof(v.l));
}

// op: 0 add, 1 sub, 2 mul, 3 div, 4 recip(a), 5 sqrt(a), 6 renormalize(a)
template <int N>
void elementwise(int op, const double* a, const double* b, double* out, size_t n) {
  for (size_t i = 0; i < n; ++i) {
    MFN<N> x = ld<N>(a, i), r = x;
    switch (op) {
      case 0: r = add<N>(x, ld<N>(b, i)); break;
      case 1: r = sub<N>(x, ld<N>(b, i)); break;
      case 2: r = mul<N>(x, ld<N>(b, i)); break;
      case 3: r = div<N>(x, ld<N>(b, i)); break;
      case 4: r = recip<N>(x); break;
      case 5: r = sqrt_<N>(x); break;
      case 6: renormalize<N>(r.l); break;
    }
    st<N>(out, i, r);
  }
}

// mf_gemm / mf_gemm_atb (linalg.rs:39-85): per element s = ZERO; s = s + a * b over p
template <int N>
void gemm(int trans_a, const double* a, const double* b, double* c, size_t m, size_t n, size_t k, size_t row0,
          size_t nrows) {
#pragma omp parallel for schedule(dynamic, 1)
  for (long ii = 0; ii < (long)nrows; ++ii) {
    const size_t i = row0 + (size_t)ii;
    for (size_t j = 0; j < n; ++j) {
      MFN<N> s = zero<N>();
      for (size_t p = 0; p < k; ++p)
        s = add<N>(s, mul<N>(ld<N>(a, trans_a ? p * m + i : i * k + p), ld<N>(b, p * n + j)));
      st<N>(c, i * n + j, s);
    }
  }
}
// mf_gemv (linalg.rs:17-36) and mf_dot (7-14, the one-row case)
template <int N>
void gemv(const double* a, const double* x, double* y, size_t n, size_t m) {
#pragma omp parallel for schedule(dynamic, 4)
  for (long i = 0; i < (long)n; ++i) {
    MFN<N> s = zero<N>();
    for (size_t j = 0; j < m; ++j) s = add<N>(s, mul<N>(ld<N>(a, (size_t)i * m + j), ld<N>(x, j)));
    st<N>(y, (size_t)i, s);
  }
}
// mf_cholesky (linalg.rs:89-141); returns 1 + the failing column, 0 on success
template <int N>
long cholesky(double* a, size_t n) {
  std::vector<MFN<N>> row_j(n);
  for (size_t j = 0; j < n; ++j) {
    MFN<N> sum = zero<N>();
    for (size_t k = 0; k < j; ++k) {
      MFN<N> ljk = ld<N>(a, j * n + k);
      sum = add<N>(sum, mul<N>(ljk, ljk));
    }
    MFN<N> diag = sub<N>(ld<N>(a, j * n + j), sum);
    if (diag.l[0] <= 0.0) return (long)j + 1;
    MFN<N> ljj = sqrt_<N>(diag);
    st<N>(a, j * n + j, ljj);
    MFN<N> ljj_inv = recip<N>(ljj);
    for (size_t k = 0; k < j; ++k) row_j[k] = ld<N>(a, j * n + k);
#pragma omp parallel for schedule(static) if (n - j - 1 >= 64)
    for (long i = (long)j + 1; i < (long)n; ++i) {
      MFN<N> s = zero<N>();
      for (size_t k = 0; k < j; ++k) s = add<N>(s, mul<N>(ld<N>(a, (size_t)i * n + k), row_j[k]));
      st<N>(a, (size_t)i * n + j, mul<N>(sub<N>(ld<N>(a, (size_t)i * n + j), s), ljj_inv));
    }
  }
  return 0;
}
// mf_cholesky_solve (linalg.rs:145-172)
template <int N>
void cholesky_solve(const double* l, const double* b, double* x, size_t n) {
  std::vector<MFN<N>> y(n);
  for (size_t i = 0; i < n; ++i) y[i] = ld<N>(b, i);
  for (size_t i = 0; i < n; ++i) {
    for (size_t j = 0; j < i; ++j) y[i] = sub<N>(y[i], mul<N>(ld<N>(l, i * n + j), y[j]));
    y[i] = div<N>(y[i], ld<N>(l, i * n + i));
  }
  std::vector<MFN<N>> xv(y);
  for (size_t ii = n; ii-- > 0;) {
    for (size_t j = ii + 1; j < n; ++j) xv[ii] = sub<N>(xv[ii], mul<N>(ld<N>(l, j * n + ii), xv[j]));
    xv[ii] = div<N>(xv[ii], ld<N>(l, ii * n + ii));
  }
  for (size_t i = 0; i < n; ++i) st<N>(x, i, xv[i]);
}

template <int N>
long count_mac(const double* s, const double* a, const double* b) {
  MFN<N> x = ld<N>(s, 0), y = ld<N>(a, 0), z = ld<N>(b, 0);
  t_ops = 0;
  x = add<N>(x, mul<N>(y, z));
  (void)x;
  return t_ops;
}

}  // namespace

#define HPO_MFN_DISPATCH(N, CALL)  \
  switch (N) {                     \
    case 1: CALL(1); break;        \
    case 2: CALL(2); break;        \
    case 3: CALL(3); break;        \
    case 4: CALL(4); break;        \
    case 5: CALL(5); break;        \
    case 6: CALL(6); break;        \
    case 7: CALL(7); break;        \
    case 8: CALL(8); break;        \
    default: return -1;            \
  }

extern "C" {

int hpo_mfn_elementwise(int nl, int op, const double* a, const double* b, double* out, size_t n) {
#define C_(K) elementwise<K>(op, a, b, out, n)
  HPO_MFN_DISPATCH(nl, C_)
#undef C_
  return 0;
}
int hpo_mfn_gemm_rows(int nl, int trans_a, const double* a, const double* b, double* c, size_t m, size_t n, size_t k,
                      size_t row0, size_t nrows) {
#define C_(K) gemm<K>(trans_a, a, b, c, m, n, k, row0, nrows)
  HPO_MFN_DISPATCH(nl, C_)
#undef C_
  return 0;
}
int hpo_mfn_gemv(int nl, const double* a, const double* x, double* y, size_t n, size_t m) {
#define C_(K) gemv<K>(a, x, y, n, m)
  HPO_MFN_DISPATCH(nl, C_)
#undef C_
  return 0;
}
long hpo_mfn_cholesky(int nl, double* a, size_t n) {
  long r = 0;
#define C_(K) r = cholesky<K>(a, n)
  HPO_MFN_DISPATCH(nl, C_)
#undef C_
  return r;
}
// FP64 operations of one s = s + a * b (the GEMM's MAC) on these values
long hpo_count_mfn_mac(int nl, const double* s, const double* a, const double* b) {
  long r = 0;
#define C_(K) r = count_mac<K>(s, a, b)
  HPO_MFN_DISPATCH(nl, C_)
#undef C_
  return r;
}
int hpo_mfn_cholesky_solve(int nl, const double* l, const double* b, double* x, size_t n) {
#define C_(K) cholesky_solve<K>(l, b, x, n)
  HPO_MFN_DISPATCH(nl, C_)
#undef C_
  return 0;
}

}  // extern "C"
