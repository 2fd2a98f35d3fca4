// hpo.cpp -- the oracle's linear algebra, solver and overlap restatements and the C ABI
// the test harness loads.  TEST INFRASTRUCTURE ONLY (see hpo.h).
//
// Loop orders follow the reference exactly; rows are split across OpenMP threads where
// the reference uses rayon par_chunks_mut (results per element do not depend on the
// thread count, as in the reference).
#include "hpo.h"
#include "hpo_core.hpp"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cctype>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

using namespace hpo;
using D = DD<double>;
using M4 = MF<double>;

static inline D ld(const double* p, size_t i) { return D{p[2 * i], p[2 * i + 1]}; }
static inline void st(double* p, size_t i, D v) { p[2 * i] = v.hi; p[2 * i + 1] = v.lo; }
static inline M4 ldm(const double* p, size_t i) { M4 r; std::memcpy(r.l, p + 4 * i, 32); return r; }
static inline void stm(double* p, size_t i, const M4& v) { std::memcpy(p + 4 * i, v.l, 32); }
static inline hpo_dd to_c(D v) { return hpo_dd{v.hi, v.lo}; }
static inline D from_c(hpo_dd v) { return D{v.hi, v.lo}; }
static inline hpo_mf4 to_cm(const M4& v) { hpo_mf4 r; std::memcpy(r.l, v.l, 32); return r; }
static inline M4 from_cm(const hpo_mf4& v) { M4 r; std::memcpy(r.l, v.l, 32); return r; }

extern "C" {

int hpo_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

// ------------------------------------------------------------------ dd scalar ops
hpo_dd hpo_dd_add(hpo_dd a, hpo_dd b) { return to_c(add(from_c(a), from_c(b))); }
hpo_dd hpo_dd_sub(hpo_dd a, hpo_dd b) { return to_c(sub(from_c(a), from_c(b))); }
hpo_dd hpo_dd_mul(hpo_dd a, hpo_dd b) { return to_c(mul(from_c(a), from_c(b))); }
hpo_dd hpo_dd_div(hpo_dd a, hpo_dd b) { return to_c(div(from_c(a), from_c(b))); }
hpo_dd hpo_dd_add_f64(hpo_dd a, double b) { return to_c(add_f(from_c(a), b)); }
hpo_dd hpo_dd_mul_f64(hpo_dd a, double b) { return to_c(mul_f(from_c(a), b)); }
hpo_dd hpo_dd_div_f64(hpo_dd a, double b) { return to_c(div_f(from_c(a), b)); }
hpo_dd hpo_dd_sqrt(hpo_dd a) { return to_c(dsqrt(from_c(a))); }
hpo_dd hpo_dd_recip(hpo_dd a) { return to_c(recip(from_c(a))); }
hpo_dd hpo_dd_exp(hpo_dd a) { return to_c(dexp(from_c(a))); }
hpo_dd hpo_dd_ln(hpo_dd a) { return to_c(dln(from_c(a))); }

// op: 0 add, 1 sub, 2 mul, 3 div, 4 add_f64 (b = hi of b), 5 mul_f64, 6 div_f64
void hpo_dd_binop_n(int op, const double* a, const double* b, double* out, size_t n) {
  for (size_t i = 0; i < n; ++i) {
    D x = ld(a, i), y = ld(b, i), r{0, 0};
    switch (op) {
      case 0: r = add(x, y); break;
      case 1: r = sub(x, y); break;
      case 2: r = mul(x, y); break;
      case 3: r = div(x, y); break;
      case 4: r = add_f(x, y.hi); break;
      case 5: r = mul_f(x, y.hi); break;
      case 6: r = div_f(x, y.hi); break;
    }
    st(out, i, r);
  }
}
// op: 0 sqrt, 1 recip, 2 exp, 3 ln, 4 neg, 5 abs
void hpo_dd_unop_n(int op, const double* a, double* out, size_t n) {
  for (size_t i = 0; i < n; ++i) {
    D x = ld(a, i), r{0, 0};
    switch (op) {
      case 0: r = dsqrt(x); break;
      case 1: r = recip(x); break;
      case 2: r = dexp(x); break;
      case 3: r = dln(x); break;
      case 4: r = neg(x); break;
      case 5: r = dabs(x); break;
    }
    st(out, i, r);
  }
}

// ------------------------------------------------------------------ MF4 scalar ops
void hpo_mf4_renormalize(double* limbs4) { renormalize(limbs4); }
void hpo_mf4_renormalize_from8(const double* in8, double* out4) { renormalize_from8(in8, out4); }
hpo_mf4 hpo_mf4_add(hpo_mf4 a, hpo_mf4 b) { return to_cm(mf_add(from_cm(a), from_cm(b))); }
hpo_mf4 hpo_mf4_sub(hpo_mf4 a, hpo_mf4 b) { return to_cm(mf_sub(from_cm(a), from_cm(b))); }
hpo_mf4 hpo_mf4_mul(hpo_mf4 a, hpo_mf4 b) { return to_cm(mf_mul(from_cm(a), from_cm(b))); }
hpo_mf4 hpo_mf4_mul_f64(hpo_mf4 a, double b) { return to_cm(mf_mul_f(from_cm(a), b)); }
hpo_mf4 hpo_mf4_recip(hpo_mf4 a) { return to_cm(mf_recip(from_cm(a))); }
hpo_mf4 hpo_mf4_div(hpo_mf4 a, hpo_mf4 b) { return to_cm(mf_div(from_cm(a), from_cm(b))); }
hpo_mf4 hpo_mf4_sqrt(hpo_mf4 a) { return to_cm(mf_sqrt(from_cm(a))); }
hpo_mf4 hpo_mf4_exp(hpo_mf4 a) { return to_cm(mf_exp(from_cm(a))); }

// op: 0 add, 1 sub, 2 mul, 3 div, 4 mul_f64 (b = limb 0 of b)
void hpo_mf4_binop_n(int op, const double* a, const double* b, double* out, size_t n) {
  for (size_t i = 0; i < n; ++i) {
    M4 x = ldm(a, i), y = ldm(b, i), r = mf_zero<double>();
    switch (op) {
      case 0: r = mf_add(x, y); break;
      case 1: r = mf_sub(x, y); break;
      case 2: r = mf_mul(x, y); break;
      case 3: r = mf_div(x, y); break;
      case 4: r = mf_mul_f(x, y.l[0]); break;
    }
    stm(out, i, r);
  }
}
// op: 0 sqrt, 1 recip, 2 exp, 3 renormalize
void hpo_mf4_unop_n(int op, const double* a, double* out, size_t n) {
  for (size_t i = 0; i < n; ++i) {
    M4 x = ldm(a, i), r = x;
    switch (op) {
      case 0: r = mf_sqrt(x); break;
      case 1: r = mf_recip(x); break;
      case 2: r = mf_exp(x); break;
      case 3: renormalize(r.l); break;
    }
    stm(out, i, r);
  }
}

// ------------------------------------------------------------------ dd linalg
// gemm: crates/quad/src/linalg.rs:150-165 (row-parallel; per element a sequential dd sum)
void hpo_gemm_rows(const double* a, const double* b, double* c, size_t m, size_t n, size_t k,
                   size_t row0, size_t nrows) {
  (void)m;
#pragma omp parallel for schedule(dynamic, 1)
  for (long ii = 0; ii < (long)nrows; ++ii) {
    size_t i = row0 + (size_t)ii;
    for (size_t j = 0; j < n; ++j) {
      D s{0.0, 0.0};
      for (size_t p = 0; p < k; ++p) s = add(s, mul(ld(a, i * k + p), ld(b, p * n + j)));
      st(c, i * n + j, s);
    }
  }
}
void hpo_gemm(const double* a, const double* b, double* c, size_t m, size_t n, size_t k) {
  hpo_gemm_rows(a, b, c, m, n, k, 0, m);
}
// gemm_atb: linalg.rs:168-182 (A stored k x m)
void hpo_gemm_atb_rows(const double* a, const double* b, double* c, size_t m, size_t n, size_t k,
                       size_t row0, size_t nrows) {
#pragma omp parallel for schedule(dynamic, 1)
  for (long ii = 0; ii < (long)nrows; ++ii) {
    size_t i = row0 + (size_t)ii;
    for (size_t j = 0; j < n; ++j) {
      D s{0.0, 0.0};
      for (size_t p = 0; p < k; ++p) s = add(s, mul(ld(a, p * m + i), ld(b, p * n + j)));
      st(c, i * n + j, s);
    }
  }
}
void hpo_gemm_atb(const double* a, const double* b, double* c, size_t m, size_t n, size_t k) {
  hpo_gemm_atb_rows(a, b, c, m, n, k, 0, m);
}

// cholesky (unblocked): linalg.rs:42-96.  Only the lower triangle is written.
int hpo_cholesky(double* a, size_t n, size_t* pivot, double* value) {
  std::vector<D> row_j(n);
  for (size_t j = 0; j < n; ++j) {
    D sum{0.0, 0.0};
    for (size_t k = 0; k < j; ++k) {
      D ljk = ld(a, j * n + k);
      sum = add(sum, mul(ljk, ljk));
    }
    D diag = sub(ld(a, j * n + j), sum);
    if (diag.hi <= 0.0) {
      if (pivot) *pivot = j;
      if (value) *value = diag.hi + diag.lo;
      return 1;
    }
    D ljj = dsqrt(diag);
    st(a, j * n + j, ljj);
    D ljj_inv = recip(ljj);
    for (size_t k = 0; k < j; ++k) row_j[k] = ld(a, j * n + k);
    size_t remaining = n - j - 1;
    if (remaining == 0) continue;
#pragma omp parallel for schedule(static) if (remaining >= 64 && j >= 4)
    for (long idx = 0; idx < (long)remaining; ++idx) {
      size_t i = j + 1 + (size_t)idx;
      D s{0.0, 0.0};
      for (size_t k = 0; k < j; ++k) s = add(s, mul(ld(a, i * n + k), row_j[k]));
      st(a, i * n + j, mul(sub(ld(a, i * n + j), s), ljj_inv));
    }
  }
  return 0;
}

// trsm_right_lt: linalg.rs:224-235 (l is jb x jb, b is m x jb, both dense row-major)
static void trsm_right_lt(const D* l, size_t jb, D* b, size_t m) {
  for (size_t j = 0; j < jb; ++j) {
    D ljj_inv = recip(l[j * jb + j]);
    for (size_t i = 0; i < m; ++i) {
      D s = b[i * jb + j];
      for (size_t k = 0; k < j; ++k) s = sub(s, mul(b[i * jb + k], l[j * jb + k]));
      b[i * jb + j] = mul(s, ljj_inv);
    }
  }
}

// cholesky_blocked: linalg.rs:238-293
int hpo_cholesky_blocked(double* a, size_t n, size_t nb, size_t* pivot, double* value) {
  if (n == 0) return 0;
  if (nb == 0) nb = 64;
  std::vector<D> diag(nb * nb), panel(n * nb);
  for (size_t jblk = 0; jblk < n; jblk += nb) {
    size_t jb = std::min(nb, n - jblk);
    for (size_t i = 0; i < jb; ++i)
      for (size_t j = 0; j < jb; ++j) diag[i * jb + j] = ld(a, (jblk + i) * n + jblk + j);
    {
      std::vector<double> tmp(2 * jb * jb);
      for (size_t i = 0; i < jb * jb; ++i) { tmp[2 * i] = diag[i].hi; tmp[2 * i + 1] = diag[i].lo; }
      size_t pv = 0; double vv = 0;
      if (hpo_cholesky(tmp.data(), jb, &pv, &vv)) {
        if (pivot) *pivot = pv + jblk;
        if (value) *value = vv;
        return 1;
      }
      for (size_t i = 0; i < jb * jb; ++i) diag[i] = D{tmp[2 * i], tmp[2 * i + 1]};
    }
    for (size_t i = 0; i < jb; ++i)
      for (size_t j = 0; j < jb; ++j) st(a, (jblk + i) * n + jblk + j, diag[i * jb + j]);
    size_t remaining = n - jblk - jb;
    if (remaining == 0) continue;
    for (size_t i = 0; i < remaining; ++i)
      for (size_t j = 0; j < jb; ++j) panel[i * jb + j] = ld(a, (jblk + jb + i) * n + jblk + j);
    trsm_right_lt(diag.data(), jb, panel.data(), remaining);
    for (size_t i = 0; i < remaining; ++i)
      for (size_t j = 0; j < jb; ++j) st(a, (jblk + jb + i) * n + jblk + j, panel[i * jb + j]);
    size_t t0 = jblk + jb;
#pragma omp parallel for schedule(dynamic, 4) if (remaining >= 64)
    for (long ii = 0; ii < (long)remaining; ++ii) {
      size_t i = (size_t)ii;
      for (size_t cj = 0; cj <= i; ++cj) {
        D s{0.0, 0.0};
        for (size_t k = 0; k < jb; ++k) s = add(s, mul(panel[i * jb + k], panel[cj * jb + k]));
        size_t idx = (t0 + i) * n + t0 + cj;
        st(a, idx, sub(ld(a, idx), s));
      }
    }
  }
  return 0;
}

// forward_solve / backward_solve / solve_cholesky: linalg.rs:99-126
void hpo_forward_solve(const double* l, const double* b, double* y, size_t n) {
  std::vector<D> v(n);
  for (size_t i = 0; i < n; ++i) v[i] = ld(b, i);
  for (size_t i = 0; i < n; ++i) {
    for (size_t j = 0; j < i; ++j) v[i] = sub(v[i], mul(ld(l, i * n + j), v[j]));
    v[i] = div(v[i], ld(l, i * n + i));
  }
  for (size_t i = 0; i < n; ++i) st(y, i, v[i]);
}
void hpo_backward_solve(const double* l, const double* y, double* x, size_t n) {
  std::vector<D> v(n);
  for (size_t i = 0; i < n; ++i) v[i] = ld(y, i);
  for (size_t ii = n; ii-- > 0;) {
    for (size_t j = ii + 1; j < n; ++j) v[ii] = sub(v[ii], mul(ld(l, j * n + ii), v[j]));
    v[ii] = div(v[ii], ld(l, ii * n + ii));
  }
  for (size_t i = 0; i < n; ++i) st(x, i, v[i]);
}
void hpo_solve_cholesky(const double* l, const double* b, double* x, size_t n) {
  std::vector<double> y(2 * n);
  hpo_forward_solve(l, b, y.data(), n);
  hpo_backward_solve(l, y.data(), x, n);
}
// matvec: linalg.rs:129-147
void hpo_matvec(const double* a, const double* x, double* y, size_t n) {
#pragma omp parallel for schedule(static) if (n >= 64)
  for (long ii = 0; ii < (long)n; ++ii) {
    size_t i = (size_t)ii;
    D s{0.0, 0.0};
    for (size_t j = 0; j < n; ++j) s = add(s, mul(ld(a, i * n + j), ld(x, j)));
    st(y, i, s);
  }
}
// cond_estimate: linalg.rs:185-194
double hpo_cond_estimate(const double* l, size_t n) {
  double dmax = 0.0, dmin = DBL_MAX;
  for (size_t i = 0; i < n; ++i) {
    D v = ld(l, i * n + i);
    double d = std::fabs(v.hi + v.lo);
    dmax = std::fmax(dmax, d);
    dmin = std::fmin(dmin, d);
  }
  double r = dmax / dmin;
  return r * r;
}

// ------------------------------------------------------------------ reductions
// dot_serial: crates/quad/src/linalg.rs:20-27 (sum += a[i] * b[i] from ZERO)
static D dot_serial(const double* a, const double* b, size_t n) {
  D s{0.0, 0.0};
  for (size_t i = 0; i < n; ++i) s = add(s, mul(ld(a, i), ld(b, i)));
  return s;
}
// dot: linalg.rs:7-18.  Below 1024 elements the serial sum; else serial sums p_i of
// 1024-element chunks combined by rayon's reduce(|| ZERO, |acc, x| acc + x).  Rayon's
// indexed bridge splits a range of len items at len / 2 (left = the first half) and folds
// an unsplit range from ZERO, joining halves as left + right; which ranges stay unsplit is
// a scheduling choice.  Fixed here as the split carried down to single items:
//   tree(lo, 1)   = ZERO + p_lo
//   tree(lo, len) = tree(lo, len / 2) + tree(lo + len / 2, len - len / 2)
// a tree rayon produces when every range is split (e.g. enough idle workers stealing).
static D dot_tree(const std::vector<D>& p, size_t lo, size_t len) {
  if (len == 1) return add(D{0.0, 0.0}, p[lo]);
  const size_t half = len / 2;
  const D l = dot_tree(p, lo, half);
  return add(l, dot_tree(p, lo + half, len - half));
}
hpo_dd hpo_dot(const double* a, const double* b, size_t n) {
  const size_t CHUNK = 1024;
  if (n < CHUNK) return to_c(dot_serial(a, b, n));
  const size_t m = (n + CHUNK - 1) / CHUNK;
  std::vector<D> p(m);
#pragma omp parallel for schedule(static)
  for (long c = 0; c < (long)m; ++c)
    p[c] = dot_serial(a + 2 * c * CHUNK, b + 2 * c * CHUNK, std::min(CHUNK, n - c * CHUNK));
  return to_c(dot_tree(p, 0, m));
}
// gemv: linalg.rs:29-40 (y_i = dot_serial(row_i, x))
void hpo_gemv(const double* a, size_t rows, size_t cols, const double* x, double* y) {
#pragma omp parallel for schedule(static)
  for (long i = 0; i < (long)rows; ++i) st(y, (size_t)i, dot_serial(a + 2 * (size_t)i * cols, x, cols));
}
// neumaier_sum_f64: crates/compensated-sum/src/lib.rs:6-21
double hpo_neumaier_sum_f64(const double* v, size_t n) {
  double sum = 0.0, comp = 0.0;
  for (size_t i = 0; i < n; ++i) {
    double t = sum + v[i];
    if (std::fabs(sum) >= std::fabs(v[i])) comp += (sum - t) + v[i];
    else comp += (v[i] - t) + sum;
    sum = t;
  }
  return sum + comp;
}
// neumaier_sum: lib.rs:24-39 (dd)
hpo_dd hpo_neumaier_sum(const double* v, size_t n) {
  D sum{0.0, 0.0}, comp{0.0, 0.0};
  for (size_t i = 0; i < n; ++i) {
    D x = ld(v, i);
    D t = add(sum, x);
    if (dabs(sum).hi >= dabs(x).hi) comp = add(comp, add(sub(sum, t), x));
    else comp = add(comp, add(sub(x, t), sum));
    sum = t;
  }
  return to_c(add(sum, comp));
}
// par_neumaier_sum_f64 / par_neumaier_sum: lib.rs:43-66 (4096-element chunk sums, in order,
// then the Neumaier sum of the partials)
double hpo_par_neumaier_sum_f64(const double* v, size_t n) {
  const size_t CHUNK = 4096;
  if (n < CHUNK) return hpo_neumaier_sum_f64(v, n);
  std::vector<double> part;
  for (size_t c0 = 0; c0 < n; c0 += CHUNK) part.push_back(hpo_neumaier_sum_f64(v + c0, std::min(CHUNK, n - c0)));
  return hpo_neumaier_sum_f64(part.data(), part.size());
}
hpo_dd hpo_par_neumaier_sum(const double* v, size_t n) {
  const size_t CHUNK = 4096;
  if (n < CHUNK) return hpo_neumaier_sum(v, n);
  std::vector<double> part;
  for (size_t c0 = 0; c0 < n; c0 += CHUNK) {
    hpo_dd p = hpo_neumaier_sum(v + 2 * c0, std::min(CHUNK, n - c0));
    part.push_back(p.hi);
    part.push_back(p.lo);
  }
  return hpo_neumaier_sum(part.data(), part.size() / 2);
}

// ------------------------------------------------------------------ eigensolvers
// dd ordering of the reference (dd.rs:349-356): by hi, then by lo
static inline bool dd_lt(D a, D b) { return a.hi < b.hi || (a.hi == b.hi && a.lo < b.lo); }
static const D ONE{1.0, 0.0};

// jacobi_eigen: crates/quad/src/linalg.rs:299-362 (cyclic sweeps, stop when the dd
// Frobenius norm of the strict upper triangle is below 1e-30, checked before each sweep).
// mistake (tests of the checks): 1 = eigenvector update with the sine's sign flipped,
// 2 = tau from (a_pp - a_qq).
int hpo_jacobi_eigen_planted(const double* a, size_t n, size_t max_iter, double* evals, double* evecs, int mistake) {
  std::vector<D> s(n * n), v(n * n, D{0.0, 0.0});
  for (size_t i = 0; i < n * n; ++i) s[i] = ld(a, i);
  for (size_t i = 0; i < n; ++i) v[i * n + i] = ONE;
  const D threshold{1e-30, 0.0};
  for (size_t it = 0; it < max_iter; ++it) {
    D off{0.0, 0.0};
    for (size_t i = 0; i < n; ++i)
      for (size_t j = i + 1; j < n; ++j) {
        D val = s[i * n + j];
        off = add(off, mul(val, val));
      }
    off = dsqrt(off);
    if (dd_lt(off, threshold)) {
      for (size_t i = 0; i < n; ++i) st(evals, i, s[i * n + i]);
      for (size_t i = 0; i < n; ++i)
        for (size_t j = 0; j < n; ++j) st(evecs, i * n + j, v[j * n + i]);
      return 0;
    }
    for (size_t p = 0; p < n; ++p)
      for (size_t q = p + 1; q < n; ++q) {
        D apq = s[p * n + q];
        if (dd_lt(dabs(apq), D{1e-60, 0.0})) continue;
        D app = s[p * n + p], aqq = s[q * n + q];
        D tau = div(mistake == 2 ? sub(app, aqq) : sub(aqq, app), mul(D{2.0, 0.0}, apq));
        D tau_abs = dabs(tau);
        D t_denom = add(tau_abs, dsqrt(add(ONE, mul(tau, tau))));
        D t = tau.hi >= 0.0 ? div(ONE, t_denom) : div(neg(ONE), t_denom);
        D c = div(ONE, dsqrt(add(ONE, mul(t, t))));
        D s_rot = mul(t, c);
        s[p * n + p] = sub(app, mul(t, apq));
        s[q * n + q] = add(aqq, mul(t, apq));
        s[p * n + q] = D{0.0, 0.0};
        s[q * n + p] = D{0.0, 0.0};
        for (size_t r = 0; r < n; ++r) {
          if (r == p || r == q) continue;
          D srp = s[r * n + p], srq = s[r * n + q];
          s[r * n + p] = sub(mul(c, srp), mul(s_rot, srq));
          s[p * n + r] = sub(mul(c, srp), mul(s_rot, srq));
          s[r * n + q] = add(mul(s_rot, srp), mul(c, srq));
          s[q * n + r] = add(mul(s_rot, srp), mul(c, srq));
        }
        D sv = mistake == 1 ? neg(s_rot) : s_rot;
        for (size_t r = 0; r < n; ++r) {
          D vrp = v[r * n + p], vrq = v[r * n + q];
          v[r * n + p] = sub(mul(c, vrp), mul(sv, vrq));
          v[r * n + q] = add(mul(sv, vrp), mul(c, vrq));
        }
      }
  }
  return 1;  // "Jacobi eigenvalue solver did not converge within max iterations"
}
int hpo_jacobi_eigen(const double* a, size_t n, size_t max_iter, double* evals, double* evecs) {
  return hpo_jacobi_eigen_planted(a, n, max_iter, evals, evecs, 0);
}

// jacobi_eigen of the solver crate: crates/solver/src/ortho.rs:61-152 (classic Jacobi on
// the largest off-diagonal |a_pq| as f64, at most 100 n^2 rotations, stop below 1e-60).
// a is destroyed; eigenvectors are the columns of v (row-major n x n).
static void jacobi_maxpivot(std::vector<D>& a, size_t n, std::vector<D>& evals, std::vector<D>& v) {
  v.assign(n * n, D{0.0, 0.0});
  for (size_t i = 0; i < n; ++i) v[i * n + i] = ONE;
  const size_t max_iter = 100 * n * n;
  for (size_t it = 0; it < max_iter; ++it) {
    double max_val = 0.0;
    size_t p = 0, q = 1;
    for (size_t i = 0; i < n; ++i)
      for (size_t j = i + 1; j < n; ++j) {
        double val = std::fabs(a[i * n + j].hi + a[i * n + j].lo);
        if (val > max_val) {
          max_val = val;
          p = i;
          q = j;
        }
      }
    if (max_val < 1e-60) break;
    D app = a[p * n + p], aqq = a[q * n + q], apq = a[p * n + q];
    D diff = sub(aqq, app);
    D c = ONE, s = D{0.0, 0.0};
    if (!(std::fabs(apq.hi + apq.lo) < 1e-70)) {
      D tau = div(diff, add(apq, apq));
      D tau_abs = dabs(tau);
      D t = tau.hi >= 0.0 ? div(ONE, add(tau_abs, dsqrt(add(ONE, mul(tau, tau)))))
                          : div(neg(ONE), add(tau_abs, dsqrt(add(ONE, mul(tau, tau)))));
      c = div(ONE, dsqrt(add(ONE, mul(t, t))));
      s = mul(t, c);
    }
    for (size_t j = 0; j < n; ++j) {
      if (j == p || j == q) continue;
      D ajp = a[j * n + p], ajq = a[j * n + q];
      D new_jp = sub(mul(c, ajp), mul(s, ajq));
      D new_jq = add(mul(s, ajp), mul(c, ajq));
      a[j * n + p] = new_jp;
      a[p * n + j] = new_jp;
      a[j * n + q] = new_jq;
      a[q * n + j] = new_jq;
    }
    const D two{2.0, 0.0};
    D cs2apq = mul(mul(mul(two, c), s), apq);
    D new_pp = add(sub(mul(mul(c, c), app), cs2apq), mul(mul(s, s), aqq));
    D new_qq = add(add(mul(mul(s, s), app), cs2apq), mul(mul(c, c), aqq));
    a[p * n + p] = new_pp;
    a[q * n + q] = new_qq;
    a[p * n + q] = D{0.0, 0.0};
    a[q * n + p] = D{0.0, 0.0};
    for (size_t i = 0; i < n; ++i) {
      D vip = v[i * n + p], viq = v[i * n + q];
      v[i * n + p] = sub(mul(c, vip), mul(s, viq));
      v[i * n + q] = add(mul(s, vip), mul(c, viq));
    }
  }
  evals.resize(n);
  for (size_t i = 0; i < n; ++i) evals[i] = a[i * n + i];
}
void hpo_jacobi_eigen_maxpivot(const double* a, size_t n, double* evals, double* evecs) {
  std::vector<D> w(n * n), ev, v;
  for (size_t i = 0; i < n * n; ++i) w[i] = ld(a, i);
  jacobi_maxpivot(w, n, ev, v);
  for (size_t i = 0; i < n; ++i) st(evals, i, ev[i]);
  for (size_t i = 0; i < n * n; ++i) st(evecs, i, v[i]);
}

// ortho.rs:24-56 gemm_atb / gemm_rect (zero entries of A skipped, sums in k order)
static std::vector<D> ortho_gemm_atb(const std::vector<D>& a, const std::vector<D>& b, size_t m, size_t n, size_t p) {
  std::vector<D> c(n * p, D{0.0, 0.0});
  for (size_t k = 0; k < m; ++k)
    for (size_t i = 0; i < n; ++i) {
      D aki = a[k * n + i];
      if (aki.hi == 0.0 && aki.lo == 0.0) continue;
      for (size_t j = 0; j < p; ++j) c[i * p + j] = add(c[i * p + j], mul(aki, b[k * p + j]));
    }
  return c;
}
static std::vector<D> ortho_gemm_rect(const std::vector<D>& a, const std::vector<D>& b, size_t m, size_t n,
                                      size_t p) {
  std::vector<D> c(m * p, D{0.0, 0.0});
  for (size_t i = 0; i < m; ++i)
    for (size_t k = 0; k < n; ++k) {
      D aik = a[i * n + k];
      if (aik.hi == 0.0 && aik.lo == 0.0) continue;
      for (size_t j = 0; j < p; ++j) c[i * p + j] = add(c[i * p + j], mul(aik, b[k * p + j]));
    }
  return c;
}

// canonical_orthogonalize: ortho.rs:158-193.  x: n x n_orth (row-major), returns n_orth.
static size_t canonical_ortho(const std::vector<D>& s, size_t n, double threshold, std::vector<D>& x) {
  std::vector<D> w = s, evals, evecs;
  jacobi_maxpivot(w, n, evals, evecs);
  std::vector<std::pair<size_t, D>> surv;
  for (size_t i = 0; i < n; ++i)
    if (evals[i].hi + evals[i].lo > threshold) surv.push_back({i, evals[i]});
  std::stable_sort(surv.begin(), surv.end(), [](const std::pair<size_t, D>& a, const std::pair<size_t, D>& b) {
    return (a.second.hi + a.second.lo) < (b.second.hi + b.second.lo);
  });
  const size_t m = surv.size();
  x.assign(n * m, D{0.0, 0.0});
  for (size_t k = 0; k < m; ++k) {
    D inv_sqrt = recip(dsqrt(surv[k].second));
    for (size_t i = 0; i < n; ++i) x[i * m + k] = mul(evecs[i * n + surv[k].first], inv_sqrt);
  }
  return m;
}
size_t hpo_canonical_orthogonalize(const double* s, size_t n, double threshold, double* x) {
  std::vector<D> sv(n * n), xv;
  for (size_t i = 0; i < n * n; ++i) sv[i] = ld(s, i);
  size_t m = canonical_ortho(sv, n, threshold, xv);
  for (size_t i = 0; i < n * m; ++i) st(x, i, xv[i]);
  return m;
}

// solve_gen_eigen: ortho.rs:201-248.  evals: n_basis dd (ascending), c: n x n_basis.
size_t hpo_solve_gen_eigen(const double* f, const double* s, size_t n, double threshold, double* evals, double* c) {
  std::vector<D> fv(n * n), sv(n * n), x;
  for (size_t i = 0; i < n * n; ++i) {
    fv[i] = ld(f, i);
    sv[i] = ld(s, i);
  }
  const size_t m = canonical_ortho(sv, n, threshold, x);
  std::vector<D> tmp = ortho_gemm_rect(fv, x, n, n, m);
  std::vector<D> fp = ortho_gemm_atb(x, tmp, n, m, m);
  for (size_t i = 0; i < m; ++i)
    for (size_t j = i + 1; j < m; ++j) {
      D avg = mul(add(fp[i * m + j], fp[j * m + i]), D{0.5, 0.0});
      fp[i * m + j] = avg;
      fp[j * m + i] = avg;
    }
  std::vector<D> ev, evp;
  jacobi_maxpivot(fp, m, ev, evp);
  std::vector<size_t> order(m);
  for (size_t i = 0; i < m; ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(),
                   [&](size_t a, size_t b) { return (ev[a].hi + ev[a].lo) < (ev[b].hi + ev[b].lo); });
  std::vector<D> sorted(m * m);
  for (size_t nk = 0; nk < m; ++nk)
    for (size_t i = 0; i < m; ++i) sorted[i * m + nk] = evp[i * m + order[nk]];
  std::vector<D> cv = ortho_gemm_rect(x, sorted, n, m, m);
  for (size_t k = 0; k < m; ++k) st(evals, k, ev[order[k]]);
  for (size_t i = 0; i < n * m; ++i) st(c, i, cv[i]);
  return m;
}

// ------------------------------------------------------------------ MF4 linalg
// mf_gemm: multifloat/linalg.rs:39-61
void hpo_mf4_gemm_rows(const double* a, const double* b, double* c, size_t m, size_t n, size_t k,
                       size_t row0, size_t nrows) {
  (void)m;
#pragma omp parallel for schedule(dynamic, 1)
  for (long ii = 0; ii < (long)nrows; ++ii) {
    size_t i = row0 + (size_t)ii;
    for (size_t j = 0; j < n; ++j) {
      M4 s = mf_zero<double>();
      for (size_t p = 0; p < k; ++p) s = mf_add(s, mf_mul(ldm(a, i * k + p), ldm(b, p * n + j)));
      stm(c, i * n + j, s);
    }
  }
}
void hpo_mf4_gemm(const double* a, const double* b, double* c, size_t m, size_t n, size_t k) {
  hpo_mf4_gemm_rows(a, b, c, m, n, k, 0, m);
}
// mf_gemm_atb: multifloat/linalg.rs:64-85
void hpo_mf4_gemm_atb(const double* a, const double* b, double* c, size_t m, size_t n, size_t k) {
#pragma omp parallel for schedule(dynamic, 1)
  for (long ii = 0; ii < (long)m; ++ii) {
    size_t i = (size_t)ii;
    for (size_t j = 0; j < n; ++j) {
      M4 s = mf_zero<double>();
      for (size_t p = 0; p < k; ++p) s = mf_add(s, mf_mul(ldm(a, p * m + i), ldm(b, p * n + j)));
      stm(c, i * n + j, s);
    }
  }
}
// mf_cholesky: multifloat/linalg.rs:89-141.  Returns 0 ok, 1 not positive definite.
int hpo_mf4_cholesky(double* a, size_t n) {
  std::vector<M4> row_j(n);
  for (size_t j = 0; j < n; ++j) {
    M4 sum = mf_zero<double>();
    for (size_t k = 0; k < j; ++k) { M4 v = ldm(a, j * n + k); sum = mf_add(sum, mf_mul(v, v)); }
    M4 dg = mf_sub(ldm(a, j * n + j), sum);
    if (dg.l[0] <= 0.0) return 1;
    M4 ljj = mf_sqrt(dg);
    stm(a, j * n + j, ljj);
    M4 inv = mf_recip(ljj);
    for (size_t k = 0; k < j; ++k) row_j[k] = ldm(a, j * n + k);
    size_t remaining = n - j - 1;
    if (remaining == 0) continue;
#pragma omp parallel for schedule(static) if (remaining >= 64 && j >= 4)
    for (long idx = 0; idx < (long)remaining; ++idx) {
      size_t i = j + 1 + (size_t)idx;
      M4 s = mf_zero<double>();
      for (size_t k = 0; k < j; ++k) s = mf_add(s, mf_mul(ldm(a, i * n + k), row_j[k]));
      stm(a, i * n + j, mf_mul(mf_sub(ldm(a, i * n + j), s), inv));
    }
  }
  return 0;
}
// mf_cholesky_solve: multifloat/linalg.rs:145-172
void hpo_mf4_cholesky_solve(const double* l, const double* b, double* x, size_t n) {
  std::vector<M4> y(n);
  for (size_t i = 0; i < n; ++i) y[i] = ldm(b, i);
  for (size_t i = 0; i < n; ++i) {
    for (size_t j = 0; j < i; ++j) y[i] = mf_sub(y[i], mf_mul(ldm(l, i * n + j), y[j]));
    y[i] = mf_div(y[i], ldm(l, i * n + i));
  }
  for (size_t ii = n; ii-- > 0;) {
    for (size_t j = ii + 1; j < n; ++j) y[ii] = mf_sub(y[ii], mf_mul(ldm(l, j * n + ii), y[j]));
    y[ii] = mf_div(y[ii], ldm(l, ii * n + ii));
  }
  for (size_t i = 0; i < n; ++i) stm(x, i, y[i]);
}
// mf_gemv: multifloat/linalg.rs:17-36 (A is n x m)
void hpo_mf4_gemv(const double* a, const double* x, double* y, size_t n, size_t m) {
#pragma omp parallel for schedule(static)
  for (long ii = 0; ii < (long)n; ++ii) {
    size_t i = (size_t)ii;
    M4 s = mf_zero<double>();
    for (size_t j = 0; j < m; ++j) s = mf_add(s, mf_mul(ldm(a, i * m + j), ldm(x, j)));
    stm(y, i, s);
  }
}

// ------------------------------------------------------------------ solver crate
// norm_f128: solver/src/lib.rs:243-249
double hpo_norm_f128(const double* x, size_t n) {
  D s{0.0, 0.0};
  for (size_t i = 0; i < n; ++i) { D v = ld(x, i); s = add(s, mul(v, v)); }
  D r = dsqrt(s);
  return r.hi + r.lo;
}
// residual_norm_f128_from_f128: lib.rs:237-241
static double residual_norm_dd(const double* a, const double* x, const double* b, size_t n) {
  std::vector<double> ax(2 * n), r(2 * n);
  hpo_matvec(a, x, ax.data(), n);
  for (size_t i = 0; i < n; ++i) st(r.data(), i, sub(ld(b, i), ld(ax.data(), i)));
  return hpo_norm_f128(r.data(), n);
}
// solve_pure_f128: lib.rs:127-144
static void solve_pure_f128(const double* a, const double* b, size_t n, hpo_solution* sol,
                            double* x_dd, double* x_f64) {
  std::vector<double> l(a, a + 2 * n * n);
  size_t pv = 0; double vv = 0;
  if (hpo_cholesky(l.data(), n, &pv, &vv)) { sol->error = 1; sol->pivot = pv; return; }
  sol->kappa = hpo_cond_estimate(l.data(), n);
  hpo_solve_cholesky(l.data(), b, x_dd, n);
  for (size_t i = 0; i < n; ++i) x_f64[i] = x_dd[2 * i] + x_dd[2 * i + 1];
  sol->residual_norm = residual_norm_dd(a, x_dd, b, n);
  sol->strategy = 2;
  sol->refinement_iters = 0;
  sol->error = 0;
}
// iterative_refinement: lib.rs:75-125
static void iterative_refinement(const double* a, const double* b, const std::vector<double>& x0,
                                 size_t n, hpo_solution* sol, double* x_dd, double* x_f64) {
  std::vector<double> l(a, a + 2 * n * n);
  size_t pv = 0; double vv = 0;
  if (hpo_cholesky(l.data(), n, &pv, &vv)) { sol->error = 1; sol->pivot = pv; return; }
  sol->kappa = hpo_cond_estimate(l.data(), n);
  for (size_t i = 0; i < n; ++i) { x_dd[2 * i] = x0[i]; x_dd[2 * i + 1] = 0.0; }
  double bnorm = hpo_norm_f128(b, n);
  double best = DBL_MAX;
  int iters = 0;
  std::vector<double> ax(2 * n), r(2 * n), dx(2 * n);
  for (int it = 0; it < 30; ++it) {
    hpo_matvec(a, x_dd, ax.data(), n);
    for (size_t i = 0; i < n; ++i) st(r.data(), i, sub(ld(b, i), ld(ax.data(), i)));
    double rnorm = hpo_norm_f128(r.data(), n);
    iters += 1;
    if (rnorm < 1e-28 * bnorm) break;
    if (iters > 3 && rnorm > 0.5 * best) break;
    if (rnorm < best) best = rnorm;
    hpo_solve_cholesky(l.data(), r.data(), dx.data(), n);
    for (size_t i = 0; i < n; ++i) st(x_dd, i, add(ld(x_dd, i), ld(dx.data(), i)));
  }
  for (size_t i = 0; i < n; ++i) x_f64[i] = x_dd[2 * i] + x_dd[2 * i + 1];
  sol->residual_norm = residual_norm_dd(a, x_dd, b, n);
  sol->strategy = 1;
  sol->refinement_iters = iters;
  sol->error = 0;
}
// cholesky_f64 / solve_cholesky_f64 / cond_estimate_f64: lib.rs:169-224 (non-macOS path)
static bool cholesky_f64(const double* a, size_t n, std::vector<double>& l) {
  l.assign(a, a + n * n);
  for (size_t j = 0; j < n; ++j) {
    double sum = 0.0;
    for (size_t k = 0; k < j; ++k) sum += l[j * n + k] * l[j * n + k];
    double diag = l[j * n + j] - sum;
    if (diag <= 0.0) return false;
    double ljj = std::sqrt(diag);
    l[j * n + j] = ljj;
    double inv = 1.0 / ljj;
    for (size_t i = j + 1; i < n; ++i) {
      double s = 0.0;
      for (size_t k = 0; k < j; ++k) s += l[i * n + k] * l[j * n + k];
      l[i * n + j] = (l[i * n + j] - s) * inv;
    }
    for (size_t i = 0; i < j; ++i) l[i * n + j] = 0.0;
  }
  return true;
}
static void solve_cholesky_f64(const std::vector<double>& l, const double* b, size_t n, std::vector<double>& y) {
  y.assign(b, b + n);
  for (size_t i = 0; i < n; ++i) {
    for (size_t j = 0; j < i; ++j) y[i] -= l[i * n + j] * y[j];
    y[i] /= l[i * n + i];
  }
  for (size_t ii = n; ii-- > 0;) {
    for (size_t j = ii + 1; j < n; ++j) y[ii] -= l[j * n + ii] * y[j];
    y[ii] /= l[ii * n + ii];
  }
}
static double cond_estimate_f64(const std::vector<double>& l, size_t n) {
  double dmax = 0.0, dmin = DBL_MAX;
  for (size_t i = 0; i < n; ++i) {
    double d = std::fabs(l[i * n + i]);
    dmax = std::fmax(dmax, d);
    dmin = std::fmin(dmin, d);
  }
  double r = dmax / dmin;
  return r * r;
}
// solve_spd: lib.rs:38-68
void hpo_solve_spd(const double* a, const double* b, size_t n, hpo_solution* sol, double* x_dd, double* x_f64) {
  std::memset(sol, 0, sizeof(*sol));
  std::vector<double> a128(2 * n * n, 0.0), b128(2 * n, 0.0);
  for (size_t i = 0; i < n * n; ++i) a128[2 * i] = a[i];
  for (size_t i = 0; i < n; ++i) b128[2 * i] = b[i];
  std::vector<double> l64, x64;
  if (cholesky_f64(a, n, l64)) {
    solve_cholesky_f64(l64, b, n, x64);
    double kappa = cond_estimate_f64(l64, n);
    if (kappa < 1e10) {
      for (size_t i = 0; i < n; ++i) { x_dd[2 * i] = x64[i]; x_dd[2 * i + 1] = 0.0; x_f64[i] = x64[i]; }
      sol->kappa = kappa;
      sol->residual_norm = residual_norm_dd(a128.data(), x_dd, b128.data(), n);
      sol->strategy = 0;
      sol->refinement_iters = 0;
      return;
    }
    iterative_refinement(a128.data(), b128.data(), x64, n, sol, x_dd, x_f64);
    return;
  }
  solve_pure_f128(a128.data(), b128.data(), n, sol, x_dd, x_f64);
}
// solve_spd_f128: lib.rs:71-73
void hpo_solve_spd_f128(const double* a_dd, const double* b_dd, size_t n, hpo_solution* sol, double* x_dd, double* x_f64) {
  std::memset(sol, 0, sizeof(*sol));
  solve_pure_f128(a_dd, b_dd, n, sol, x_dd, x_f64);
}

// ------------------------------------------------------------------ overlap crate
static const double PI_F64 = 3.141592653589793;
// libm exp, optionally moved by a deterministic pseudo-random number of ulps in
// [-shift, shift] (a function of x's bits).  Shift 0 (the default) is libm's exp, as the
// reference calls it; a non-zero shift stands in for another correctly-rounded-or-nearly
// exp (the device library's) when the tests check the per-entry error bound of the
// mixed-tier assembly.
static int g_exp_ulp_shift = 0;
void hpo_set_exp_ulp_shift(int ulps) { g_exp_ulp_shift = ulps; }
static double oexp(double x) {
  double e = std::exp(x);
  if (g_exp_ulp_shift == 0 || !(e > 0.0) || std::isinf(e)) return e;
  uint64_t b;
  std::memcpy(&b, &x, 8);
  b ^= b >> 29;
  b *= 0xbf58476d1ce4e5b9ull;
  b ^= b >> 32;
  int k = (int)(b % (uint64_t)(2 * g_exp_ulp_shift + 1)) - g_exp_ulp_shift;
  for (; k > 0; --k) e = std::nextafter(e, INFINITY);
  for (; k < 0; ++k) e = std::nextafter(e, 0.0);
  return e;
}
// overlap_primitive_f64: overlap/src/integral.rs:15-25
double hpo_overlap_primitive_f64(double ai, const double* ci, double aj, const double* cj) {
  double gamma = ai + aj;
  double mu = ai * aj / gamma;
  double dx = ci[0] - cj[0], dy = ci[1] - cj[1], dz = ci[2] - cj[2];
  double r2 = dx * dx + dy * dy + dz * dz;  // integral.rs:354-360
  double sq = std::sqrt(PI_F64 / gamma);
  double prefactor = (sq * sq) * sq;         // powi(3)
  return prefactor * oexp(-mu * r2);
}
// overlap_primitive_f128: integral.rs:29-56 (f64 pi, f64 centre differences, libm exp)
hpo_dd hpo_overlap_primitive_f128(double ai, const double* ci, double aj, const double* cj) {
  D a = dd_from(ai), b = dd_from(aj);
  D gamma = add(a, b);
  D mu = div(mul(a, b), gamma);
  D dx = dd_from(ci[0] - cj[0]), dy = dd_from(ci[1] - cj[1]), dz = dd_from(ci[2] - cj[2]);
  D r2 = add(add(mul(dx, dx), mul(dy, dy)), mul(dz, dz));
  D pi = dd_from(PI_F64);
  D ratio = div(pi, gamma);
  D prefactor = mul(ratio, dsqrt(ratio));
  D nmr2 = mul(sub(D{0.0, 0.0}, mu), r2);
  D exp_hi = dd_from(oexp(nmr2.hi));
  D exp_val = mul(exp_hi, add(D{1.0, 0.0}, dd_from(nmr2.lo)));
  return to_c(mul(prefactor, exp_val));
}
// screen_shell_pair for single-primitive S shells: screen.rs:57-101
int hpo_screen_s_pair(double ai, const double* ci, double aj, const double* cj, double diffuse_threshold,
                      double ratio_threshold, double eps_skip, double safe_distance) {
  double dx = ci[0] - cj[0], dy = ci[1] - cj[1], dz = ci[2] - cj[2];
  double r2 = dx * dx + dy * dy + dz * dz;
  double amin_a = std::fmin(INFINITY, ai), amin_b = std::fmin(INFINITY, aj);
  double alpha_sum = amin_a + amin_b;
  double gamma = amin_a * amin_b / alpha_sum;
  double exponent = gamma * r2;
  if (exponent > -std::log(eps_skip)) return 2;
  bool ad = amin_a < diffuse_threshold, bd = amin_b < diffuse_threshold;
  if (!ad && !bd) return 0;
  double amax_a = std::fmax(0.0, ai), amax_b = std::fmax(0.0, aj);
  double ratio = std::fmax(amax_a / amin_b, amax_b / amin_a);
  double r = std::sqrt(r2);
  if (ratio > ratio_threshold && r < safe_distance) return 1;
  if (ad && bd && r < safe_distance) return 1;
  return 0;
}
// assemble_overlap_mixed over single-primitive S shells: assemble.rs:35-85 with
// overlap_contracted_{f64,f128} s-s branches (integral.rs:93-142)
void hpo_assemble_overlap_s(const double* centers, const double* alpha, const double* coeff, size_t n,
                            const double* params4, double* s, size_t* counts) {
  std::memset(s, 0, sizeof(double) * n * n);
  counts[0] = counts[1] = counts[2] = 0;
  for (size_t i = 0; i < n; ++i)
    for (size_t j = 0; j <= i; ++j) {
      const double* ci = centers + 3 * i;
      const double* cj = centers + 3 * j;
      int tier = hpo_screen_s_pair(alpha[i], ci, alpha[j], cj, params4[0], params4[1], params4[2], params4[3]);
      double v = 0.0;
      if (tier == 0) {
        double sp = hpo_overlap_primitive_f64(alpha[i], ci, alpha[j], cj);
        v = coeff[i] * coeff[j] * sp;
      } else if (tier == 1) {
        D sp = from_c(hpo_overlap_primitive_f128(alpha[i], ci, alpha[j], cj));
        D sum = add(D{0.0, 0.0}, mul(mul(dd_from(coeff[i]), dd_from(coeff[j])), sp));
        v = sum.hi + sum.lo;
      }
      counts[tier] += 1;
      if (tier == 2) continue;
      s[i * n + j] = v;
      if (i != j) s[j * n + i] = v;
    }
}

// ---- full assemble_overlap_mixed: contracted s/p/d/f shells (crates/overlap/src) ----
// A basis is flattened: shell s has centre centers[3s..3s+2], angular momentum l[s] in
// 0..3 and primitives prim_off[s] .. prim_off[s+1]-1 (alpha[], coeff[]).
struct ShellView {
  const double* c;
  int l;
  long p0, p1;
};
static double shell_alpha_min(const ShellView& s, const double* alpha) {  // basis.rs:40-43
  double m = INFINITY;
  for (long p = s.p0; p < s.p1; ++p) m = std::fmin(m, alpha[p]);
  return m;
}
static double shell_alpha_max(const ShellView& s, const double* alpha) {  // basis.rs:45-48
  double m = 0.0;
  for (long p = s.p0; p < s.p1; ++p) m = std::fmax(m, alpha[p]);
  return m;
}
// screen_shell_pair: screen.rs:57-101 (0 F64, 1 F128, 2 Skip)
static int screen_pair(const ShellView& a, const ShellView& b, const double* alpha, const double* prm) {
  double dx = a.c[0] - b.c[0], dy = a.c[1] - b.c[1], dz = a.c[2] - b.c[2];
  double r2 = dx * dx + dy * dy + dz * dz;
  double amin_a = shell_alpha_min(a, alpha), amin_b = shell_alpha_min(b, alpha);
  double alpha_sum = amin_a + amin_b;
  double gamma = amin_a * amin_b / alpha_sum;
  double exponent = gamma * r2;
  if (exponent > -std::log(prm[2])) return 2;
  bool ad = amin_a < prm[0], bd = amin_b < prm[0];
  if (!ad && !bd) return 0;
  double amax_a = shell_alpha_max(a, alpha), amax_b = shell_alpha_max(b, alpha);
  double ratio = std::fmax(amax_a / amin_b, amax_b / amin_a);
  double r = std::sqrt(r2);
  if (ratio > prm[1] && r < prm[3]) return 1;
  if (ad && bd && r < prm[3]) return 1;
  return 0;
}
// cart_components: integral.rs:220-231
static int cart_components(int l, int (*out)[3]) {
  int n = 0;
  for (int lx = l; lx >= 0; --lx)
    for (int ly = l - lx; ly >= 0; --ly) {
      out[n][0] = lx;
      out[n][1] = ly;
      out[n][2] = l - lx - ly;
      ++n;
    }
  return n;
}
// obara_saika_1d_f64: integral.rs:158-185 (s[i][j], nb = lb + 1)
static void os1d_f64(int la, int lb, double pa, double pb, double gamma, double* s) {
  const int nb = lb + 1;
  const double oo2g = 0.5 / gamma;
  for (int k = 0; k < (la + 1) * nb; ++k) s[k] = 0.0;
  s[0] = 1.0;
  for (int i = 0; i < la; ++i) s[(i + 1) * nb] = pa * s[i * nb] + oo2g * (double)i * (i > 0 ? s[(i - 1) * nb] : 0.0);
  for (int j = 0; j < lb; ++j)
    for (int i = 0; i <= la; ++i) {
      double tim1 = oo2g * (double)i * (i > 0 ? s[(i - 1) * nb + j] : 0.0);
      double tjm1 = oo2g * (double)j * (j > 0 ? s[i * nb + j - 1] : 0.0);
      s[i * nb + j + 1] = pb * s[i * nb + j] + tim1 + tjm1;
    }
}
// obara_saika_1d_f128: integral.rs:187-216
static void os1d_dd(int la, int lb, D pa, D pb, D gamma, D* s) {
  const int nb = lb + 1;
  const D oo2g = div(D{0.5, 0.0}, gamma);
  for (int k = 0; k < (la + 1) * nb; ++k) s[k] = D{0.0, 0.0};
  s[0] = D{1.0, 0.0};
  for (int i = 0; i < la; ++i) {
    D fi = dd_from((double)i);
    D prev = i > 0 ? s[(i - 1) * nb] : D{0.0, 0.0};
    s[(i + 1) * nb] = add(mul(pa, s[i * nb]), mul(mul(oo2g, fi), prev));
  }
  for (int j = 0; j < lb; ++j) {
    D fj = dd_from((double)j);
    for (int i = 0; i <= la; ++i) {
      D fi = dd_from((double)i);
      D im1 = i > 0 ? s[(i - 1) * nb + j] : D{0.0, 0.0};
      D jm1 = j > 0 ? s[i * nb + j - 1] : D{0.0, 0.0};
      s[i * nb + j + 1] = add(add(mul(pb, s[i * nb + j]), mul(mul(oo2g, fi), im1)), mul(mul(oo2g, fj), jm1));
    }
  }
}
// overlap_shell_pair: integral.rs:65-142 and 233-346; block row-major ni x nj (f64)
static void shell_pair_block(const ShellView& A, const ShellView& B, const double* alpha, const double* coeff,
                             int tier, double* block) {
  int ca[10][3], cb[10][3];
  const int ni = cart_components(A.l, ca), nj = cart_components(B.l, cb);
  for (int k = 0; k < ni * nj; ++k) block[k] = 0.0;
  if (tier == 2) return;
  const int nb = B.l + 1;
  if (tier == 0) {
    if (A.l == 0 && B.l == 0) {  // integral.rs:99-110
      double sum = 0.0;
      for (long p = A.p0; p < A.p1; ++p)
        for (long q = B.p0; q < B.p1; ++q)
          sum += coeff[p] * coeff[q] * hpo_overlap_primitive_f64(alpha[p], A.c, alpha[q], B.c);
      block[0] = sum;
      return;
    }
    for (long p = A.p0; p < A.p1; ++p)  // integral.rs:233-282
      for (long q = B.p0; q < B.p1; ++q) {
        double ai = alpha[p], aj = alpha[q];
        double gamma = ai + aj;
        double mu = ai * aj / gamma;
        double dx = A.c[0] - B.c[0], dy = A.c[1] - B.c[1], dz = A.c[2] - B.c[2];
        double r2 = dx * dx + dy * dy + dz * dz;
        double sq = std::sqrt(PI_F64 / gamma);
        double prefactor = (sq * sq) * sq * oexp(-mu * r2);
        double pa[3], pb[3];
        for (int k = 0; k < 3; ++k) {
          double pk = (ai * A.c[k] + aj * B.c[k]) / gamma;
          pa[k] = pk - A.c[k];
          pb[k] = pk - B.c[k];
        }
        double sx[16], sy[16], sz[16];
        os1d_f64(A.l, B.l, pa[0], pb[0], gamma, sx);
        os1d_f64(A.l, B.l, pa[1], pb[1], gamma, sy);
        os1d_f64(A.l, B.l, pa[2], pb[2], gamma, sz);
        for (int u = 0; u < ni; ++u)
          for (int v = 0; v < nj; ++v) {
            double val = sx[ca[u][0] * nb + cb[v][0]] * sy[ca[u][1] * nb + cb[v][1]] * sz[ca[u][2] * nb + cb[v][2]] *
                         prefactor * coeff[p] * coeff[q];
            block[u * nj + v] += val;
          }
      }
    return;
  }
  if (A.l == 0 && B.l == 0) {  // integral.rs:119-136
    D sum{0.0, 0.0};
    for (long p = A.p0; p < A.p1; ++p)
      for (long q = B.p0; q < B.p1; ++q) {
        D sp = from_c(hpo_overlap_primitive_f128(alpha[p], A.c, alpha[q], B.c));
        sum = add(sum, mul(mul(dd_from(coeff[p]), dd_from(coeff[q])), sp));
      }
    block[0] = sum.hi + sum.lo;
    return;
  }
  D acc[100];
  for (int k = 0; k < ni * nj; ++k) acc[k] = D{0.0, 0.0};
  for (long p = A.p0; p < A.p1; ++p)  // integral.rs:284-346
    for (long q = B.p0; q < B.p1; ++q) {
      D ai = dd_from(alpha[p]), aj = dd_from(alpha[q]);
      D gamma = add(ai, aj);
      D mu = div(mul(ai, aj), gamma);
      D dx = dd_from(A.c[0] - B.c[0]), dy = dd_from(A.c[1] - B.c[1]), dz = dd_from(A.c[2] - B.c[2]);
      D r2 = add(add(mul(dx, dx), mul(dy, dy)), mul(dz, dz));
      D ratio = div(dd_from(PI_F64), gamma);
      D prefactor = mul(ratio, dsqrt(ratio));
      D nmr2 = mul(sub(D{0.0, 0.0}, mu), r2);
      D exp_val = mul(dd_from(oexp(nmr2.hi)), add(D{1.0, 0.0}, dd_from(nmr2.lo)));
      D total_pre = mul(prefactor, exp_val);
      D pa[3], pb[3];
      for (int k = 0; k < 3; ++k) {
        D cik = dd_from(A.c[k]), cjk = dd_from(B.c[k]);
        D pk = div(add(mul(ai, cik), mul(aj, cjk)), gamma);
        pa[k] = sub(pk, cik);
        pb[k] = sub(pk, cjk);
      }
      D sx[16], sy[16], sz[16];
      os1d_dd(A.l, B.l, pa[0], pb[0], gamma, sx);
      os1d_dd(A.l, B.l, pa[1], pb[1], gamma, sy);
      os1d_dd(A.l, B.l, pa[2], pb[2], gamma, sz);
      D cf = mul(mul(dd_from(coeff[p]), dd_from(coeff[q])), total_pre);
      for (int u = 0; u < ni; ++u)
        for (int v = 0; v < nj; ++v) {
          D val = mul(mul(mul(sx[ca[u][0] * nb + cb[v][0]], sy[ca[u][1] * nb + cb[v][1]]), sz[ca[u][2] * nb + cb[v][2]]),
                      cf);
          acc[u * nj + v] = add(acc[u * nj + v], val);
        }
    }
  for (int k = 0; k < ni * nj; ++k) block[k] = acc[k].hi + acc[k].lo;
}
// assemble_overlap_mixed: assemble.rs:35-85.  Returns n_ao; s: n_ao x n_ao (capacity given
// by the caller: sum of (l+1)(l+2)/2); counts: f64, f128, skipped shell pairs.
size_t hpo_assemble_overlap_mixed(const double* centers, const int* l, const long* prim_off, const double* alpha,
                                  const double* coeff, size_t n_shells, const double* params4, double* s,
                                  size_t* counts) {
  std::vector<long> off(n_shells + 1, 0);
  for (size_t i = 0; i < n_shells; ++i) off[i + 1] = off[i] + (l[i] + 1) * (l[i] + 2) / 2;
  const size_t n_ao = (size_t)off[n_shells];
  std::memset(s, 0, sizeof(double) * n_ao * n_ao);
  size_t c0 = 0, c1 = 0, c2 = 0;
  // parallel over the lower-triangle pairs as the reference's par_iter (assemble.rs:45-55);
  // each pair writes its own two blocks, so the result does not depend on the schedule
#pragma omp parallel for schedule(dynamic, 4) reduction(+ : c0, c1, c2)
  for (long ii = 0; ii < (long)n_shells; ++ii) {
    const size_t i = (size_t)ii;
    double block[100];
    for (size_t j = 0; j <= i; ++j) {
      ShellView A{centers + 3 * i, l[i], prim_off[i], prim_off[i + 1]};
      ShellView B{centers + 3 * j, l[j], prim_off[j], prim_off[j + 1]};
      int tier = screen_pair(A, B, alpha, params4);
      shell_pair_block(A, B, alpha, coeff, tier, block);
      (tier == 0 ? c0 : tier == 1 ? c1 : c2) += 1;
      if (tier == 2) continue;
      const int ni = (l[i] + 1) * (l[i] + 2) / 2, nj = (l[j] + 1) * (l[j] + 2) / 2;
      for (int u = 0; u < ni; ++u)
        for (int v = 0; v < nj; ++v) {
          double val = block[u * nj + v];
          s[(off[i] + u) * n_ao + (off[j] + v)] = val;
          if (i != j) s[(off[j] + v) * n_ao + (off[i] + u)] = val;
        }
    }
  }
  counts[0] = c0;
  counts[1] = c1;
  counts[2] = c2;
  return n_ao;
}

// Per-entry envelope of assemble_overlap_mixed: env_uv = sum over the primitive pairs of
// |term_pq|, the magnitudes of the terms whose signed sum is S_uv, each term formed in its
// pair's tier arithmetic as the assembly forms it (f64, or dd for the F128 tier, where
// P - A keeps the bits that f64 cancels).  A relative change d in each term's exp moves
// S_uv by about d * env_uv, so env_uv scales a per-entry bound that is tight where the
// contraction does not cancel and wider only where it does.
void hpo_assemble_overlap_envelope(const double* centers, const int* l, const long* prim_off, const double* alpha,
                                   const double* coeff, size_t n_shells, const double* params4, double* env) {
  std::vector<long> off(n_shells + 1, 0);
  for (size_t i = 0; i < n_shells; ++i) off[i + 1] = off[i] + (l[i] + 1) * (l[i] + 2) / 2;
  const size_t n_ao = (size_t)off[n_shells];
  std::memset(env, 0, sizeof(double) * n_ao * n_ao);
#pragma omp parallel for schedule(dynamic, 4)
  for (long ii = 0; ii < (long)n_shells; ++ii) {
    const size_t i = (size_t)ii;
    for (size_t j = 0; j <= i; ++j) {
      ShellView A{centers + 3 * i, l[i], prim_off[i], prim_off[i + 1]};
      ShellView B{centers + 3 * j, l[j], prim_off[j], prim_off[j + 1]};
      const int tier = screen_pair(A, B, alpha, params4);
      if (tier == 2) continue;
      int ca[10][3], cb[10][3];
      const int ni = cart_components(A.l, ca), nj = cart_components(B.l, cb), nb = B.l + 1;
      double block[100] = {};
      for (long p = A.p0; p < A.p1; ++p)
        for (long q = B.p0; q < B.p1; ++q) {
          double sx[16], sy[16], sz[16], w;
          if (tier == 0) {
            double ai = alpha[p], aj = alpha[q], gamma = ai + aj, mu = ai * aj / gamma;
            double dx = A.c[0] - B.c[0], dy = A.c[1] - B.c[1], dz = A.c[2] - B.c[2];
            double sq = std::sqrt(PI_F64 / gamma);
            w = std::fabs(coeff[p] * coeff[q]) * sq * sq * sq * std::exp(-mu * (dx * dx + dy * dy + dz * dz));
            double pa[3], pb[3];
            for (int k = 0; k < 3; ++k) {
              double pk = (ai * A.c[k] + aj * B.c[k]) / gamma;
              pa[k] = pk - A.c[k];
              pb[k] = pk - B.c[k];
            }
            os1d_f64(A.l, B.l, pa[0], pb[0], gamma, sx);
            os1d_f64(A.l, B.l, pa[1], pb[1], gamma, sy);
            os1d_f64(A.l, B.l, pa[2], pb[2], gamma, sz);
          } else {
            D ai = dd_from(alpha[p]), aj = dd_from(alpha[q]);
            D gamma = add(ai, aj);
            w = std::fabs(coeff[p] * coeff[q]) *
                std::fabs(hpo_overlap_primitive_f128(alpha[p], A.c, alpha[q], B.c).hi);
            D pa[3], pb[3], tx[16], ty[16], tz[16];
            for (int k = 0; k < 3; ++k) {
              D cik = dd_from(A.c[k]), cjk = dd_from(B.c[k]);
              D pk = div(add(mul(ai, cik), mul(aj, cjk)), gamma);
              pa[k] = sub(pk, cik);
              pb[k] = sub(pk, cjk);
            }
            os1d_dd(A.l, B.l, pa[0], pb[0], gamma, tx);
            os1d_dd(A.l, B.l, pa[1], pb[1], gamma, ty);
            os1d_dd(A.l, B.l, pa[2], pb[2], gamma, tz);
            for (int k = 0; k < (A.l + 1) * nb; ++k) {
              sx[k] = tx[k].hi;
              sy[k] = ty[k].hi;
              sz[k] = tz[k].hi;
            }
          }
          for (int u = 0; u < ni; ++u)
            for (int v = 0; v < nj; ++v)
              block[u * nj + v] +=
                  w * std::fabs(sx[ca[u][0] * nb + cb[v][0]] * sy[ca[u][1] * nb + cb[v][1]] * sz[ca[u][2] * nb + cb[v][2]]);
        }
      for (int u = 0; u < ni; ++u)
        for (int v = 0; v < nj; ++v) {
          env[(off[i] + u) * n_ao + (off[j] + v)] = block[u * nj + v];
          env[(off[j] + v) * n_ao + (off[i] + u)] = block[u * nj + v];
        }
    }
  }
}

// Normalised s-type overlap built from the reference's dd operations.  The reference's
// dd exp (dd.rs:119) wraps the exponent field for k < -1022, so for arguments below
// -708 the entry is returned as 0 (the true value is below 1e-307, under a dd ulp of
// the unit diagonal by ~275 orders of magnitude).
hpo_dd hpo_overlap_s_dd_entry(double ai, const double* ci, double aj, const double* cj) {
  D a = dd_from(ai), b = dd_from(aj);
  D g = add(a, b);
  D ab = mul(a, b);
  D p = div(ab, g);
  D ratio = div(mul_f(ab, 4.0), mul(g, g));
  D t = dsqrt(ratio);
  D pref = mul(t, dsqrt(t));
  D d[3];
  for (int c = 0; c < 3; ++c) two_sum(ci[c], -cj[c], d[c].hi, d[c].lo);
  D r2 = add(add(mul(d[0], d[0]), mul(d[1], d[1])), mul(d[2], d[2]));
  D arg = neg(mul(p, r2));
  if (arg.hi < -708.0) return hpo_dd{0.0, 0.0};
  return to_c(mul(pref, dexp(arg)));
}
void hpo_overlap_s_dd(const double* centers, const double* alpha, size_t n, double* s) {
#pragma omp parallel for schedule(dynamic, 8)
  for (long ii = 0; ii < (long)n; ++ii) {
    size_t i = (size_t)ii;
    for (size_t j = 0; j <= i; ++j) {
      hpo_dd v = hpo_overlap_s_dd_entry(alpha[i], centers + 3 * i, alpha[j], centers + 3 * j);
      s[2 * (i * n + j)] = v.hi; s[2 * (i * n + j) + 1] = v.lo;
      s[2 * (j * n + i)] = v.hi; s[2 * (j * n + i) + 1] = v.lo;
    }
  }
}

// ---- adjoints of the overlap build: F1 = log det S, F2 = b^T S^-1 b ----
// S_uu = 1 for every parameter value, and S_uv = S_vu depends on Gaussian u's parameters
// only through row / column u, so with G = dF/dS (symmetric)
//   dF/dtheta_u = 2 sum_{v != u} G_uv dS_uv/dtheta_u,
//   dS_uv/dR_uc = -2 p (R_uc - R_vc) S_uv,                         p = a_u a_v / g, g = a_u + a_v
//   dS_uv/da_u  = S_uv (3/4 (1/a_u - 2/g) - |R_u - R_v|^2 a_v^2 / g^2)
// and G = S^-1 for F1, G = -x x^T (x = S^-1 b) for F2.  Everything below is built from the
// reference's dd operations (crates/quad/src/dd.rs) and linalg (cholesky_blocked,
// solve_cholesky: crates/quad/src/linalg.rs:99-126, 238-293).

// the per-pair factors: t[c] = p (R_uc - R_vc) S_uv (c < 3), t[3] = dS_uv/da_u.
// `mistake` plants a known error for the tests that check the checkers (0: none; 1: a_u for
// a_v in the |R|^2 term; 2: the 3/4 term dropped; 3: the centre factor's sign flipped).
static void adj_pair(double au, const double* cu, double av, const double* cv, D s, D t[4], int mistake) {
  D a = dd_from(au), b = dd_from(av);
  D g = add(a, b);
  D p = div(mul(a, b), g);
  D d[3];
  for (int c = 0; c < 3; ++c) two_sum(cu[c], -cv[c], d[c].hi, d[c].lo);
  D r2 = add(add(mul(d[0], d[0]), mul(d[1], d[1])), mul(d[2], d[2]));
  D ps = mul(p, s);
  if (mistake == 3) ps = neg(ps);
  for (int c = 0; c < 3; ++c) t[c] = mul(ps, d[c]);
  D bg = div(mistake == 1 ? a : b, g);
  D t34 = mul_f(sub(recip(a), div(dd_from(2.0), g)), 0.75);
  if (mistake == 2) t34 = dd_from(0.0);
  D fa = sub(t34, mul(r2, mul(bg, bg)));
  t[3] = mul(s, fa);
}

int hpo_overlap_adjoint_dd_planted(const double* centers, const double* alpha, const double* b, size_t n,
                                   double* logdet, double* quad, double* g_ld, double* g_q, double* x_out,
                                   int mistake) {
  std::vector<double> s(2 * n * n), l;
  hpo_overlap_s_dd(centers, alpha, n, s.data());
  l = s;
  size_t piv;
  double val_;
  if (hpo_cholesky_blocked(l.data(), n, 64, &piv, &val_)) return 1;
  std::vector<double> x(2 * n), w(2 * n * n);
  hpo_solve_cholesky(l.data(), b, x.data(), n);
#pragma omp parallel for schedule(dynamic, 4)
  for (long jj = 0; jj < (long)n; ++jj) {  // column j of S^-1 = solve_cholesky(L, e_j)
    size_t j = (size_t)jj;
    std::vector<double> e(2 * n, 0.0), col(2 * n);
    e[2 * j] = 1.0;
    hpo_solve_cholesky(l.data(), e.data(), col.data(), n);
    for (size_t i = 0; i < n; ++i) { w[2 * (i * n + j)] = col[2 * i]; w[2 * (i * n + j) + 1] = col[2 * i + 1]; }
  }
  D ldv = dd_from(0.0), qv = dd_from(0.0);
  for (size_t i = 0; i < n; ++i) ldv = add(ldv, mul_f(dln(ld(l.data(), i * n + i)), 2.0));
  for (size_t i = 0; i < n; ++i) qv = add(qv, mul(ld(b, i), ld(x.data(), i)));
  st(logdet, 0, ldv);
  st(quad, 0, qv);
#pragma omp parallel for schedule(dynamic, 8)
  for (long uu = 0; uu < (long)n; ++uu) {
    size_t u = (size_t)uu;
    D a1[4], a2[4], t[4];
    for (int c = 0; c < 4; ++c) a1[c] = a2[c] = dd_from(0.0);
    for (size_t v = 0; v < n; ++v) {
      if (v == u) continue;
      adj_pair(alpha[u], centers + 3 * u, alpha[v], centers + 3 * v, ld(s.data(), u * n + v), t, mistake);
      D wv = ld(w.data(), u * n + v), xv = ld(x.data(), v);
      for (int c = 0; c < 4; ++c) {
        a1[c] = add(a1[c], mul(wv, t[c]));
        a2[c] = add(a2[c], mul(xv, t[c]));
      }
    }
    D xu = ld(x.data(), u);
    for (int c = 0; c < 3; ++c) {
      st(g_ld, 3 * u + c, mul_f(a1[c], -4.0));
      st(g_q, 3 * u + c, mul(mul_f(a2[c], 4.0), xu));
    }
    st(g_ld, 3 * n + u, mul_f(a1[3], 2.0));
    st(g_q, 3 * n + u, mul(mul_f(a2[3], -2.0), xu));
  }
  if (x_out) std::memcpy(x_out, x.data(), sizeof(double) * 2 * n);
  return 0;
}
int hpo_overlap_adjoint_dd(const double* centers, const double* alpha, const double* b, size_t n, double* logdet,
                           double* quad, double* g_ld, double* g_q, double* x_out) {
  return hpo_overlap_adjoint_dd_planted(centers, alpha, b, n, logdet, quad, g_ld, g_q, x_out, 0);
}

// f256 overlap entry from MultiFloat<4> parameters (the dd entry's formula in the
// reference's MF4 ops, crates/quad/src/multifloat); arguments below -708 give 0.
static M4 mf_overlap_entry(const M4& a, const M4* cu, const M4& b, const M4* cv) {
  M4 g = mf_add(a, b), ab = mf_mul(a, b);
  M4 p = mf_div(ab, g);
  M4 ratio = mf_div(mf_mul_f(ab, 4.0), mf_mul(g, g));
  M4 t = mf_sqrt(ratio);
  M4 pref = mf_mul(t, mf_sqrt(t));
  M4 r2 = mf_zero<double>();
  for (int c = 0; c < 3; ++c) {
    M4 d = mf_sub(cu[c], cv[c]);
    r2 = mf_add(r2, mf_mul(d, d));
  }
  M4 arg = mf_neg(mf_mul(p, r2));
  if (arg.l[0] < -708.0) return mf_zero<double>();
  return mf_mul(pref, mf_exp(arg));
}
static M4 mf_log1p_small(const M4& z) {  // log(1+z) = 2 atanh(z/(2+z)), |z| small
  M4 y = mf_div(z, mf_add(mf_from(2.0), z));
  M4 y2 = mf_mul(y, y), term = y, sum = y;
  for (int k = 1; k < 60; ++k) {
    term = mf_mul(term, y2);
    M4 q = mf_div(term, mf_from((double)(2 * k + 1)));
    sum = mf_add(sum, q);
    if (std::fabs(q.l[0]) <= 1e-70 * std::fabs(y.l[0])) break;
  }
  return mf_mul_f(sum, 2.0);
}

int hpo_overlap_fd_f256(const double* centers, const double* alpha, const double* b, size_t n, const long* coords,
                        size_t ncoords, double h, double* fd_ld, double* fd_q) {
  std::vector<M4> A(n), C(3 * n);
  for (size_t i = 0; i < n; ++i) {
    A[i] = mf_from(alpha[i]);
    for (int c = 0; c < 3; ++c) C[3 * i + c] = mf_from(centers[3 * i + c]);
  }
  std::vector<double> S(4 * n * n);
#pragma omp parallel for schedule(dynamic, 8)
  for (long ii = 0; ii < (long)n; ++ii) {
    size_t i = (size_t)ii;
    stm(S.data(), i * n + i, mf_from(1.0));
    for (size_t j = 0; j < i; ++j) {
      M4 v = mf_overlap_entry(A[i], &C[3 * i], A[j], &C[3 * j]);
      stm(S.data(), i * n + j, v);
      stm(S.data(), j * n + i, v);
    }
  }
  if (hpo_mf4_cholesky(S.data(), n)) return 1;
  const double* L = S.data();
  std::vector<double> b4(4 * n, 0.0), x4(4 * n);
  for (size_t i = 0; i < n; ++i) { b4[4 * i] = b[2 * i]; b4[4 * i + 1] = b[2 * i + 1]; }
  hpo_mf4_cholesky_solve(L, b4.data(), x4.data(), n);
#pragma omp parallel for schedule(dynamic, 1)
  for (long ci = 0; ci < (long)ncoords; ++ci) {
    long k = coords[ci];
    size_t u = (size_t)(k < 3 * (long)n ? k / 3 : k - 3 * (long)n);
    int comp = k < 3 * (long)n ? (int)(k % 3) : 3;
    std::vector<double> e(4 * n, 0.0), a(4 * n), uv(4 * n), cv(4 * n);
    e[4 * u] = 1.0;
    hpo_mf4_cholesky_solve(L, e.data(), a.data(), n);  // a = S^-1 e_u
    std::vector<M4> base(n);
    for (size_t v = 0; v < n; ++v)
      base[v] = v == u ? mf_zero<double>() : mf_overlap_entry(A[u], &C[3 * u], A[v], &C[3 * v]);
    M4 au = ldm(a.data(), u), xu = ldm(x4.data(), u);
    for (int step = 0; step < 2; ++step) {
      double hh = step == 0 ? h : 2.0 * h;
      M4 dl[2], dq[2];
      for (int sg = 0; sg < 2; ++sg) {
        M4 ap = A[u], cp[3] = {C[3 * u], C[3 * u + 1], C[3 * u + 2]};
        M4 dlt = mf_from(sg == 0 ? hh : -hh);
        if (comp == 3) ap = mf_add(ap, dlt);
        else cp[comp] = mf_add(cp[comp], dlt);
        for (size_t v = 0; v < n; ++v) {
          M4 d = v == u ? mf_zero<double>() : mf_sub(mf_overlap_entry(ap, cp, A[v], &C[3 * v]), base[v]);
          stm(uv.data(), v, d);
        }
        // S + e_u u^T + u e_u^T = S + U V^T, U = [e_u, u], V = [u, e_u]:
        // det(I + V^T S^-1 U) = (1 + c_u)^2 - (u^T c) a_u with c = S^-1 u, a = S^-1 e_u
        hpo_mf4_cholesky_solve(L, uv.data(), cv.data(), n);
        M4 cu = ldm(cv.data(), u), utc = mf_zero<double>(), q = mf_zero<double>();
        for (size_t v = 0; v < n; ++v) {
          utc = mf_add(utc, mf_mul(ldm(uv.data(), v), ldm(cv.data(), v)));
          q = mf_add(q, mf_mul(ldm(x4.data(), v), ldm(uv.data(), v)));
        }
        M4 z = mf_sub(mf_add(mf_mul_f(cu, 2.0), mf_mul(cu, cu)), mf_mul(utc, au));
        dl[sg] = mf_log1p_small(z);
        // Woodbury: b^T (S + U V^T)^-1 b - b^T S^-1 b = -[p q] M^-1 [q p]^T, p = x_u, q = x^T u,
        // M = [[1 + c_u, u^T c], [a_u, 1 + c_u]]
        M4 m11 = mf_add(mf_from(1.0), cu);
        M4 num = mf_sub(mf_sub(mf_mul(mf_mul(xu, q), mf_mul_f(m11, 2.0)), mf_mul(utc, mf_mul(xu, xu))),
                        mf_mul(au, mf_mul(q, q)));
        dq[sg] = mf_neg(mf_div(num, mf_add(mf_from(1.0), z)));
      }
      M4 gl = mf_div(mf_sub(dl[0], dl[1]), mf_from(2.0 * hh));
      M4 gq = mf_div(mf_sub(dq[0], dq[1]), mf_from(2.0 * hh));
      stm(fd_ld, (size_t)ci * 2 + step, gl);
      stm(fd_q, (size_t)ci * 2 + step, gq);
    }
  }
  return 0;
}

}  // extern "C"

// ------------------------------------------------------------------ op counting
namespace {
thread_local long g_ops = 0;
struct CD {
  double v;
  CD() : v(0) {}
  CD(double x) : v(x) {}
};
inline CD operator+(CD a, CD b) { ++g_ops; return CD(a.v + b.v); }
inline CD operator-(CD a, CD b) { ++g_ops; return CD(a.v - b.v); }
inline CD operator*(CD a, CD b) { ++g_ops; return CD(a.v * b.v); }
inline CD operator/(CD a, CD b) { ++g_ops; return CD(a.v / b.v); }
inline CD operator-(CD a) { return CD(-a.v); }
// found by argument-dependent lookup from the templates in hpo_core.hpp
inline CD fma_(CD a, CD b, CD c) { ++g_ops; return CD(std::fma(a.v, b.v, c.v)); }
inline CD sqrt_(CD a) { ++g_ops; return CD(std::sqrt(a.v)); }
inline double val(CD a) { return a.v; }
}  // namespace

extern "C" {
long hpo_count_dd_mac(void) {
  DD<CD> s{CD(0.3), CD(1e-18)}, a{CD(1.1), CD(3e-17)}, b{CD(-0.7), CD(2e-17)};
  g_ops = 0;
  s = add(s, mul(a, b));
  return g_ops;
}
long hpo_count_mf4_mac(const double* s4, const double* a4, const double* b4) {
  MF<CD> s, a, b;
  for (int i = 0; i < 4; ++i) { s.l[i] = CD(s4[i]); a.l[i] = CD(a4[i]); b.l[i] = CD(b4[i]); }
  g_ops = 0;
  s = mf_add(s, mf_mul(a, b));
  return g_ops;
}
long hpo_count_overlap_entry(double ai, const double* ci, double aj, const double* cj) {
  g_ops = 0;
  DD<CD> a = dd_from(CD(ai)), b = dd_from(CD(aj));
  DD<CD> g = add(a, b);
  DD<CD> ab = mul(a, b);
  DD<CD> p = div(ab, g);
  DD<CD> ratio = div(mul_f(ab, CD(4.0)), mul(g, g));
  DD<CD> t = dsqrt(ratio);
  DD<CD> pref = mul(t, dsqrt(t));
  DD<CD> d[3];
  for (int c = 0; c < 3; ++c) two_sum(CD(ci[c]), CD(-cj[c]), d[c].hi, d[c].lo);
  DD<CD> r2 = add(add(mul(d[0], d[0]), mul(d[1], d[1])), mul(d[2], d[2]));
  DD<CD> arg = neg(mul(p, r2));
  if (arg.hi.v < -708.0) return g_ops;
  DD<CD> r = mul(pref, dexp(arg));
  (void)r;
  return g_ops;
}
}

// ------------------------------------------------------------------ text wire formats
// Rust's `{:.Pe}`: glibc's printf("%.*e") is exact with ties to even (the default rounding
// mode), as Rust's exact-mode formatter is; only the exponent spelling differs ("e-7", "e12"
// in Rust against "e-07", "e+12"), and Rust prints NaN as "NaN" without a sign.
int hpo_format_exp(double x, int prec, int plus, char* out) {
  if (std::isnan(x)) {
    std::memcpy(out, "NaN", 3);
    return 3;
  }
  int n = 0;
  if (std::signbit(x)) out[n++] = '-';
  else if (plus) out[n++] = '+';
  if (std::isinf(x)) {
    std::memcpy(out + n, "inf", 3);
    return n + 3;
  }
  char buf[128];
  std::snprintf(buf, sizeof buf, "%.*e", prec, std::fabs(x));
  const char* e = std::strchr(buf, 'e');
  const size_t mlen = (size_t)(e - buf);
  std::memcpy(out + n, buf, mlen);
  n += (int)mlen;
  n += std::sprintf(out + n, "e%d", std::atoi(e + 1));
  return n;
}
int hpo_dd_to_string(const double* v, char* out) { return hpo_format_exp(v[0] + v[1], 32, 0, out); }
int hpo_mf4_to_string(const double* v, char* out) { return hpo_format_exp(v[0] + v[1], 64, 0, out); }
int hpo_dd_serialize_str(const double* v, char* out) {
  int n = hpo_format_exp(v[0], 18, 1, out);
  out[n++] = '+';
  return n + hpo_format_exp(v[1], 18, 1, out + n);
}
int hpo_mf4_serialize_str(const double* v, char* out) {
  int n = 0;
  for (int i = 0; i < 4; ++i) {
    if (i) out[n++] = '+';
    n += hpo_format_exp(v[i], 14, 1, out + n);
  }
  return n;
}

// Rust's f64 grammar: [+-] then inf | infinity | nan (any case), or digits [. digits]
// [(e|E) [+-] digits] with at least one mantissa digit.  A literal that passes goes to
// glibc's strtod, which rounds correctly for any number of digits.
int hpo_parse_f64(const char* s, size_t len, double* out) {
  size_t i = 0;
  bool neg = false;
  if (len && (s[0] == '+' || s[0] == '-')) {
    neg = s[0] == '-';
    i = 1;
  }
  if (i >= len) return 1;
  std::string rest(s + i, len - i);
  std::string low = rest;
  for (char& c : low) c = (char)std::tolower((unsigned char)c);
  if (low == "inf" || low == "infinity") {
    *out = neg ? -INFINITY : INFINITY;
    return 0;
  }
  if (low == "nan") {
    *out = neg ? -NAN : NAN;
    return 0;
  }
  size_t j = 0, nd = 0;
  while (j < rest.size() && std::isdigit((unsigned char)rest[j])) ++j, ++nd;
  if (j < rest.size() && rest[j] == '.') {
    ++j;
    while (j < rest.size() && std::isdigit((unsigned char)rest[j])) ++j, ++nd;
  }
  if (nd == 0) return 1;
  if (j < rest.size() && (rest[j] == 'e' || rest[j] == 'E')) {
    ++j;
    if (j < rest.size() && (rest[j] == '+' || rest[j] == '-')) ++j;
    size_t ne = 0;
    while (j < rest.size() && std::isdigit((unsigned char)rest[j])) ++j, ++ne;
    if (ne == 0) return 1;
  }
  if (j != rest.size()) return 1;
  std::string lit = (neg ? "-" : "") + rest;
  *out = std::strtod(lit.c_str(), nullptr);
  return 0;
}

// str::trim: strip Unicode White_Space code points from both ends (UTF-8 input)
static bool white_space(uint32_t c) {
  return (c >= 0x09 && c <= 0x0D) || c == 0x20 || c == 0x85 || c == 0xA0 || c == 0x1680 ||
         (c >= 0x2000 && c <= 0x200A) || c == 0x2028 || c == 0x2029 || c == 0x202F || c == 0x205F || c == 0x3000;
}
static std::vector<uint32_t> utf8_decode(const char* s, size_t len, std::vector<size_t>& start) {
  std::vector<uint32_t> cp;
  for (size_t i = 0; i < len;) {
    const unsigned char c = (unsigned char)s[i];
    const int k = c < 0x80 ? 1 : c < 0xE0 ? 2 : c < 0xF0 ? 3 : 4;
    uint32_t v = k == 1 ? c : k == 2 ? (c & 0x1F) : k == 3 ? (c & 0x0F) : (c & 0x07);
    for (int t = 1; t < k && i + t < len; ++t) v = (v << 6) | ((unsigned char)s[i + t] & 0x3F);
    cp.push_back(v);
    start.push_back(i);
    i += k;
  }
  start.push_back(len);
  return cp;
}
static void rust_trim(const char* s, size_t len, size_t& b, size_t& e) {
  std::vector<size_t> st;
  std::vector<uint32_t> cp = utf8_decode(s, len, st);
  size_t lo = 0, hi = cp.size();
  while (lo < hi && white_space(cp[lo])) ++lo;
  while (hi > lo && white_space(cp[hi - 1])) --hi;
  b = st[lo];
  e = st[hi];
}
// i32::from_str
static bool parse_i32(const std::string& t, int& out) {
  size_t i = 0;
  bool neg = false;
  if (!t.empty() && (t[0] == '+' || t[0] == '-')) {
    neg = t[0] == '-';
    i = 1;
  }
  if (i >= t.size()) return false;
  long long v = 0;
  for (; i < t.size(); ++i) {
    if (!std::isdigit((unsigned char)t[i])) return false;
    v = v * 10 + (t[i] - '0');
    if (v > 2147483648LL) return false;
  }
  if (!neg && v > 2147483647LL) return false;
  out = (int)(neg ? -v : v);
  return true;
}
// FromStr of dd.rs:371-435 and multifloat/fmt.rs:32-114, generic over the value type.
// The exponent loop runs |exp| steps; i32::MIN runs none, as `0..(-exp)` does when -exp
// wraps in a release build.
template <class T, class Ops>
static int from_str_impl(const char* s0, size_t len0, T& out) {
  size_t b, e;
  rust_trim(s0, len0, b, e);
  const char* s = s0 + b;
  size_t len = e - b;
  if (len == 0) return 1;
  bool negative = false;
  if (s[0] == '-' || s[0] == '+') {
    negative = s[0] == '-';
    ++s;
    --len;
  }
  if (len == 0) return 1;
  size_t split = len;
  for (size_t i = 0; i < len; ++i)
    if (s[i] == 'e' || s[i] == 'E') {
      split = i;
      break;
    }
  if (split == 0) return 1;
  const T ten = Ops::from(10.0);
  T acc = Ops::zero(), pow10_neg = Ops::from(1.0);
  bool saw_dot = false, any_digit = false;
  for (size_t i = 0; i < split; ++i) {
    const char c = s[i];
    if (c == '.') {
      if (saw_dot) return 1;
      saw_dot = true;
      continue;
    }
    if (c < '0' || c > '9') return 1;
    any_digit = true;
    const T digit = Ops::from((double)(c - '0'));
    if (!saw_dot) {
      acc = Ops::add(Ops::mul(acc, ten), digit);
    } else {
      pow10_neg = Ops::div(pow10_neg, ten);
      acc = Ops::add(acc, Ops::mul(digit, pow10_neg));
    }
  }
  if (!any_digit) return 1;
  if (split < len) {
    int ex;
    if (!parse_i32(std::string(s + split + 1, len - split - 1), ex)) return 1;
    if (ex > 0) {
      for (int k = 0; k < ex; ++k) acc = Ops::mul(acc, ten);
    } else if (ex < 0 && ex != INT32_MIN) {
      for (int k = 0; k < -ex; ++k) acc = Ops::div(acc, ten);
    }
  }
  out = negative ? Ops::neg(acc) : acc;
  return 0;
}
struct DdTextOps {
  static D zero() { return D{0.0, 0.0}; }
  static D from(double x) { return dd_from(x); }
  static D add(D a, D b) { return hpo::add(a, b); }
  static D mul(D a, D b) { return hpo::mul(a, b); }
  static D div(D a, D b) { return hpo::div(a, b); }
  static D neg(D a) { return hpo::neg(a); }
};
struct Mf4TextOps {
  static M4 zero() { return mf_zero<double>(); }
  static M4 from(double x) { return mf_from(x); }
  static M4 add(M4 a, M4 b) { return mf_add(a, b); }
  static M4 mul(M4 a, M4 b) { return mf_mul(a, b); }
  static M4 div(M4 a, M4 b) { return mf_div(a, b); }
  static M4 neg(M4 a) { return mf_neg(a); }
};
int hpo_dd_from_str(const char* s, size_t len, double* out) {
  D v{0.0, 0.0};
  const int rc = from_str_impl<D, DdTextOps>(s, len, v);
  out[0] = rc ? 0.0 : v.hi;
  out[1] = rc ? 0.0 : v.lo;
  return rc;
}
int hpo_mf4_from_str(const char* s, size_t len, double* out) {
  M4 v = mf_zero<double>();
  const int rc = from_str_impl<M4, Mf4TextOps>(s, len, v);
  for (int k = 0; k < 4; ++k) out[k] = rc ? 0.0 : v.l[k];
  return rc;
}
// dd.rs:496-511: split at the first '+' (index 1..len-2) that is followed by a sign and
// not preceded by 'e'/'E'; both halves must parse as f64
int hpo_dd_deserialize_str(const char* s, size_t len, double* out) {
  out[0] = out[1] = 0.0;
  for (size_t i = 1; i + 1 < len; ++i) {
    if (s[i] == '+' && (s[i + 1] == '+' || s[i + 1] == '-')) {
      if (s[i - 1] == 'e' || s[i - 1] == 'E') continue;
      double hi, lo;
      if (hpo_parse_f64(s, i, &hi) || hpo_parse_f64(s + i + 1, len - i - 1, &lo)) return 1;
      out[0] = hi;
      out[1] = lo;
      return 0;
    }
  }
  return 1;
}
// traits.rs:104-148: limb boundaries are '+' followed by a sign, not preceded by 'e'/'E';
// exactly four limbs
int hpo_mf4_deserialize_str(const char* s, size_t len, double* out) {
  double limbs[4] = {0.0, 0.0, 0.0, 0.0};
  size_t count = 0, start = 0, i = 1;
  for (int k = 0; k < 4; ++k) out[k] = 0.0;
  while (i < len) {
    const bool boundary = s[i] == '+' && i + 1 < len && (s[i + 1] == '+' || s[i + 1] == '-') && s[i - 1] != 'e' &&
                          s[i - 1] != 'E';
    if (boundary) {
      if (count >= 4) return 1;
      if (hpo_parse_f64(s + start, i - start, &limbs[count])) return 1;
      ++count;
      start = i + 1;
      i = start + 1;
    } else {
      ++i;
    }
  }
  if (count >= 4) return 1;
  if (hpo_parse_f64(s + start, len - start, &limbs[count])) return 1;
  ++count;
  if (count != 4) return 1;
  for (int k = 0; k < 4; ++k) out[k] = limbs[k];
  return 0;
}
// Bulk forms: each value is formatted / parsed independently (OpenMP over values): the
// strings go to 64-byte slots, then are packed behind their prefix-summed offsets.
size_t hpo_dd_serialize_n(const double* v, size_t n, char* text, size_t* offsets) {
  std::vector<char> slots(n * 64);
  std::vector<unsigned char> len(n);
#pragma omp parallel for schedule(static)
  for (long i = 0; i < (long)n; ++i) len[i] = (unsigned char)hpo_dd_serialize_str(v + 2 * i, slots.data() + 64 * i);
  offsets[0] = 0;
  for (size_t i = 0; i < n; ++i) offsets[i + 1] = offsets[i] + len[i];
#pragma omp parallel for schedule(static)
  for (long i = 0; i < (long)n; ++i) std::memcpy(text + offsets[i], slots.data() + 64 * i, len[i]);
  return offsets[n];
}
size_t hpo_dd_deserialize_n(const char* text, const size_t* offsets, size_t n, double* out) {
  size_t bad = 0;
#pragma omp parallel for schedule(static) reduction(+ : bad)
  for (long i = 0; i < (long)n; ++i)
    bad += hpo_dd_deserialize_str(text + offsets[i], offsets[i + 1] - offsets[i], out + 2 * i) != 0;
  return bad;
}
