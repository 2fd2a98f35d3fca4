/*
 * hpo.h -- the oracle: a CPU restatement, op for op, of the hyperprec (m0at/quadprec)
 * reference algorithms on the extended-precision path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load this library, and only as the checker / CPU baseline.
 * The product library (quadprec_amd/) never links, loads or calls it.
 *
 * Arithmetic contract: IEEE binary64, round to nearest, compiled with
 * -ffp-contract=off so that no a*b+c is fused except where the reference writes
 * mul_add (the TwoProd error term, crates/quad/src/dd.rs:33).
 *
 * Layout: a double-double value is {hi, lo} (crates/quad/src/dd.rs:8-11); arrays of
 * them are interleaved (AoS), which is how a Rust &[f128] lies in memory.  A
 * MultiFloat<4> value is 4 doubles, most significant first
 * (crates/quad/src/multifloat/core.rs:6-8).
 *
 * Parity pinning (see the test files named here):
 *   dd ops            tests/test_oracle_dd.py  (mpmath + reference unit tests)
 *   MF4 ops           tests/test_oracle_mf4.py (mpmath + reference unit tests)
 *   linalg / solver   tests/test_oracle_linalg.py (reference tests, README Hilbert table)
 *   overlap           tests/test_oracle_overlap.py (mpmath closed form, reference tests)
 */
#ifndef HPO_H
#define HPO_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct { double hi, lo; } hpo_dd;
typedef struct { double l[4]; } hpo_mf4;

/* ---- dd scalar ops (crates/quad/src/dd.rs) ---- */
hpo_dd hpo_dd_add(hpo_dd a, hpo_dd b);          /* dd.rs:147-157 */
hpo_dd hpo_dd_sub(hpo_dd a, hpo_dd b);          /* dd.rs:159-165 */
hpo_dd hpo_dd_mul(hpo_dd a, hpo_dd b);          /* dd.rs:175-183 */
hpo_dd hpo_dd_div(hpo_dd a, hpo_dd b);          /* dd.rs:185-195 */
hpo_dd hpo_dd_add_f64(hpo_dd a, double b);      /* dd.rs:228-235 */
hpo_dd hpo_dd_mul_f64(hpo_dd a, double b);      /* dd.rs:243-250 */
hpo_dd hpo_dd_div_f64(hpo_dd a, double b);      /* dd.rs:252-260 */
hpo_dd hpo_dd_sqrt(hpo_dd a);                   /* dd.rs:61-78 */
hpo_dd hpo_dd_recip(hpo_dd a);                  /* dd.rs:80-83 */
hpo_dd hpo_dd_exp(hpo_dd a);                    /* dd.rs:106-121 */
hpo_dd hpo_dd_ln(hpo_dd a);                     /* dd.rs:124-137 */
/* array forms for the Python harness: out[i] = op(a[i], b[i]) */
void hpo_dd_binop_n(int op, const double* a, const double* b, double* out, size_t n);
void hpo_dd_unop_n(int op, const double* a, double* out, size_t n);

/* ---- MultiFloat<4> ops (crates/quad/src/multifloat) ---- */
void hpo_mf4_renormalize(double* limbs4);                        /* core.rs:57-93 */
void hpo_mf4_renormalize_from8(const double* in8, double* out4); /* core.rs:99-168 */
hpo_mf4 hpo_mf4_add(hpo_mf4 a, hpo_mf4 b);   /* add.rs:19-106 */
hpo_mf4 hpo_mf4_sub(hpo_mf4 a, hpo_mf4 b);   /* add.rs:108-114 */
hpo_mf4 hpo_mf4_mul(hpo_mf4 a, hpo_mf4 b);   /* mul.rs:12-111 */
hpo_mf4 hpo_mf4_mul_f64(hpo_mf4 a, double b);/* mul.rs:124-202 */
hpo_mf4 hpo_mf4_recip(hpo_mf4 a);            /* div.rs:14-41 */
hpo_mf4 hpo_mf4_div(hpo_mf4 a, hpo_mf4 b);   /* div.rs:44-77 */
hpo_mf4 hpo_mf4_sqrt(hpo_mf4 a);             /* math.rs:22-40 */
hpo_mf4 hpo_mf4_exp(hpo_mf4 a);              /* transcendental.rs:179-243 */
void hpo_mf4_binop_n(int op, const double* a, const double* b, double* out, size_t n);
void hpo_mf4_unop_n(int op, const double* a, double* out, size_t n);

/* ---- dd linear algebra (crates/quad/src/linalg.rs) ---- */
void hpo_gemm(const double* a, const double* b, double* c, size_t m, size_t n, size_t k);      /* 150-165 */
void hpo_gemm_rows(const double* a, const double* b, double* c, size_t m, size_t n, size_t k,
                   size_t row0, size_t nrows);   /* rows [row0,row0+nrows) of hpo_gemm */
void hpo_gemm_atb(const double* a, const double* b, double* c, size_t m, size_t n, size_t k);  /* 168-182 */
void hpo_gemm_atb_rows(const double* a, const double* b, double* c, size_t m, size_t n, size_t k,
                       size_t row0, size_t nrows);
/* return 0 on success, else 1 with *pivot / *value set (CholeskyError, linalg.rs:196-208) */
int hpo_cholesky(double* a, size_t n, size_t* pivot, double* value);                          /* 42-96 */
int hpo_cholesky_blocked(double* a, size_t n, size_t nb, size_t* pivot, double* value);       /* 238-293 */
void hpo_forward_solve(const double* l, const double* b, double* y, size_t n);                 /* 99-108 */
void hpo_backward_solve(const double* l, const double* y, double* x, size_t n);                /* 111-120 */
void hpo_solve_cholesky(const double* l, const double* b, double* x, size_t n);                /* 123-126 */
void hpo_matvec(const double* a, const double* x, double* y, size_t n);                        /* 129-147 */
double hpo_cond_estimate(const double* l, size_t n);                                           /* 185-194 */

/* ---- MF4 linear algebra (crates/quad/src/multifloat/linalg.rs) ---- */
void hpo_mf4_gemm(const double* a, const double* b, double* c, size_t m, size_t n, size_t k);      /* 39-61 */
void hpo_mf4_gemm_rows(const double* a, const double* b, double* c, size_t m, size_t n, size_t k,
                       size_t row0, size_t nrows);
void hpo_mf4_gemm_atb(const double* a, const double* b, double* c, size_t m, size_t n, size_t k);  /* 64-85 */
int hpo_mf4_cholesky(double* a, size_t n);                                                         /* 89-141 */
void hpo_mf4_cholesky_solve(const double* l, const double* b, double* x, size_t n);                /* 145-172 */
void hpo_mf4_gemv(const double* a, const double* x, double* y, size_t n, size_t m);                /* 17-36 */

/* ---- solver crate (crates/solver/src/lib.rs) ---- */
typedef struct {
  double kappa;
  double residual_norm;
  int strategy;            /* 0 Float64, 1 MixedPrecision, 2 Float128 (lib.rs:25-30) */
  int refinement_iters;
  int error;               /* 0 ok, 1 NotPositiveDefinite */
  size_t pivot;
} hpo_solution;
/* x_dd: n dd values (AoS); x_f64: n doubles */
void hpo_solve_spd(const double* a, const double* b, size_t n, hpo_solution* sol, double* x_dd, double* x_f64); /* 38-68 */
void hpo_solve_spd_f128(const double* a_dd, const double* b_dd, size_t n, hpo_solution* sol, double* x_dd, double* x_f64); /* 71-73 */
double hpo_norm_f128(const double* x_dd, size_t n);   /* 243-249 */

/* ---- overlap crate (crates/overlap/src) ---- */
double hpo_overlap_primitive_f64(double ai, const double* ci, double aj, const double* cj);   /* integral.rs:15-25 */
hpo_dd hpo_overlap_primitive_f128(double ai, const double* ci, double aj, const double* cj);  /* integral.rs:29-56 */
/* screen_shell_pair for single-primitive S shells: 0 F64, 1 F128, 2 Skip (screen.rs:57-101) */
int hpo_screen_s_pair(double ai, const double* ci, double aj, const double* cj,
                      double diffuse_threshold, double ratio_threshold, double eps_skip, double safe_distance);
/* assemble_overlap_mixed over single-primitive S shells (assemble.rs:35-85); counts[3] = f64, f128, skipped */
void hpo_assemble_overlap_s(const double* centers, const double* alpha, const double* coeff, size_t n,
                            const double* params4, double* s_f64, size_t* counts);
/* normalised s-type overlap in dd, built from the reference's dd ops:
 *   S = (4 a b / (a+b)^2)^{3/4} exp(-(a b/(a+b)) |Ri-Rj|^2), Ri-Rj by TwoSum.  Full n x n, AoS dd. */
hpo_dd hpo_overlap_s_dd_entry(double ai, const double* ci, double aj, const double* cj);
void hpo_overlap_s_dd(const double* centers, const double* alpha, size_t n, double* s_dd);
/* gradients of log det S and b^T S^-1 b over the 4n parameters (3n centre coordinates,
 * Gaussian-major, then n exponents), from the reference's dd ops: S^-1 by solve_cholesky
 * against unit vectors, x = S^-1 b.  logdet, quad: 1 dd; g_ld, g_q: 4n dd; x_out: n dd
 * (may be NULL).  Returns 1 if S is not positive definite. */
int hpo_overlap_adjoint_dd(const double* centers, const double* alpha, const double* b_dd, size_t n, double* logdet,
                           double* quad, double* g_ld, double* g_q, double* x_out);
/* the same with a planted mistake, for tests of the checks themselves (1: a_u in place of
 * a_v in the |R|^2 a_v^2/g^2 term; 2: the 3/4 (1/a_u - 2/g) term dropped; 3: the sign of
 * dS/dR flipped) */
int hpo_overlap_adjoint_dd_planted(const double* centers, const double* alpha, const double* b_dd, size_t n,
                                   double* logdet, double* quad, double* g_ld, double* g_q, double* x_out,
                                   int mistake);
/* f256 central finite differences of log det S and b^T S^-1 b at the given coordinates
 * (indices into the 4n parameters), steps h and 2h, each evaluated exactly through a
 * rank-2 update of the f256 factor (determinant lemma, Woodbury).  fd_ld, fd_q:
 * ncoords x 2 (step) MF4 values.  Returns 1 if S is not positive definite. */
int hpo_overlap_fd_f256(const double* centers, const double* alpha, const double* b_dd, size_t n, const long* coords,
                        size_t ncoords, double h, double* fd_ld, double* fd_q);

/* assemble_overlap_mixed (crates/overlap/src/assemble.rs:35-85) over contracted s/p/d/f
 * shells: shell s has centre centers[3s..3s+2], l[s] in 0..3 and primitives
 * prim_off[s] .. prim_off[s+1]-1 of alpha / coeff; params4 = ScreenParams {diffuse,
 * ratio, eps_skip, safe_distance}.  s: n_ao x n_ao f64; counts: F64, F128, Skip shell
 * pairs.  Returns n_ao = sum (l+1)(l+2)/2. */
size_t hpo_assemble_overlap_mixed(const double* centers, const int* l, const long* prim_off, const double* alpha,
                                  const double* coeff, size_t n_shells, const double* params4, double* s,
                                  size_t* counts);
/* Per-entry envelope of the same assembly: sum over primitive pairs of the magnitudes of
 * the terms whose signed sum is S_uv (0 for skipped pairs).  Test infrastructure: the
 * scale of the per-entry bound on the device's differences. */
void hpo_assemble_overlap_envelope(const double* centers, const int* l, const long* prim_off, const double* alpha,
                                   const double* coeff, size_t n_shells, const double* params4, double* env);
/* Moves every exp of the overlap integrals by a deterministic pseudo-random number of ulps
 * in [-ulps, ulps]; 0 restores libm's exp.  Test infrastructure only. */
void hpo_set_exp_ulp_shift(int ulps);

/* ---- MultiFloat<N>, N = 1..8 (hpo_mfn.cpp): the generic expansion arithmetic and linear
 * algebra of multifloat/{core,add,mul,div,math,linalg}.rs.  AoS arrays of N doubles per
 * value.  Return -1 for N outside 1..8. */
/* op: 0 add, 1 sub, 2 mul, 3 div, 4 recip(a), 5 sqrt(a), 6 renormalize(a) */
int hpo_mfn_elementwise(int nl, int op, const double* a, const double* b, double* out, size_t n);
/* mf_gemm / mf_gemm_atb (linalg.rs:39-85) for rows row0 .. row0+nrows-1 of C */
int hpo_mfn_gemm_rows(int nl, int trans_a, const double* a, const double* b, double* c, size_t m, size_t n, size_t k,
                      size_t row0, size_t nrows);
int hpo_mfn_gemv(int nl, const double* a, const double* x, double* y, size_t n, size_t m); /* linalg.rs:17-36 */
long hpo_mfn_cholesky(int nl, double* a, size_t n); /* linalg.rs:89-141: 1 + failing column, 0 on success */
int hpo_mfn_cholesky_solve(int nl, const double* l, const double* b, double* x, size_t n); /* 145-172 */
/* FP64 operations of one MAC s = s + a * b (add.rs + mul.rs) on the given values */
long hpo_count_mfn_mac(int nl, const double* s, const double* a, const double* b);

/* ---- reductions ---- */
/* dot (crates/quad/src/linalg.rs:7-27): serial below 1024 elements, else 1024-element
 * chunk sums combined by rayon's halving tree (split at len/2 down to single chunks,
 * each leaf ZERO + chunk sum) */
hpo_dd hpo_dot(const double* a, const double* b, size_t n);
/* gemv (linalg.rs:29-40): y_i = dot_serial(row_i, x) */
void hpo_gemv(const double* a, size_t rows, size_t cols, const double* x, double* y);
/* Neumaier compensated sums (crates/compensated-sum/src/lib.rs:6-66) */
double hpo_neumaier_sum_f64(const double* v, size_t n);
hpo_dd hpo_neumaier_sum(const double* v, size_t n);
double hpo_par_neumaier_sum_f64(const double* v, size_t n);
hpo_dd hpo_par_neumaier_sum(const double* v, size_t n);

/* ---- eigensolvers ---- */
/* jacobi_eigen (crates/quad/src/linalg.rs:299-362): cyclic Jacobi, max_iter sweeps.
 * evals: n dd in diagonal order; evecs: n x n dd, row i = eigenvector i.  Returns 1 if
 * the off-diagonal norm did not fall below 1e-30 within max_iter sweeps. */
int hpo_jacobi_eigen(const double* a, size_t n, size_t max_iter, double* evals, double* evecs);
/* the same with a planted mistake, for tests of the checks themselves (1: the sine's
 * sign flipped in the eigenvector update; 2: tau from a_pp - a_qq) */
int hpo_jacobi_eigen_planted(const double* a, size_t n, size_t max_iter, double* evals, double* evecs, int mistake);
/* the solver crate's max-pivot Jacobi (crates/solver/src/ortho.rs:61-152): evals in
 * diagonal order, evecs n x n with column j = eigenvector j */
void hpo_jacobi_eigen_maxpivot(const double* a, size_t n, double* evals, double* evecs);
/* canonical_orthogonalize (ortho.rs:158-193): x receives n x n_orth dd (capacity n x n);
 * returns n_orth */
size_t hpo_canonical_orthogonalize(const double* s, size_t n, double threshold, double* x);
/* solve_gen_eigen (ortho.rs:201-248): evals n_basis dd ascending, c n x n_basis dd
 * (capacities n and n x n); returns n_basis */
size_t hpo_solve_gen_eigen(const double* f, const double* s, size_t n, double threshold, double* evals, double* c);

/* ---- text wire formats (crates/quad/src/dd.rs:331-435,441-514; multifloat/fmt.rs;
 * multifloat/traits.rs:7-150).  Formatting returns the length written to out (no NUL;
 * out holds >= prec + 16 bytes per f64); parsing returns 0 on success, 1 on a rejected
 * literal. ---- */
/* Rust's format!("{:.prec$e}") / ("{:+.prec$e}") of one f64 */
int hpo_format_exp(double x, int prec, int plus, char* out);
int hpo_dd_to_string(const double* v2, char* out);      /* Display, dd.rs:337-341 */
int hpo_mf4_to_string(const double* v4, char* out);     /* Display, multifloat/fmt.rs:7-17 */
int hpo_dd_serialize_str(const double* v2, char* out);  /* dd.rs:446-450 */
int hpo_mf4_serialize_str(const double* v4, char* out); /* traits.rs:14-24 */
/* Rust's str::parse::<f64> */
int hpo_parse_f64(const char* s, size_t len, double* out);
int hpo_dd_from_str(const char* s, size_t len, double* out2);        /* dd.rs:371-435 */
int hpo_mf4_from_str(const char* s, size_t len, double* out4);       /* multifloat/fmt.rs:32-114 */
int hpo_dd_deserialize_str(const char* s, size_t len, double* out2);  /* dd.rs:496-511 */
int hpo_mf4_deserialize_str(const char* s, size_t len, double* out4); /* traits.rs:104-148 */
/* batches over n dd values for the CPU-baseline timing: text + n + 1 offsets (Arrow
 * layout), single thread; deserialize returns the number of rejected elements */
size_t hpo_dd_serialize_n(const double* v, size_t n, char* text, size_t* offsets);
size_t hpo_dd_deserialize_n(const char* text, const size_t* offsets, size_t n, double* out);

/* ---- counting: FP64 ops per unit of work, from an instrumented restatement ---- */
/* returns the number of binary64 operations (+,-,*,/,fma,sqrt) one dd MAC s += a*b costs (29) */
long hpo_count_dd_mac(void);
/* number of binary64 ops of one MF4 MAC s = s + a*b on given operands (data dependent) */
long hpo_count_mf4_mac(const double* s4, const double* a4, const double* b4);
long hpo_count_overlap_entry(double ai, const double* ci, double aj, const double* cj);

int hpo_num_threads(void);

#ifdef __cplusplus
}
#endif
#endif
