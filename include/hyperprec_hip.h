/*
 * hyperprec_hip.h -- C ABI of the MI355X (gfx950) implementation of hyperprec's
 * extended-precision path: double-double (f128) and MultiFloat<4> (f256) GEMM, the
 * blocked Cholesky factorisation with its triangular solves, the solver crate's SPD
 * solves and the Gaussian overlap-matrix build.
 *
 * Plain pointers and sizes only.  Two families of entry points:
 *
 *   hp_*      drop-in replacements for the reference's Rust functions.  Arrays are host
 *             memory in the reference's layout: a dd value is {hi, lo} (two doubles,
 *             crates/quad/src/dd.rs:8-11), an MF4 value is 4 doubles most significant
 *             first (crates/quad/src/multifloat/core.rs:6-8), matrices are row-major.
 *             Each call stages the operands through HBM on the calling thread's default
 *             stream and returns when the results are back in host memory.
 *
 *   hp_dev_*  the same operations on operands already resident in HBM, as
 *             structure-of-arrays planes (dd: a hi plane and a lo plane; MF4: four limb
 *             planes, plane l of an r x c matrix at offset l*r*c), enqueued on the
 *             caller's HIP stream (`stream` is a hipStream_t; NULL is the null stream).
 *
 * Every function returns a status code (HP_OK on success).  On failure
 * hp_last_error() returns a message for the calling thread.  Shape errors, which the
 * reference turns into assert_eq! panics (crates/quad/src/linalg.rs:151-153), are
 * HP_ESHAPE; a Cholesky breakdown, which the reference reports as
 * CholeskyError{pivot, value} (linalg.rs:196-208), is HP_ENOTSPD with the pivot and the
 * f64 value of the failing diagonal written through the out-parameters.
 *
 * All functions are thread-safe.  State kept between calls, all process-wide and guarded:
 *   - the split GEMM's device workspaces (two slots, grown on demand, each ordered by a
 *     completion event) and its CRT tables;
 *   - the autotuner's choices (int8 kernel per shape class and moduli count, the
 *     HP_GEMM_AUTO plan per shape class and accuracy), under a mutex;
 *   - the split-dd factorisation's panel width (hp_set_split_factor_panel, atomic);
 *   - the per-thread error message and the phase timer of hp_set_phase_timing.
 * Everything else is allocated per call and freed before the call returns.
 */
#ifndef HYPERPREC_HIP_H
#define HYPERPREC_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum hp_status {
  HP_OK = 0,
  HP_ESHAPE = 1,   /* inconsistent sizes (reference: assert_eq! panic) */
  HP_ENOTSPD = 2,  /* Cholesky breakdown (reference: CholeskyError / SolveError) */
  HP_EDEVICE = 3,  /* HIP runtime error or no usable gfx950 device */
  HP_EINVAL = 4,   /* invalid argument (null pointer, unknown mode) */
  HP_ENOMEM = 5,   /* device allocation failed */
  HP_ENOCONV = 6   /* iteration limit reached (reference: jacobi_eigen's Err, linalg.rs:360) */
};

/* Message describing the last failure on the calling thread ("" if none). */
const char* hp_last_error(void);
/* Library ABI version (major*100 + minor). */
int hp_version(void);

/* ------------------------------------------------------------------ GEMM */

/* C = A*B, A m x k, B k x n, C m x n, dd, row-major.
 * Replaces gemm (crates/quad/src/linalg.rs:150-165).  Bit-identical to the reference. */
int hp_gemm(const double* a, const double* b, double* c, size_t m, size_t n, size_t k);

/* C = A^T * B with A stored k x m.  Replaces gemm_atb (linalg.rs:168-182). */
int hp_gemm_atb(const double* a, const double* b, double* c, size_t m, size_t n, size_t k);

/* MultiFloat<4> C = A*B.  Replaces mf_gemm::<4> (crates/quad/src/multifloat/linalg.rs:39-61). */
int hp_mf_gemm(const double* a, const double* b, double* c, size_t m, size_t n, size_t k);

/* MultiFloat<4> C = A^T * B, A stored k x m.  Replaces mf_gemm_atb::<4> (multifloat/linalg.rs:64-85). */
int hp_mf_gemm_atb(const double* a, const double* b, double* c, size_t m, size_t n, size_t k);

/* Device dd GEMM on SoA planes.  trans_a = 0: A is m x k (gemm); 1: A is k x m (gemm_atb).
 *   HP_GEMM_PLAIN     the reference's per-element order on FP64 VALU (bit-identical);
 *   HP_GEMM_SPLIT_I8  split precision on INT8 matrix cores: rows of A / columns of B are
 *                     scaled to W-bit integers (W = 105 bits for `slices` = 30 at k = 4096), the
 *                     integer product is formed exactly from `slices` int8 GEMMs modulo
 *                     coprime moduli <= 256 and rebuilt by the Chinese remainder theorem.
 *                     Error <= 2^(1-W) (max_p|A_ip| sum_p|B_pj| + sum_p|A_ip| max_p|B_pj|)
 *                     + 2^-106 |C_ij|.  slices in [8, 32]; 0 selects the fewest moduli
 *                     reaching dd accuracy (W >= 105) at this k (30 at k = 4096).
 *                     The bound is normwise per row of A and column of B, NOT
 *                     componentwise: an entry far below that norm (graded or cancelling
 *                     data) keeps only ~W bits of the norm, where the plain kernel keeps
 *                     ~2^-102 k of (|A||B|)_ij.  It is not a componentwise drop-in for
 *                     gemm on such data.  A row of A or a column of B holding an inf or a
 *                     NaN yields NaN in C's row / column, as the reference's dd arithmetic
 *                     does; results beyond the binary64 range may be inf where the
 *                     reference gives NaN.
 *   HP_GEMM_AUTO      `slices` is the accuracy asked for, in bits (0 selects 104): the
 *                     result's error is at most 2^-bits of the split bound's norm.  Picks
 *                     the fewest moduli whose W >= bits + 1 (hp_split_slices_for_bits) and,
 *                     when the plain kernel's bound (k+1) 2^-102 (|A||B|)_ij meets bits too,
 *                     the faster of the two, timed once per shape class on the first call's
 *                     operands (hp_gemm_plan_get reports the choice).  HP_EINVAL above 112.
 * The split modes keep a device workspace between calls, grown as needed
 * (~slices*(m*k + k*n + m*n) bytes).  Calls on one stream are ordered by the stream; a call
 * on another stream first waits (on the device) for the previous call's completion event.
 * The call returns once its kernels are enqueued, except when the workspace must grow
 * while a previous call is still running (it then waits for that call). */
enum hp_gemm_mode {
  HP_GEMM_PLAIN = 0,
  HP_GEMM_SPLIT_I8 = 2,
  HP_GEMM_AUTO = 3
};
int hp_dev_gemm(int trans_a, const double* a_hi, const double* a_lo, const double* b_hi, const double* b_lo,
                double* c_hi, double* c_lo, size_t m, size_t n, size_t k, int mode, int slices, void* stream);
/* The fewest split moduli whose integer width W reaches bits + 1 at inner dimension k
 * (0 if none does).  No counterpart in the reference: the split form is this library's. */
int hp_split_slices_for_bits(size_t k, int bits);
/* The choice HP_GEMM_AUTO made for a shape class and accuracy: returns 1 and fills `plan`,
 * or 0 when that class has not been tuned yet.  ms_* are the timings it chose on (0 when
 * only one form met the accuracy). */
typedef struct {
  int mode;        /* HP_GEMM_PLAIN or HP_GEMM_SPLIT_I8 */
  int slices;      /* moduli of the split form (0: none meets the accuracy) */
  int w_bits;      /* its integer width W */
  int variant;     /* its int8 kernel (hp_split_gemm_variant) */
  float ms_plain;  /* event-timed run of each candidate */
  float ms_split;
} hp_gemm_plan;
int hp_gemm_plan_get(int trans_a, size_t m, size_t n, size_t k, int bits, hp_gemm_plan* plan);

/* Device MF4 GEMM on limb-major planes (see above). */
int hp_dev_mf_gemm(int trans_a, const double* a, const double* b, double* c, size_t m, size_t n, size_t k,
                   void* stream);

/* ------------------------------------------------------------------ MultiFloat<4> linear algebra */

/* In-place MF4 Cholesky A = L L^T (n x n, row-major, AoS), replaces mf_cholesky::<4>
 * (crates/quad/src/multifloat/linalg.rs:89-141): lower triangle written, upper keeps A,
 * bit-identical.  On a pivot with limb 0 <= 0 returns HP_ENOTSPD (the reference's Err) with
 * *pivot set and the columns before it factored, as the reference leaves them. */
int hp_mf_cholesky(double* a, size_t n, size_t* pivot);
/* L L^T x = b by forward then backward substitution, replaces mf_cholesky_solve::<4>
 * (multifloat/linalg.rs:145-172); bit-identical. */
int hp_mf_cholesky_solve(const double* l, const double* b, double* x, size_t n);
/* y = A x, A n x m row-major, replaces mf_gemv::<4> (multifloat/linalg.rs:17-36). */
int hp_mf_gemv(const double* a, const double* x, double* y, size_t n, size_t m);
/* Elementwise MF4 ops over n values: op 0 recip(a) (multifloat/div.rs:14-41), 1 sqrt(a)
 * (multifloat/math.rs:22-40), 2 a / b (div.rs:44-77); b is read only for op 2. */
int hp_mf4_elementwise(int op, const double* a, const double* b, double* out, size_t n);
/* MultiFloat<4>::from_f128 / to_f128 (multifloat/conv.rs:120-131) over n values: limbs
 * (hi, lo, 0, 0), and (limb 0, limb 1) with limbs 2 and 3 dropped unrounded. */
int hp_mf4_from_dd(const double* dd, double* mf4, size_t n);
int hp_dd_from_mf4(const double* mf4, double* dd, size_t n);

/* ---- MultiFloat<N>, N = 1..8 (f512 = MultiFloat<8>; N = 4 is f256 and takes the kernels
 * above).  The reference's generic expansion arithmetic (multifloat/{core,add,mul,div,
 * math}.rs) and linear algebra (multifloat/linalg.rs), bit-identical per element.  Values
 * are AoS arrays of nl doubles (limb 0 most significant), as [MultiFloat<N>] is in memory;
 * nl outside 1..8 is HP_EINVAL. */
/* C = A*B, A m x k, B k x n.  Replaces mf_gemm::<N> (multifloat/linalg.rs:39-61). */
int hp_mfn_gemm(int nl, const double* a, const double* b, double* c, size_t m, size_t n, size_t k);
/* C = A^T B, A stored k x m.  Replaces mf_gemm_atb::<N> (multifloat/linalg.rs:64-85). */
int hp_mfn_gemm_atb(int nl, const double* a, const double* b, double* c, size_t m, size_t n, size_t k);
/* The same on device limb planes (plane l of an r x c matrix at offset l*r*c), enqueued
 * on `stream` (a hipStream_t; NULL is the default stream). */
int hp_dev_mfn_gemm(int nl, int trans_a, const double* a, const double* b, double* c, size_t m, size_t n, size_t k,
                    void* stream);
/* y = A x, A n x m.  Replaces mf_gemv::<N> (multifloat/linalg.rs:17-36). */
int hp_mfn_gemv(int nl, const double* a, const double* x, double* y, size_t n, size_t m);
/* *out = sum a_i b_i from ZERO in index order.  Replaces mf_dot::<N> (multifloat/linalg.rs:7-14). */
int hp_mfn_dot(int nl, const double* a, const double* b, size_t n, double* out);
/* In-place Cholesky, replaces mf_cholesky::<N> (multifloat/linalg.rs:89-141): HP_ENOTSPD
 * with *pivot at a column whose diag.limbs[0] <= 0 (the columns before it factored, as the
 * reference leaves them; the upper triangle keeps A). */
int hp_mfn_cholesky(int nl, double* a, size_t n, size_t* pivot);
/* L L^T x = b, replaces mf_cholesky_solve::<N> (multifloat/linalg.rs:145-172). */
int hp_mfn_cholesky_solve(int nl, const double* l, const double* b, double* x, size_t n);
/* out_i = op(a_i, b_i): 0 a + b, 1 a - b, 2 a * b, 3 a / b (add.rs:19-114, mul.rs:12-111,
 * div.rs:44-77), 4 a.recip() (div.rs:14-41), 5 a.sqrt() (math.rs:22-40), 6 a.renorm()
 * (core.rs:231, renormalize); b is unused (may be NULL) for ops 4-6. */
int hp_mfn_elementwise(int nl, int op, const double* a, const double* b, double* out, size_t n);
/* MultiFloat<nl_to> from MultiFloat<nl_from>: the first min(nl_from, nl_to) limbs, the rest
 * zero: from_f128 / to_f128 / from_f256 / to_f256 (multifloat/conv.rs:119-157; f128 is the
 * 2-limb case). */
int hp_mfn_convert(int nl_from, int nl_to, const double* in, double* out, size_t n);

/* ------------------------------------------------------------------ Cholesky and solves */

/* In-place unblocked Cholesky A = L L^T of an n x n dd matrix (row-major, AoS).
 * Replaces cholesky (crates/quad/src/linalg.rs:42-96): only the lower triangle is
 * written, the upper keeps A.  Bit-identical to the reference.  On a non-positive pivot
 * returns HP_ENOTSPD with *pivot / *value = CholeskyError{pivot, value}. */
int hp_cholesky(double* a, size_t n, size_t* pivot, double* value);

/* Blocked Cholesky with block size nb (0 -> 64).  Replaces cholesky_blocked
 * (linalg.rs:238-293); bit-identical to the reference for the same nb. */
int hp_cholesky_blocked(double* a, size_t n, size_t nb, size_t* pivot, double* value);

/* Device form on dd planes: unblocked != 0 selects `cholesky`, else `cholesky_blocked(nb)`.
 * Synchronises `stream` to read the pivot status. */
int hp_dev_cholesky(double* a_hi, double* a_lo, size_t n, size_t nb, int unblocked, size_t* pivot, double* value,
                    void* stream);

/* L y = b, L^T x = y and both (linalg.rs:99-126); L is the dd factor (lower triangle),
 * b / y / x are n dd values.  Reference operation order (bit-identical). */
int hp_forward_solve(const double* l, const double* b, double* y, size_t n);
int hp_backward_solve(const double* l, const double* y, double* x, size_t n);
int hp_solve_cholesky(const double* l, const double* b, double* x, size_t n);

/* y = A x in dd (linalg.rs:129-147), reference order. */
int hp_matvec(const double* a, const double* x, double* y, size_t n);

/* (max|L_ii| / min|L_ii|)^2 of a dd factor (linalg.rs:185-194). */
double hp_cond_estimate(const double* l, size_t n);

/* ------------------------------------------------------------------ SPD solver */

/* Strategy (crates/solver/src/lib.rs:25-30) */
enum hp_strategy { HP_STRATEGY_FLOAT64 = 0, HP_STRATEGY_MIXED = 1, HP_STRATEGY_FLOAT128 = 2 };

/* Solution{kappa, residual_norm, strategy, refinement_iters} (lib.rs:10-23); x and x_f64
 * are returned through the x_dd (n dd values) and x_f64 (n doubles) arguments.  On
 * SolveError::NotPositiveDefinite{pivot} (lib.rs:251-265) the call returns HP_ENOTSPD
 * and `pivot` is set. */
typedef struct {
  double kappa;
  double residual_norm;
  int strategy;
  int refinement_iters;
  size_t pivot;
} hp_solution;

/* solve_spd (lib.rs:38-68): f64 A and b; f64 Cholesky, then f64 / mixed / f128 strategy. */
int hp_solve_spd(const double* a, const double* b, size_t n, hp_solution* sol, double* x_dd, double* x_f64);

/* solve_spd_f128 (lib.rs:71-73): dd A and b, unblocked dd Cholesky and solves, no refinement. */
int hp_solve_spd_f128(const double* a_dd, const double* b_dd, size_t n, hp_solution* sol, double* x_dd,
                      double* x_f64);

/* Mixed-precision SPD solve to a target accuracy (extends iterative_refinement,
 * lib.rs:75-125, with a quad-double residual).  S and b are dd; x is returned in dd.
 * The factor is chosen by `factor_kind`:
 *   HP_FACTOR_F64       binary64 Cholesky of S.hi (needs kappa(S) * 2^-53 < 1)
 *   HP_FACTOR_DD        dd cholesky_blocked(nb = 64), reference order on FP64 VALU
 *   HP_FACTOR_DD_SPLIT  dd blocked Cholesky (nb = 256) with the trailing updates on the
 *                       INT8 matrix cores (split GEMM)
 * Iterates r = b - S x (qd), x += solve(L, r) and stops on the first of:
 *   - the measured correction max|dx| / max|x| <= tol                 (converged = 1);
 *   - after two or more corrections, the next one predicted from the last two,
 *     last * (last / prev), <= tol / 10 while the corrections shrink at least 4x per
 *     step                                                              (converged = 2);
 *   - after four or more corrections, a correction above half the smallest so far
 *     (the reference's stagnation test, lib.rs:102-106)                (converged = 0);
 *   - max_iter steps                                                   (converged = 0).
 * The call returns HP_OK in every case, like the reference's iterative_refinement; the
 * caller reads `converged`. */
enum hp_factor_kind { HP_FACTOR_F64 = 0, HP_FACTOR_DD = 1, HP_FACTOR_DD_SPLIT = 2 };
enum hp_converged { HP_NOT_CONVERGED = 0, HP_CONVERGED_MEASURED = 1, HP_CONVERGED_PREDICTED = 2 };
typedef struct {
  int factor_kind;
  int iterations;
  double factor_ms;        /* wall time of the factorisation (device-synchronised) */
  double refine_ms;        /* wall time of the refinement loop */
  double last_correction;  /* max|dx| / max|x| of the last step */
  size_t pivot;            /* failing pivot on HP_ENOTSPD */
  int converged;           /* hp_converged: which stop rule ended the loop */
} hp_refine_stats;
int hp_dev_solve_spd_refined(int factor_kind, const double* s_hi, const double* s_lo, const double* b_hi,
                             const double* b_lo, size_t n, double tol, int max_iter, double* x_hi, double* x_lo,
                             hp_refine_stats* stats, void* stream);
int hp_solve_spd_refined(int factor_kind, const double* s_dd, const double* b_dd, size_t n, double tol, int max_iter,
                         double* x_dd, hp_refine_stats* stats);
/* Panel width of HP_FACTOR_DD_SPLIT (64, 128, 256 or 512; process-wide, default 256). */
int hp_set_split_factor_panel(int nb);

/* ------------------------------------------------------------------ batched factor + solve */

/* Many dd SPD systems of uneven size at once: each instance is copied to L and factored
 * with cholesky_blocked(nb = 64) (bit-identical to hp_cholesky_blocked; the upper triangle
 * keeps S), then L L^T X = B is solved for `nrhs` (1..8) right-hand sides by blocked
 * forward and backward substitution (64-row diagonal blocks, pivots applied as dd
 * reciprocals; not the reference's solve order, accurate to its error bound).  Two
 * schedules, same bits:
 *   HP_BATCH_PER_INSTANCE  one persistent launch, one workgroup per instance, taken
 *                          largest first from an atomic queue;
 *   HP_BATCH_PANEL_STEPS   panel step p of every instance at once: its diagonal blocks
 *                          (one workgroup each), the rows below (8 per workgroup), the
 *                          trailing 64 x 64 tiles (one per workgroup) across the whole chip,
 *                          then one solve workgroup per instance.
 * The autotuner times both on the first batch of a class (instance count, largest n and
 * total n^3, each to the next power of two; nrhs) and keeps the faster for the class;
 * hp_batch_set_schedule pins one.  An instance that breaks down leaves its L unspecified.
 * Layout (device): matrices packed in instance order, instance i at sum_{j<i} n_j^2 in each
 * of the hi / lo planes (s_*, l_*); right-hand sides packed likewise, instance i's r-th
 * vector at sum_{j<i} nrhs*n_j + r*n_i (b_*, x_*).  status[i] = 0, or p+1 when instance i
 * broke down at pivot p. */
int hp_dev_batch_cholesky_solve(const int* sizes, size_t count, int nrhs, const double* s_hi, const double* s_lo,
                                double* l_hi, double* l_lo, const double* b_hi, const double* b_lo, double* x_hi,
                                double* x_lo, int* status, void* stream);

/* The same batch captured once in a hipGraph (queue reset by a kernel node + the
 * persistent kernel) and replayed by hp_batch_graph_launch on any stream. */
typedef struct hp_batch hp_batch;
enum hp_batch_schedule { HP_BATCH_PER_INSTANCE = 1, HP_BATCH_PANEL_STEPS = 2 };
/* 0: the autotuner chooses (default); HP_BATCH_PER_INSTANCE / HP_BATCH_PANEL_STEPS: pinned
 * for every later batch of the process. */
int hp_batch_set_schedule(int schedule);
hp_batch* hp_batch_graph_create(const int* sizes, size_t count, int nrhs, const double* s_hi, const double* s_lo,
                                double* l_hi, double* l_lo, const double* b_hi, const double* b_lo, double* x_hi,
                                double* x_lo, int* status, void* stream);
int hp_batch_graph_launch(hp_batch* batch, void* stream);
/* The schedule the graph captured (HP_BATCH_PER_INSTANCE or HP_BATCH_PANEL_STEPS). */
int hp_batch_graph_schedule(const hp_batch* batch);
void hp_batch_graph_destroy(hp_batch* batch);

/* ------------------------------------------------------------------ overlap build */

/* Normalised s-type Gaussian overlap in dd: S_uv = (4 a_u a_v/(a_u+a_v)^2)^{3/4}
 * exp(-(a_u a_v/(a_u+a_v)) |R_u - R_v|^2), the single-primitive s-s case of
 * assemble_overlap_mixed (crates/overlap/src/assemble.rs:35-85, integral.rs:15-142) built
 * in dd without screening.  centers: n x 3 (x, y, z), alpha: n, s_dd: n x n dd (AoS). */
int hp_overlap_s(const double* centers, const double* alpha, size_t n, double* s_dd);
int hp_dev_overlap_s(const double* centers, const double* alpha, size_t n, double* s_hi, double* s_lo, void* stream);

/* assemble_overlap_mixed (crates/overlap/src/assemble.rs:35-85) over contracted s/p/d/f
 * shells: shell i has centre centers[3i..3i+2], angular momentum l[i] (0..3 = S, P, D, F,
 * Cartesian components in cart_components order, integral.rs:220-231) and primitives
 * prim_off[i] .. prim_off[i+1]-1 of alpha / coeff (prim_off: n_shells + 1 entries).
 * params4 = ScreenParams {diffuse_threshold, ratio_threshold, eps_skip, safe_distance}
 * (screen.rs:18-46).  s: n_ao x n_ao f64 (row-major), n_ao = sum (l+1)(l+2)/2 (HP_ESHAPE
 * otherwise); counts: TierCounts {f64_pairs, f128_pairs, skipped_pairs}.  Each shell pair
 * is screened and its block formed in f64 or dd (integral.rs:65-346) exactly as the
 * reference; only libm exp may differ in its last bit. */
int hp_assemble_overlap_mixed(const double* centers, const int* l, const int64_t* prim_off, const double* alpha,
                              const double* coeff, size_t n_shells, const double* params4, double* s, size_t n_ao,
                              size_t* counts);
/* the same with S left in device memory (s_dev: n_ao x n_ao f64), for device-resident
 * consumers such as the factorizations; shell data and counts are host memory */
int hp_dev_assemble_overlap_mixed(const double* centers, const int* l, const int64_t* prim_off, const double* alpha,
                                  const double* coeff, size_t n_shells, const double* params4, double* s_dev,
                                  size_t n_ao, size_t* counts, void* stream);

/* ------------------------------------------------------------------ eigensolvers */

/* Symmetric dd eigendecomposition, replaces jacobi_eigen (crates/quad/src/linalg.rs:299-362):
 * a n x n dd (AoS, symmetric); evals n dd in diagonal order of the original indices;
 * evecs n x n dd, row i = eigenvector i.  Stops when the dd Frobenius norm of the strict
 * upper triangle is below 1e-30, tested before each of at most max_iter sweeps;
 * HP_ENOCONV (the reference's Err) if that never happens.  The sweeps rotate n/2 disjoint
 * pairs at a time (round-robin order), so the results agree with the reference's cyclic
 * order to dd rounding, not bit for bit. */
int hp_jacobi_eigen(const double* a, size_t n, size_t max_iter, double* evals, double* evecs);
/* the same on device planes; stop_rule 0: the Frobenius test above (HP_ENOCONV on
 * failure), 1: max |a_pq| < 1e-60 (the solver crate's test, ortho.rs:84), at most
 * max_sweeps sweeps, no error; vec_rows 1: row i = eigenvector i, 0: column i.  sweeps
 * (optional) receives the number of sweeps run. */
int hp_dev_jacobi_eigen(const double* a_hi, const double* a_lo, size_t n, size_t max_sweeps, int stop_rule,
                        double* evals_hi, double* evals_lo, double* vecs_hi, double* vecs_lo, int vec_rows,
                        size_t* sweeps, void* stream);
/* From n = min_n on (default 768; 0 restores it, SIZE_MAX disables it) the eigensolvers
 * above (and canonical_orthogonalize, solve_gen_eigen) run block Jacobi: 16 x 16 blocks
 * paired in round-robin order, each 32 x 32 subproblem taken through one inner parallel
 * sweep, A <- Q^T A Q and V <- V Q as dd block products.  Same stop rules and results to
 * dd rounding; a sweep makes n/16 passes over A instead of n.  Process-wide. */
int hp_eigen_set_block_min(size_t min_n);
/* canonical_orthogonalize (crates/solver/src/ortho.rs:158-193): diagonalise S (stop rule 1),
 * keep eigenvalues above threshold in ascending order, X = U diag(1/sqrt(lambda)).
 * x_dd: capacity n x n dd, receives n x n_orth (row-major). */
int hp_canonical_orthogonalize(const double* s_dd, size_t n, double threshold, double* x_dd, size_t* n_orth);
/* solve_gen_eigen (ortho.rs:201-248): F C = S C eps.  X from S, F' = X^T F X (symmetrised),
 * F' = C' diag(eps) C'^T, C = X C'.  evals_dd: capacity n, receives n_basis ascending;
 * c_dd: capacity n x n, receives n x n_basis (column k = eigenvector k). */
int hp_solve_gen_eigen(const double* f_dd, const double* s_dd, size_t n, double threshold, double* evals_dd,
                       double* c_dd, size_t* n_basis);
int hp_dev_solve_gen_eigen(const double* f_hi, const double* f_lo, const double* s_hi, const double* s_lo, size_t n,
                           double threshold, double* evals_hi, double* evals_lo, double* c_hi, double* c_lo,
                           size_t* n_basis, size_t* sweeps, void* stream);

/* ------------------------------------------------------------------ reductions */

/* dd dot product, replaces dot (crates/quad/src/linalg.rs:7-18): below 1024 elements the
 * serial sum from ZERO (linalg.rs:20-27); else the serial sums p_i of 1024-element chunks
 * combined as the reference's rayon reduce does when every range is split at len / 2:
 * tree(lo, 1) = ZERO + p_lo, tree(lo, len) = tree(lo, len/2) + tree(lo + len/2, len - len/2).
 * a, b: n dd (AoS); out_dd: 1 dd. */
int hp_dot(const double* a, const double* b, size_t n, double* out_dd);
int hp_dev_dot(const double* a_hi, const double* a_lo, const double* b_hi, const double* b_lo, size_t n,
               double* out_hi, double* out_lo, void* stream);
/* hp_dev_dot without any allocation, for callers that cannot allocate (code under HIP graph
 * capture, tight loops, harnesses that forbid hipMalloc); no counterpart in the reference.
 * Same kernels and bits as hp_dev_dot; its scratch (chunk sums, tree nodes) is a ~1 MiB
 * device array the library holds on each GPU until the process ends, so n <= 2^26
 * (HP_EINVAL above).  Enqueued on `stream` without waiting.  All calls on one GPU share
 * the array: calls that overlap in time on one GPU (two streams, two threads) race. */
int hp_dev_dot_noalloc(const double* a_hi, const double* a_lo, const double* b_hi, const double* b_lo, size_t n,
                       double* out_hi, double* out_lo, void* stream);
/* y = A x with A row-major rows x cols, replaces gemv (linalg.rs:29-40): y_i is the serial
 * dot of row i with x. */
int hp_gemv(const double* a, size_t rows, size_t cols, const double* x, double* y);
int hp_dev_gemv(const double* a_hi, const double* a_lo, size_t rows, size_t cols, const double* x_hi,
                const double* x_lo, double* y_hi, double* y_lo, void* stream);
/* Neumaier compensated sums, replace neumaier_sum_f64 / neumaier_sum / par_neumaier_sum_f64 /
 * par_neumaier_sum (crates/compensated-sum/src/lib.rs:6-66).  The par forms sum 4096-element
 * chunks and then the partials, in order (the serial sum below 4096 elements); the plain
 * forms run the whole sum in order on one GPU thread.  v: n f64 or n dd (AoS); out: one
 * f64 or one dd. */
int hp_neumaier_sum_f64(const double* v, size_t n, double* out);
int hp_neumaier_sum(const double* v, size_t n, double* out_dd);
int hp_par_neumaier_sum_f64(const double* v, size_t n, double* out);
int hp_par_neumaier_sum(const double* v, size_t n, double* out_dd);

/* ------------------------------------------------------------------ text wire formats */

/* Text arrays are Arrow-style: n strings concatenated (no separators, no NUL) in `text`,
 * string i at text[offsets[i] .. offsets[i+1]); offsets has n + 1 entries.
 *
 * Formatting: values are AoS (1, 2 or 4 doubles per element).  offsets is always written;
 * when text is NULL or cap < offsets[n] the call returns HP_ESHAPE without writing text
 * (call again with cap >= offsets[n]).  Maximum lengths per element: f64 with precision P:
 * P + 8 bytes; dd Display 40; MF4 Display 72; dd serde 53; MF4 serde 91.
 *
 * Parsing: out receives 1, 2 or 4 doubles per element (zeros for a rejected element);
 * status (may be NULL) receives per element 0 ok, 1 rejected literal (the reference's
 * Err), 2 the literal has more than 40 significant digits and lies so close to a halfway
 * point that they do not settle the rounding.  Returns HP_EINVAL, naming the first bad
 * element, when any status is nonzero. */

/* Rust's format!("{:.P$e}") (force_sign = 0) or ("{:+.P$e}") of f64 values, 0 <= P <= 64:
 * exact decimal, ties to even (the building block of the formats below). */
int hp_format_f64(const double* v, size_t n, int precision, int force_sign, char* text, size_t cap, size_t* offsets);
/* Display of f128, "{:.32e}" of hi + lo (crates/quad/src/dd.rs:337-341) */
int hp_dd_to_string(const double* v_dd, size_t n, char* text, size_t cap, size_t* offsets);
/* Display of MultiFloat<4>, "{:.64e}" of limb0 + limb1 (crates/quad/src/multifloat/fmt.rs:7-17) */
int hp_mf4_to_string(const double* v_mf4, size_t n, char* text, size_t cap, size_t* offsets);
/* human-readable serde of f128, "{:+.18e}+{:+.18e}" of hi, lo (dd.rs:446-450) */
int hp_dd_serialize_str(const double* v_dd, size_t n, char* text, size_t cap, size_t* offsets);
/* human-readable serde of MultiFloat<4>, "{:+.14e}" limbs joined by '+'
 * (crates/quad/src/multifloat/traits.rs:14-24) */
int hp_mf4_serialize_str(const double* v_mf4, size_t n, char* text, size_t cap, size_t* offsets);
/* Rust's str::parse::<f64> (the building block of the serde parsers) */
int hp_parse_f64(const char* text, const size_t* offsets, size_t n, double* out, uint8_t* status);
/* FromStr of f128 (dd.rs:371-435) and of MultiFloat<4> (multifloat/fmt.rs:32-114) */
int hp_dd_from_str(const char* text, const size_t* offsets, size_t n, double* out_dd, uint8_t* status);
int hp_mf4_from_str(const char* text, const size_t* offsets, size_t n, double* out_mf4, uint8_t* status);
/* serde's visit_str of f128 (dd.rs:475-479, parse_dd_string 496-511) and of MultiFloat<4>
 * (traits.rs:53-57, parse_limbs_string 104-148) */
int hp_dd_deserialize_str(const char* text, const size_t* offsets, size_t n, double* out_dd, uint8_t* status);
int hp_mf4_deserialize_str(const char* text, const size_t* offsets, size_t n, double* out_mf4, uint8_t* status);
/* Device-resident forms.  kind (format): 0 f64 "{:.Pe}", 1 dd Display, 2 MF4 Display, 3 dd
 * serde, 4 MF4 serde; planes[k] = limb plane k (n doubles each); text / offsets are device
 * memory, *need = offsets[n].  kind (parse): 0 f64, 1 dd FromStr, 2 MF4 FromStr, 3 dd serde,
 * 4 MF4 serde; status is device memory (n bytes), *n_bad counts nonzero statuses. */
int hp_dev_format(int kind, const double* const* planes, size_t n, int precision, int force_sign, char* text,
                  size_t cap, size_t* offsets, size_t* need, void* stream);
int hp_dev_parse(int kind, const char* text, const size_t* offsets, size_t n, double* const* planes, uint8_t* status,
                 size_t* n_bad, void* stream);

/* ------------------------------------------------------------------ adjoints */

/* Gradients of F1 = log det S and F2 = b^T S^-1 b (S as hp_overlap_s, b dd) with respect to
 * the 4n parameters: the 3n centre coordinates (Gaussian-major) then the n exponents.
 * No reference counterpart: the reference has no differentiation.  The functions
 * differentiated are its overlap entry (crates/overlap/src/integral.rs:29-56, closed form
 * crates/demo/src/main.rs:104-124) and log det / solve of its cholesky_blocked factor
 * (crates/quad/src/linalg.rs:238-293, 123-126).  Forward/backward follow
 * torch.autograd.Function: the forward saves the factor L and x = S^-1 b in an opaque
 * context; the backward forms a vector-Jacobian product from them, recomputing S^-1 and the
 * overlap entries instead of storing them. */
typedef struct hp_adjoint hp_adjoint;
typedef struct {
  double logdet[2];        /* dd log det S = 2 sum ln L_ii */
  double quad[2];          /* dd b^T S^-1 b */
  int refine_iters;        /* quad-double refinement steps of x = S^-1 b */
  double last_correction;  /* max|dx| / max|x| of the last step */
  size_t pivot;            /* failing pivot when the forward returns HP_ENOTSPD */
  double forward_ms;       /* host wall time of the forward (build, factor, solve) */
  double inverse_ms;       /* last backward, when phase timing is on: L^-1 */
  double gram_ms;          /* S^-1 = L^-T L^-1 */
  double contract_ms;      /* per-row contraction with recomputed overlap entries */
} hp_adjoint_info;
/* Forward on device memory: centers (3n), alpha (n), b planes (n each); factor_kind is
 * HP_FACTOR_DD (reference-order factor) or HP_FACTOR_DD_SPLIT.  *ctx receives the saved
 * state (free with hp_adjoint_destroy); info receives F1, F2.  HP_ENOTSPD if S is not
 * positive definite (info->pivot set). */
int hp_dev_overlap_forward(const double* centers, const double* alpha, const double* b_hi, const double* b_lo,
                           size_t n, int factor_kind, hp_adjoint** ctx, hp_adjoint_info* info, void* stream);
/* Backward: cot = host dd cotangents {c1.hi, c1.lo, c2.hi, c2.lo} of (F1, F2); writes
 * c1 dF1 + c2 dF2 as dd planes: centres gc (3n), exponents ga (n), and b: gb = 2 c2 x
 * (gb may be NULL).  gemm_mode HP_GEMM_PLAIN or HP_GEMM_SPLIT_I8 for the S^-1 products;
 * with c1 = 0 no S^-1 is formed. */
int hp_dev_overlap_backward(hp_adjoint* ctx, const double* cot, int gemm_mode, double* gc_hi, double* gc_lo,
                            double* ga_hi, double* ga_lo, double* gb_hi, double* gb_lo, void* stream);
/* Both full gradients (4n dd planes each; gld may be NULL to skip F1). */
int hp_dev_overlap_gradients(hp_adjoint* ctx, int gemm_mode, double* gld_hi, double* gld_lo, double* gq_hi,
                             double* gq_lo, void* stream);
int hp_adjoint_get_info(const hp_adjoint* ctx, hp_adjoint_info* info);
void hp_adjoint_destroy(hp_adjoint* ctx);
/* Host convenience: forward and both gradients; g_ld_dd, g_q_dd: 4n dd (AoS). */
int hp_overlap_adjoint(const double* centers, const double* alpha, const double* b_dd, size_t n, int factor_kind,
                       int gemm_mode, hp_adjoint_info* info, double* g_ld_dd, double* g_q_dd);

/* ------------------------------------------------------------------ measurement */

/* Per-phase HIP-event timing of the multi-kernel entry points (no reference
 * counterpart; for benchmarks).  on = 1: each split GEMM records events on its stream at
 * its phase boundaries (scaling + residues | int8 GEMMs | CRT), waits for them, and
 * hp_get_phase_times copies the elapsed milliseconds of the last call, returning the phase
 * count.  on = 2: the calls record their boundaries without waiting, and
 * hp_collect_phase_times waits for the last one, sums each phase over the calls since the
 * previous collect (their number in *calls) and returns the phase count.  0: off. */
void hp_set_phase_timing(int on);
int hp_collect_phase_times(double* ms_sum, int max, int* calls);
/* Int8 kernel the split GEMM's autotuner chose for this shape class (0 = not tuned yet).
 * A shape class is (m, n, k) bucketed to powers of two and the moduli count; slices = 0
 * means the count the GEMM takes by default at this k (dd accuracy).  The first split GEMM
 * of a shape class times each candidate twice on its real operands and keeps the fastest. */
int hp_split_gemm_variant(size_t m, size_t n, size_t k, int slices);
/* Pins the int8 kernel for a shape class: 32, 64 or 128 (K stages of that depth staged
 * through registers), 3064 (operand tiles streamed straight into a 4-deep LDS ring of
 * 64-deep stages), 4064 / 5064 (that ring software-pipelined, 8 / 4 wavefronts), 6064
 * (register-staged, three stages ahead) or 7064 (4064 as a persistent kernel, one block per
 * CU); 0 clears it (re-tune). */
int hp_split_gemm_set_variant(size_t m, size_t n, size_t k, int slices, int variant);
int hp_get_phase_times(double* ms, int max);

/* ------------------------------------------------------------------ synthetic inputs */

/* Deterministic workload generators (host memory).  SplitMix64 stream from `seed`;
 * N(0,1) by Box-Muller.  dd: hi ~ N(0,1), lo = hi * U(-2^-53, 2^-53), then renormalised
 * with quick_two_sum.  MF4: limb0 ~ N(0,1), limb i+1 = limb i * U(-2^-53, 2^-53), then
 * renormalised (multifloat/core.rs:57-93).  Outputs are AoS like the hp_* entry points. */
void hp_gen_dd_normal(uint64_t seed, size_t count, double* out_aos);
void hp_gen_mf4_normal(uint64_t seed, size_t count, double* out_aos);
/* Gaussian basis: the first n points (lexicographic, z fastest) of an L^3 lattice,
 * L = ceil(n^{1/3}), with `spacing` bohr plus U(-jitter, jitter) per coordinate, and
 * exponents 10^U(log10 alpha_lo, log10 alpha_hi).  centers: n x 3, alpha: n. */
void hp_gen_overlap_basis(uint64_t seed, size_t n, double alpha_lo, double alpha_hi, double spacing, double jitter,
                          double* centers, double* alpha);

#ifdef __cplusplus
}
#endif
#endif /* HYPERPREC_HIP_H */
