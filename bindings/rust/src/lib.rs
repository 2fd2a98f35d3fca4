//! hyperprec-hip: the reference's extended-precision API on one MI355X (gfx950).
//!
//! Same names, argument meaning and error behaviour as the reference's public functions
//! (crates/quad/src/lib.rs:9-14, crates/quad/src/multifloat/mod.rs:19-22,
//! crates/solver/src/lib.rs:38,71), implemented by libhyperprec_hip through the C ABI of
//! include/hyperprec_hip.h (raw items in `ffi`).  A caller switches with
//! `use hyperprec_hip as quad;` (or `as quadprec_solver` for the solver functions).
//!
//! Layout: the functions pass `&[f128]` / `&[MultiFloat<N>]` as plain `f64` arrays
//! ({hi, lo} and N limbs per value).  The reference's structs are not `#[repr(C)]`
//! (crates/quad/src/dd.rs:6-11, multifloat/core.rs:4-8); the drop-in patch adds
//! `#[repr(C)]` to both (same fields, same size: non-breaking).  The size checks below
//! catch a layout change at compile time.
//!
//! Errors: shape mismatches panic like the reference's `assert_eq!`
//! (crates/quad/src/linalg.rs:151-153); a Cholesky breakdown is the reference's
//! `CholeskyError` / `SolveError`; a HIP failure (no representation in the reference API)
//! panics with the library's message.
//!
//! The `MultiFloat<N>` functions take any N in 1..=8 (f256 = N 4, f512 = N 8), as the
//! reference's mul does (multifloat/mul.rs:21, N <= 8); a larger N panics with a message.

pub mod ffi;

use quad::{f128, CholeskyError, MultiFloat};
use std::os::raw::c_int;

const _: () = assert!(std::mem::size_of::<f128>() == 16 && std::mem::align_of::<f128>() == 8);
const _: () = assert!(std::mem::size_of::<MultiFloat<4>>() == 32 && std::mem::size_of::<MultiFloat<8>>() == 64);

fn last_error() -> String {
    unsafe { std::ffi::CStr::from_ptr(ffi::hp_last_error()) }.to_string_lossy().into_owned()
}

fn check(rc: c_int) {
    match rc {
        ffi::HP_OK => {}
        ffi::HP_ESHAPE => panic!("assertion failed: {}", last_error()),
        _ => panic!("hyperprec_hip: status {rc}: {}", last_error()),
    }
}

fn dd(p: &[f128]) -> *const f64 {
    p.as_ptr() as *const f64
}
fn dd_mut(p: &mut [f128]) -> *mut f64 {
    p.as_mut_ptr() as *mut f64
}
fn limbs<const N: usize>() -> c_int {
    assert!((1..=8).contains(&N), "hyperprec_hip provides MultiFloat<N> for N in 1..=8, not {N}");
    N as c_int
}
fn mf<const N: usize>(p: &[MultiFloat<N>]) -> *const f64 {
    p.as_ptr() as *const f64
}
fn mf_mut<const N: usize>(p: &mut [MultiFloat<N>]) -> *mut f64 {
    p.as_mut_ptr() as *mut f64
}

/// C = A*B (crates/quad/src/linalg.rs:150-165), bit-identical.
pub fn gemm(a: &[f128], b: &[f128], c: &mut [f128], m: usize, n: usize, k: usize) {
    assert_eq!(a.len(), m * k);
    assert_eq!(b.len(), k * n);
    assert_eq!(c.len(), m * n);
    check(unsafe { ffi::hp_gemm(dd(a), dd(b), dd_mut(c), m, n, k) });
}

/// C = A^T*B with A stored k x m (linalg.rs:168-182), bit-identical.
pub fn gemm_atb(a: &[f128], b: &[f128], c: &mut [f128], m: usize, n: usize, k: usize) {
    assert_eq!(a.len(), k * m);
    assert_eq!(b.len(), k * n);
    assert_eq!(c.len(), m * n);
    check(unsafe { ffi::hp_gemm_atb(dd(a), dd(b), dd_mut(c), m, n, k) });
}

/// mf_gemm::<N> (crates/quad/src/multifloat/linalg.rs:39-61), bit-identical.
pub fn mf_gemm<const N: usize>(a: &[MultiFloat<N>], b: &[MultiFloat<N>], c: &mut [MultiFloat<N>], m: usize, n: usize,
                               k: usize) {
    assert_eq!(a.len(), m * k);
    assert_eq!(b.len(), k * n);
    assert_eq!(c.len(), m * n);
    check(unsafe { ffi::hp_mfn_gemm(limbs::<N>(), mf(a), mf(b), mf_mut(c), m, n, k) });
}

/// mf_gemm_atb::<N> (multifloat/linalg.rs:64-85), bit-identical.
pub fn mf_gemm_atb<const N: usize>(a: &[MultiFloat<N>], b: &[MultiFloat<N>], c: &mut [MultiFloat<N>], m: usize,
                                   n: usize, k: usize) {
    assert_eq!(a.len(), k * m);
    assert_eq!(b.len(), k * n);
    assert_eq!(c.len(), m * n);
    check(unsafe { ffi::hp_mfn_gemm_atb(limbs::<N>(), mf(a), mf(b), mf_mut(c), m, n, k) });
}

/// In-place unblocked Cholesky (linalg.rs:42-96): lower triangle = L, upper keeps A.
pub fn cholesky(a: &mut [f128], n: usize) -> Result<(), CholeskyError> {
    assert_eq!(a.len(), n * n);
    let (mut pivot, mut value) = (0usize, 0f64);
    let rc = unsafe { ffi::hp_cholesky(dd_mut(a), n, &mut pivot, &mut value) };
    if rc == ffi::HP_ENOTSPD {
        return Err(CholeskyError { pivot, value });
    }
    check(rc);
    Ok(())
}

/// Blocked Cholesky, block size nb (0 -> 64) (linalg.rs:238-293), bit-identical for the same nb.
pub fn cholesky_blocked(a: &mut [f128], n: usize, nb: usize) -> Result<(), CholeskyError> {
    assert_eq!(a.len(), n * n);
    let (mut pivot, mut value) = (0usize, 0f64);
    let rc = unsafe { ffi::hp_cholesky_blocked(dd_mut(a), n, nb, &mut pivot, &mut value) };
    if rc == ffi::HP_ENOTSPD {
        return Err(CholeskyError { pivot, value });
    }
    check(rc);
    Ok(())
}

fn vec_op(f: unsafe extern "C" fn(*const f64, *const f64, *mut f64, usize) -> c_int, l: &[f128], b: &[f128],
          n: usize) -> Vec<f128> {
    assert_eq!(l.len(), n * n);
    assert_eq!(b.len(), n);
    let mut x = vec![f128::ZERO; n];
    check(unsafe { f(dd(l), dd(b), dd_mut(&mut x), n) });
    x
}

/// L y = b (linalg.rs:99-108).
pub fn forward_solve(l: &[f128], b: &[f128], n: usize) -> Vec<f128> {
    vec_op(ffi::hp_forward_solve, l, b, n)
}
/// L^T x = y (linalg.rs:111-120).
pub fn backward_solve(l: &[f128], y: &[f128], n: usize) -> Vec<f128> {
    vec_op(ffi::hp_backward_solve, l, y, n)
}
/// L L^T x = b (linalg.rs:123-126).
pub fn solve_cholesky(l: &[f128], b: &[f128], n: usize) -> Vec<f128> {
    vec_op(ffi::hp_solve_cholesky, l, b, n)
}
/// y = A x (linalg.rs:129-147).
pub fn matvec(a: &[f128], x: &[f128], n: usize) -> Vec<f128> {
    vec_op(ffi::hp_matvec, a, x, n)
}

/// (max|L_ii| / min|L_ii|)^2 (linalg.rs:185-194).
pub fn cond_estimate(l: &[f128], n: usize) -> f64 {
    assert_eq!(l.len(), n * n);
    unsafe { ffi::hp_cond_estimate(dd(l), n) }
}

/// dd dot product (linalg.rs:7-18).
pub fn dot(a: &[f128], b: &[f128]) -> f128 {
    assert_eq!(a.len(), b.len());
    let mut out = [f128::ZERO; 1];
    check(unsafe { ffi::hp_dot(dd(a), dd(b), a.len(), dd_mut(&mut out)) });
    out[0]
}

/// y = A x, A rows x cols (linalg.rs:29-40).
pub fn gemv(a: &[f128], rows: usize, cols: usize, x: &[f128], y: &mut [f128]) {
    assert_eq!(a.len(), rows * cols);
    assert_eq!(x.len(), cols);
    assert_eq!(y.len(), rows);
    check(unsafe { ffi::hp_gemv(dd(a), rows, cols, dd(x), dd_mut(y)) });
}

/// Symmetric dd eigendecomposition (linalg.rs:299-362): (eigenvalues, eigenvectors row-wise),
/// or the reference's Err when the sweeps do not converge.
pub fn jacobi_eigen(a: &[f128], n: usize, max_iter: usize) -> Result<(Vec<f128>, Vec<f128>), &'static str> {
    assert_eq!(a.len(), n * n);
    let mut evals = vec![f128::ZERO; n];
    let mut evecs = vec![f128::ZERO; n * n];
    let rc = unsafe { ffi::hp_jacobi_eigen(dd(a), n, max_iter, dd_mut(&mut evals), dd_mut(&mut evecs)) };
    if rc == ffi::HP_ENOCONV {
        return Err("Jacobi eigenvalue did not converge");
    }
    check(rc);
    Ok((evals, evecs))
}

/// mf_cholesky::<N> (multifloat/linalg.rs:89-141).
pub fn mf_cholesky<const N: usize>(a: &mut [MultiFloat<N>], n: usize) -> Result<(), usize> {
    assert_eq!(a.len(), n * n);
    let mut pivot = 0usize;
    let rc = unsafe { ffi::hp_mfn_cholesky(limbs::<N>(), mf_mut(a), n, &mut pivot) };
    if rc == ffi::HP_ENOTSPD {
        return Err(pivot);
    }
    check(rc);
    Ok(())
}

/// mf_cholesky_solve::<N> (multifloat/linalg.rs:145-172).
pub fn mf_cholesky_solve<const N: usize>(l: &[MultiFloat<N>], b: &[MultiFloat<N>], n: usize) -> Vec<MultiFloat<N>> {
    assert_eq!(l.len(), n * n);
    assert_eq!(b.len(), n);
    let mut x = vec![MultiFloat::<N> { limbs: [0.0; N] }; n];
    check(unsafe { ffi::hp_mfn_cholesky_solve(limbs::<N>(), mf(l), mf(b), mf_mut(&mut x), n) });
    x
}

/// mf_gemv::<N> (multifloat/linalg.rs:17-36), A n x m.
pub fn mf_gemv<const N: usize>(a: &[MultiFloat<N>], x: &[MultiFloat<N>], y: &mut [MultiFloat<N>], n: usize,
                               m: usize) {
    assert_eq!(a.len(), n * m);
    assert_eq!(x.len(), m);
    assert_eq!(y.len(), n);
    check(unsafe { ffi::hp_mfn_gemv(limbs::<N>(), mf(a), mf(x), mf_mut(y), n, m) });
}

/// mf_dot::<N> (multifloat/linalg.rs:7-14).
pub fn mf_dot<const N: usize>(a: &[MultiFloat<N>], b: &[MultiFloat<N>]) -> MultiFloat<N> {
    assert_eq!(a.len(), b.len());
    let mut out = MultiFloat::<N> { limbs: [0.0; N] };
    check(unsafe { ffi::hp_mfn_dot(limbs::<N>(), mf(a), mf(b), a.len(), out.limbs.as_mut_ptr()) });
    out
}

/// The normalised s-type Gaussian overlap matrix in dd (the single-primitive s-s case of
/// assemble_overlap_mixed, crates/overlap/src/assemble.rs:35-85, without screening).
pub fn overlap_s(centers: &[[f64; 3]], alpha: &[f64]) -> Vec<f128> {
    let n = alpha.len();
    assert_eq!(centers.len(), n);
    let mut s = vec![f128::ZERO; n * n];
    check(unsafe { ffi::hp_overlap_s(centers.as_ptr() as *const f64, alpha.as_ptr(), n, dd_mut(&mut s)) });
    s
}

/// Mixed-precision SPD solve to a target accuracy (extends iterative_refinement,
/// crates/solver/src/lib.rs:75-125, with quad-double residuals); `factor_kind` is
/// ffi::HP_FACTOR_F64, HP_FACTOR_DD or HP_FACTOR_DD_SPLIT.  Err(pivot) on a breakdown.
pub fn solve_spd_refined(factor_kind: c_int, s: &[f128], b: &[f128], n: usize, tol: f64, max_iter: usize)
                         -> Result<(Vec<f128>, ffi::HpRefineStats), usize> {
    assert_eq!(s.len(), n * n);
    assert_eq!(b.len(), n);
    let mut x = vec![f128::ZERO; n];
    let mut st = ffi::HpRefineStats::default();
    let rc = unsafe {
        ffi::hp_solve_spd_refined(factor_kind, dd(s), dd(b), n, tol, max_iter as c_int, dd_mut(&mut x), &mut st)
    };
    if rc == ffi::HP_ENOTSPD {
        return Err(st.pivot);
    }
    check(rc);
    Ok((x, st))
}

/// The solver crate's entry points (crates/solver/src/lib.rs:38-73).
#[cfg(feature = "solver")]
pub mod solver {
    use super::{check, dd, dd_mut, f128, ffi};
    use quadprec_solver::{Solution, SolveError, Strategy};

    fn solution(rc: std::os::raw::c_int, s: ffi::HpSolution, x: Vec<f128>, x_f64: Vec<f64>)
                -> Result<Solution, SolveError> {
        if rc == ffi::HP_ENOTSPD {
            return Err(SolveError::NotPositiveDefinite { pivot: s.pivot });
        }
        check(rc);
        let strategy = match s.strategy {
            ffi::HP_STRATEGY_FLOAT64 => Strategy::Float64,
            ffi::HP_STRATEGY_MIXED => Strategy::MixedPrecision,
            _ => Strategy::Float128,
        };
        Ok(Solution { x, x_f64, kappa: s.kappa, residual_norm: s.residual_norm, strategy,
                      refinement_iters: s.refinement_iters as usize })
    }

    /// solve_spd (lib.rs:38-68).
    pub fn solve_spd(a_f64: &[f64], b_f64: &[f64], n: usize) -> Result<Solution, SolveError> {
        assert_eq!(a_f64.len(), n * n);
        assert_eq!(b_f64.len(), n);
        let (mut x, mut x64, mut s) = (vec![f128::ZERO; n], vec![0.0; n], ffi::HpSolution::default());
        let rc = unsafe { ffi::hp_solve_spd(a_f64.as_ptr(), b_f64.as_ptr(), n, &mut s, dd_mut(&mut x), x64.as_mut_ptr()) };
        solution(rc, s, x, x64)
    }

    /// solve_spd_f128 (lib.rs:71-73).
    pub fn solve_spd_f128(a: &[f128], b: &[f128], n: usize) -> Result<Solution, SolveError> {
        assert_eq!(a.len(), n * n);
        assert_eq!(b.len(), n);
        let (mut x, mut x64, mut s) = (vec![f128::ZERO; n], vec![0.0; n], ffi::HpSolution::default());
        let rc = unsafe { ffi::hp_solve_spd_f128(dd(a), dd(b), n, &mut s, dd_mut(&mut x), x64.as_mut_ptr()) };
        solution(rc, s, x, x64)
    }
}
