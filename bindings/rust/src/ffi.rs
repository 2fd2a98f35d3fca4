//! Raw declarations of include/hyperprec_hip.h, one `extern "C"` item per function of the
//! header (tests/test_rust_binding.py checks names, argument counts and pointer / scalar
//! kinds against the header).  Status codes, structs and enums mirror the header.
#![allow(non_camel_case_types, dead_code)]
use std::os::raw::{c_char, c_int, c_void};

pub const HP_OK: c_int = 0;
pub const HP_ESHAPE: c_int = 1;
pub const HP_ENOTSPD: c_int = 2;
pub const HP_EDEVICE: c_int = 3;
pub const HP_EINVAL: c_int = 4;
pub const HP_ENOMEM: c_int = 5;
pub const HP_ENOCONV: c_int = 6;

pub const HP_GEMM_PLAIN: c_int = 0;
pub const HP_GEMM_SPLIT_I8: c_int = 2;
pub const HP_GEMM_AUTO: c_int = 3;

pub const HP_STRATEGY_FLOAT64: c_int = 0;
pub const HP_STRATEGY_MIXED: c_int = 1;
pub const HP_STRATEGY_FLOAT128: c_int = 2;

pub const HP_FACTOR_F64: c_int = 0;
pub const HP_FACTOR_DD: c_int = 1;
pub const HP_FACTOR_DD_SPLIT: c_int = 2;

pub const HP_BATCH_PER_INSTANCE: c_int = 1;
pub const HP_BATCH_PANEL_STEPS: c_int = 2;

pub const HP_NOT_CONVERGED: c_int = 0;
pub const HP_CONVERGED_MEASURED: c_int = 1;
pub const HP_CONVERGED_PREDICTED: c_int = 2;

#[repr(C)]
#[derive(Clone, Copy, Default, Debug)]
pub struct HpSolution {
    pub kappa: f64,
    pub residual_norm: f64,
    pub strategy: c_int,
    pub refinement_iters: c_int,
    pub pivot: usize,
}

#[repr(C)]
#[derive(Clone, Copy, Default, Debug)]
pub struct HpRefineStats {
    pub factor_kind: c_int,
    pub iterations: c_int,
    pub factor_ms: f64,
    pub refine_ms: f64,
    pub last_correction: f64,
    pub pivot: usize,
    pub converged: c_int,
}

#[repr(C)]
#[derive(Clone, Copy, Default, Debug)]
pub struct HpGemmPlan {
    pub mode: c_int,
    pub slices: c_int,
    pub w_bits: c_int,
    pub variant: c_int,
    pub ms_plain: f32,
    pub ms_split: f32,
}

#[repr(C)]
#[derive(Clone, Copy, Default, Debug)]
pub struct HpAdjointInfo {
    pub logdet: [f64; 2],
    pub quad: [f64; 2],
    pub refine_iters: c_int,
    pub last_correction: f64,
    pub pivot: usize,
    pub forward_ms: f64,
    pub inverse_ms: f64,
    pub gram_ms: f64,
    pub contract_ms: f64,
}

/// Opaque handles (hp_batch, hp_adjoint).
#[repr(C)]
pub struct HpBatch {
    _private: [u8; 0],
}
#[repr(C)]
pub struct HpAdjoint {
    _private: [u8; 0],
}

#[link(name = "hyperprec_hip")]
unsafe extern "C" {
    pub fn hp_last_error() -> *const c_char;
    pub fn hp_version() -> c_int;
    pub fn hp_gemm(a: *const f64, b: *const f64, c: *mut f64, m: usize, n: usize, k: usize) -> c_int;
    pub fn hp_gemm_atb(a: *const f64, b: *const f64, c: *mut f64, m: usize, n: usize, k: usize) -> c_int;
    pub fn hp_mf_gemm(a: *const f64, b: *const f64, c: *mut f64, m: usize, n: usize, k: usize) -> c_int;
    pub fn hp_mf_gemm_atb(a: *const f64, b: *const f64, c: *mut f64, m: usize, n: usize, k: usize) -> c_int;
    pub fn hp_dev_gemm(trans_a: c_int, a_hi: *const f64, a_lo: *const f64, b_hi: *const f64, b_lo: *const f64, c_hi: *mut f64, c_lo: *mut f64, m: usize, n: usize, k: usize, mode: c_int, slices: c_int, stream: *mut c_void) -> c_int;
    pub fn hp_split_slices_for_bits(k: usize, bits: c_int) -> c_int;
    pub fn hp_gemm_plan_get(trans_a: c_int, m: usize, n: usize, k: usize, bits: c_int, plan: *mut HpGemmPlan) -> c_int;
    pub fn hp_dev_mf_gemm(trans_a: c_int, a: *const f64, b: *const f64, c: *mut f64, m: usize, n: usize, k: usize, stream: *mut c_void) -> c_int;
    pub fn hp_mf_cholesky(a: *mut f64, n: usize, pivot: *mut usize) -> c_int;
    pub fn hp_mf_cholesky_solve(l: *const f64, b: *const f64, x: *mut f64, n: usize) -> c_int;
    pub fn hp_mf_gemv(a: *const f64, x: *const f64, y: *mut f64, n: usize, m: usize) -> c_int;
    pub fn hp_mf4_elementwise(op: c_int, a: *const f64, b: *const f64, out: *mut f64, n: usize) -> c_int;
    pub fn hp_mf4_from_dd(dd: *const f64, mf4: *mut f64, n: usize) -> c_int;
    pub fn hp_dd_from_mf4(mf4: *const f64, dd: *mut f64, n: usize) -> c_int;
    pub fn hp_mfn_gemm(nl: c_int, a: *const f64, b: *const f64, c: *mut f64, m: usize, n: usize, k: usize) -> c_int;
    pub fn hp_mfn_gemm_atb(nl: c_int, a: *const f64, b: *const f64, c: *mut f64, m: usize, n: usize, k: usize) -> c_int;
    pub fn hp_dev_mfn_gemm(nl: c_int, trans_a: c_int, a: *const f64, b: *const f64, c: *mut f64, m: usize, n: usize, k: usize, stream: *mut c_void) -> c_int;
    pub fn hp_mfn_gemv(nl: c_int, a: *const f64, x: *const f64, y: *mut f64, n: usize, m: usize) -> c_int;
    pub fn hp_mfn_dot(nl: c_int, a: *const f64, b: *const f64, n: usize, out: *mut f64) -> c_int;
    pub fn hp_mfn_cholesky(nl: c_int, a: *mut f64, n: usize, pivot: *mut usize) -> c_int;
    pub fn hp_mfn_cholesky_solve(nl: c_int, l: *const f64, b: *const f64, x: *mut f64, n: usize) -> c_int;
    pub fn hp_mfn_elementwise(nl: c_int, op: c_int, a: *const f64, b: *const f64, out: *mut f64, n: usize) -> c_int;
    pub fn hp_mfn_convert(nl_from: c_int, nl_to: c_int, in_: *const f64, out: *mut f64, n: usize) -> c_int;
    pub fn hp_cholesky(a: *mut f64, n: usize, pivot: *mut usize, value: *mut f64) -> c_int;
    pub fn hp_cholesky_blocked(a: *mut f64, n: usize, nb: usize, pivot: *mut usize, value: *mut f64) -> c_int;
    pub fn hp_dev_cholesky(a_hi: *mut f64, a_lo: *mut f64, n: usize, nb: usize, unblocked: c_int, pivot: *mut usize, value: *mut f64, stream: *mut c_void) -> c_int;
    pub fn hp_forward_solve(l: *const f64, b: *const f64, y: *mut f64, n: usize) -> c_int;
    pub fn hp_backward_solve(l: *const f64, y: *const f64, x: *mut f64, n: usize) -> c_int;
    pub fn hp_solve_cholesky(l: *const f64, b: *const f64, x: *mut f64, n: usize) -> c_int;
    pub fn hp_matvec(a: *const f64, x: *const f64, y: *mut f64, n: usize) -> c_int;
    pub fn hp_cond_estimate(l: *const f64, n: usize) -> f64;
    pub fn hp_solve_spd(a: *const f64, b: *const f64, n: usize, sol: *mut HpSolution, x_dd: *mut f64, x_f64: *mut f64) -> c_int;
    pub fn hp_solve_spd_f128(a_dd: *const f64, b_dd: *const f64, n: usize, sol: *mut HpSolution, x_dd: *mut f64, x_f64: *mut f64) -> c_int;
    pub fn hp_dev_solve_spd_refined(factor_kind: c_int, s_hi: *const f64, s_lo: *const f64, b_hi: *const f64, b_lo: *const f64, n: usize, tol: f64, max_iter: c_int, x_hi: *mut f64, x_lo: *mut f64, stats: *mut HpRefineStats, stream: *mut c_void) -> c_int;
    pub fn hp_solve_spd_refined(factor_kind: c_int, s_dd: *const f64, b_dd: *const f64, n: usize, tol: f64, max_iter: c_int, x_dd: *mut f64, stats: *mut HpRefineStats) -> c_int;
    pub fn hp_set_split_factor_panel(nb: c_int) -> c_int;
    pub fn hp_dev_batch_cholesky_solve(sizes: *const c_int, count: usize, nrhs: c_int, s_hi: *const f64, s_lo: *const f64, l_hi: *mut f64, l_lo: *mut f64, b_hi: *const f64, b_lo: *const f64, x_hi: *mut f64, x_lo: *mut f64, status: *mut c_int, stream: *mut c_void) -> c_int;
    pub fn hp_batch_set_schedule(schedule: c_int) -> c_int;
    pub fn hp_batch_graph_create(sizes: *const c_int, count: usize, nrhs: c_int, s_hi: *const f64, s_lo: *const f64, l_hi: *mut f64, l_lo: *mut f64, b_hi: *const f64, b_lo: *const f64, x_hi: *mut f64, x_lo: *mut f64, status: *mut c_int, stream: *mut c_void) -> *mut HpBatch;
    pub fn hp_batch_graph_launch(batch: *mut HpBatch, stream: *mut c_void) -> c_int;
    pub fn hp_batch_graph_schedule(batch: *const HpBatch) -> c_int;
    pub fn hp_batch_graph_destroy(batch: *mut HpBatch);
    pub fn hp_overlap_s(centers: *const f64, alpha: *const f64, n: usize, s_dd: *mut f64) -> c_int;
    pub fn hp_dev_overlap_s(centers: *const f64, alpha: *const f64, n: usize, s_hi: *mut f64, s_lo: *mut f64, stream: *mut c_void) -> c_int;
    pub fn hp_assemble_overlap_mixed(centers: *const f64, l: *const c_int, prim_off: *const i64, alpha: *const f64, coeff: *const f64, n_shells: usize, params4: *const f64, s: *mut f64, n_ao: usize, counts: *mut usize) -> c_int;
    pub fn hp_dev_assemble_overlap_mixed(centers: *const f64, l: *const c_int, prim_off: *const i64, alpha: *const f64, coeff: *const f64, n_shells: usize, params4: *const f64, s_dev: *mut f64, n_ao: usize, counts: *mut usize, stream: *mut c_void) -> c_int;
    pub fn hp_jacobi_eigen(a: *const f64, n: usize, max_iter: usize, evals: *mut f64, evecs: *mut f64) -> c_int;
    pub fn hp_dev_jacobi_eigen(a_hi: *const f64, a_lo: *const f64, n: usize, max_sweeps: usize, stop_rule: c_int, evals_hi: *mut f64, evals_lo: *mut f64, vecs_hi: *mut f64, vecs_lo: *mut f64, vec_rows: c_int, sweeps: *mut usize, stream: *mut c_void) -> c_int;
    pub fn hp_eigen_set_block_min(min_n: usize) -> c_int;
    pub fn hp_canonical_orthogonalize(s_dd: *const f64, n: usize, threshold: f64, x_dd: *mut f64, n_orth: *mut usize) -> c_int;
    pub fn hp_solve_gen_eigen(f_dd: *const f64, s_dd: *const f64, n: usize, threshold: f64, evals_dd: *mut f64, c_dd: *mut f64, n_basis: *mut usize) -> c_int;
    pub fn hp_dev_solve_gen_eigen(f_hi: *const f64, f_lo: *const f64, s_hi: *const f64, s_lo: *const f64, n: usize, threshold: f64, evals_hi: *mut f64, evals_lo: *mut f64, c_hi: *mut f64, c_lo: *mut f64, n_basis: *mut usize, sweeps: *mut usize, stream: *mut c_void) -> c_int;
    pub fn hp_dot(a: *const f64, b: *const f64, n: usize, out_dd: *mut f64) -> c_int;
    pub fn hp_dev_dot(a_hi: *const f64, a_lo: *const f64, b_hi: *const f64, b_lo: *const f64, n: usize, out_hi: *mut f64, out_lo: *mut f64, stream: *mut c_void) -> c_int;
    pub fn hp_dev_dot_noalloc(a_hi: *const f64, a_lo: *const f64, b_hi: *const f64, b_lo: *const f64, n: usize, out_hi: *mut f64, out_lo: *mut f64, stream: *mut c_void) -> c_int;
    pub fn hp_gemv(a: *const f64, rows: usize, cols: usize, x: *const f64, y: *mut f64) -> c_int;
    pub fn hp_dev_gemv(a_hi: *const f64, a_lo: *const f64, rows: usize, cols: usize, x_hi: *const f64, x_lo: *const f64, y_hi: *mut f64, y_lo: *mut f64, stream: *mut c_void) -> c_int;
    pub fn hp_neumaier_sum_f64(v: *const f64, n: usize, out: *mut f64) -> c_int;
    pub fn hp_neumaier_sum(v: *const f64, n: usize, out_dd: *mut f64) -> c_int;
    pub fn hp_par_neumaier_sum_f64(v: *const f64, n: usize, out: *mut f64) -> c_int;
    pub fn hp_par_neumaier_sum(v: *const f64, n: usize, out_dd: *mut f64) -> c_int;
    pub fn hp_format_f64(v: *const f64, n: usize, precision: c_int, force_sign: c_int, text: *mut c_char, cap: usize, offsets: *mut usize) -> c_int;
    pub fn hp_dd_to_string(v_dd: *const f64, n: usize, text: *mut c_char, cap: usize, offsets: *mut usize) -> c_int;
    pub fn hp_mf4_to_string(v_mf4: *const f64, n: usize, text: *mut c_char, cap: usize, offsets: *mut usize) -> c_int;
    pub fn hp_dd_serialize_str(v_dd: *const f64, n: usize, text: *mut c_char, cap: usize, offsets: *mut usize) -> c_int;
    pub fn hp_mf4_serialize_str(v_mf4: *const f64, n: usize, text: *mut c_char, cap: usize, offsets: *mut usize) -> c_int;
    pub fn hp_parse_f64(text: *const c_char, offsets: *const usize, n: usize, out: *mut f64, status: *mut u8) -> c_int;
    pub fn hp_dd_from_str(text: *const c_char, offsets: *const usize, n: usize, out_dd: *mut f64, status: *mut u8) -> c_int;
    pub fn hp_mf4_from_str(text: *const c_char, offsets: *const usize, n: usize, out_mf4: *mut f64, status: *mut u8) -> c_int;
    pub fn hp_dd_deserialize_str(text: *const c_char, offsets: *const usize, n: usize, out_dd: *mut f64, status: *mut u8) -> c_int;
    pub fn hp_mf4_deserialize_str(text: *const c_char, offsets: *const usize, n: usize, out_mf4: *mut f64, status: *mut u8) -> c_int;
    pub fn hp_dev_format(kind: c_int, planes: *const *const f64, n: usize, precision: c_int, force_sign: c_int, text: *mut c_char, cap: usize, offsets: *mut usize, need: *mut usize, stream: *mut c_void) -> c_int;
    pub fn hp_dev_parse(kind: c_int, text: *const c_char, offsets: *const usize, n: usize, planes: *const *mut f64, status: *mut u8, n_bad: *mut usize, stream: *mut c_void) -> c_int;
    pub fn hp_dev_overlap_forward(centers: *const f64, alpha: *const f64, b_hi: *const f64, b_lo: *const f64, n: usize, factor_kind: c_int, ctx: *mut *mut HpAdjoint, info: *mut HpAdjointInfo, stream: *mut c_void) -> c_int;
    pub fn hp_dev_overlap_backward(ctx: *mut HpAdjoint, cot: *const f64, gemm_mode: c_int, gc_hi: *mut f64, gc_lo: *mut f64, ga_hi: *mut f64, ga_lo: *mut f64, gb_hi: *mut f64, gb_lo: *mut f64, stream: *mut c_void) -> c_int;
    pub fn hp_dev_overlap_gradients(ctx: *mut HpAdjoint, gemm_mode: c_int, gld_hi: *mut f64, gld_lo: *mut f64, gq_hi: *mut f64, gq_lo: *mut f64, stream: *mut c_void) -> c_int;
    pub fn hp_adjoint_get_info(ctx: *const HpAdjoint, info: *mut HpAdjointInfo) -> c_int;
    pub fn hp_adjoint_destroy(ctx: *mut HpAdjoint);
    pub fn hp_overlap_adjoint(centers: *const f64, alpha: *const f64, b_dd: *const f64, n: usize, factor_kind: c_int, gemm_mode: c_int, info: *mut HpAdjointInfo, g_ld_dd: *mut f64, g_q_dd: *mut f64) -> c_int;
    pub fn hp_set_phase_timing(on: c_int);
    pub fn hp_collect_phase_times(ms_sum: *mut f64, max: c_int, calls: *mut c_int) -> c_int;
    pub fn hp_split_gemm_variant(m: usize, n: usize, k: usize, slices: c_int) -> c_int;
    pub fn hp_split_gemm_set_variant(m: usize, n: usize, k: usize, slices: c_int, variant: c_int) -> c_int;
    pub fn hp_get_phase_times(ms: *mut f64, max: c_int) -> c_int;
    pub fn hp_gen_dd_normal(seed: u64, count: usize, out_aos: *mut f64);
    pub fn hp_gen_mf4_normal(seed: u64, count: usize, out_aos: *mut f64);
    pub fn hp_gen_overlap_basis(seed: u64, n: usize, alpha_lo: f64, alpha_hi: f64, spacing: f64, jitter: f64, centers: *mut f64, alpha: *mut f64);
}
