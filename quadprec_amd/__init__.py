"""quadprec_amd -- MI355X (gfx950) implementation of hyperprec's extended-precision path.

Host-side mirror of the reference's Rust API (crate `hyperprec`, crates/quad/src/lib.rs:9-14;
crates/quad/src/multifloat/mod.rs:19-22; crates/solver/src/lib.rs:38,71), calling the
hand-written HIP kernels through the C ABI of include/hyperprec_hip.h.

Array conventions (numpy float64, host memory, the reference's memory layout):
  dd   : shape (count, 2) -> [hi, lo]
  MF4  : shape (count, 4) -> limbs, most significant first
Device entry points (module ``quadprec_amd.dev``) take torch tensors resident in HBM as
structure-of-arrays planes.

There is no CPU fallback: if the HIP library is missing or no gfx950 device is usable
the functions raise.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libhyperprec_hip.so")

HP_OK, HP_ESHAPE, HP_ENOTSPD, HP_EDEVICE, HP_EINVAL, HP_ENOMEM, HP_ENOCONV = range(7)
HP_GEMM_PLAIN = 0
HP_GEMM_SPLIT_I8 = 2
HP_GEMM_AUTO = 3     # slices = accuracy in bits (0: 104); see gemm_plan

_lib = None
_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_sz = C.c_size_t
_vp = C.c_void_p


class HyperprecError(RuntimeError):
    pass


class Solution(C.Structure):
    """Solution{kappa, residual_norm, strategy, refinement_iters} (crates/solver/src/lib.rs:10-23);
    x and x_f64 are attached by solve_spd / solve_spd_f128."""
    _fields_ = [("kappa", C.c_double), ("residual_norm", C.c_double), ("strategy", C.c_int),
                ("refinement_iters", C.c_int), ("pivot", C.c_size_t)]


class RefineStats(C.Structure):
    """hp_refine_stats; `converged` is HP_CONVERGED_MEASURED (the last measured correction is
    below tol), HP_CONVERGED_PREDICTED (the next one is predicted ten times below tol) or
    HP_NOT_CONVERGED (stagnation or max_iter)."""
    _fields_ = [("factor_kind", C.c_int), ("iterations", C.c_int), ("factor_ms", C.c_double),
                ("refine_ms", C.c_double), ("last_correction", C.c_double), ("pivot", C.c_size_t),
                ("converged", C.c_int)]


HP_FACTOR_F64, HP_FACTOR_DD, HP_FACTOR_DD_SPLIT = 0, 1, 2
HP_BATCH_AUTO, HP_BATCH_PER_INSTANCE, HP_BATCH_PANEL_STEPS = 0, 1, 2
HP_NOT_CONVERGED, HP_CONVERGED_MEASURED, HP_CONVERGED_PREDICTED = 0, 1, 2


class GemmPlan(C.Structure):
    """hp_gemm_plan: the form HP_GEMM_AUTO chose for a shape class and accuracy."""
    _fields_ = [("mode", C.c_int), ("slices", C.c_int), ("w_bits", C.c_int), ("variant", C.c_int),
                ("ms_plain", C.c_float), ("ms_split", C.c_float)]


class AdjointInfo(C.Structure):
    """hp_adjoint_info: dd log det S and b^T S^-1 b of the forward, refinement and timings."""
    _fields_ = [("logdet", C.c_double * 2), ("quad", C.c_double * 2), ("refine_iters", C.c_int),
                ("last_correction", C.c_double), ("pivot", C.c_size_t), ("forward_ms", C.c_double),
                ("inverse_ms", C.c_double), ("gram_ms", C.c_double), ("contract_ms", C.c_double)]


class Strategy:
    """crates/solver/src/lib.rs:25-30"""
    Float64, MixedPrecision, Float128 = 0, 1, 2


class SolveError(HyperprecError):
    """SolveError::NotPositiveDefinite{pivot} (crates/solver/src/lib.rs:251-265)."""

    def __init__(self, pivot):
        super().__init__(f"Matrix is not positive definite (failed at pivot {pivot})")
        self.pivot = pivot


class ShapeError(HyperprecError, AssertionError):
    """The reference panics via assert_eq! on inconsistent sizes (linalg.rs:151-153)."""


class CholeskyError(HyperprecError):
    """Mirror of CholeskyError{pivot, value} (crates/quad/src/linalg.rs:196-208)."""

    def __init__(self, pivot, value):
        super().__init__(f"Cholesky failed at pivot {pivot} (value={value:.4e})")
        self.pivot = pivot
        self.value = value


def lib():
    """Loads the HIP library; raises if it has not been built (no CPU fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise HyperprecError(f"{LIB_PATH} is missing: run `python -m quadprec_amd.build` first")
    L = C.CDLL(LIB_PATH)
    L.hp_last_error.restype = C.c_char_p
    L.hp_version.restype = C.c_int
    for name in ("hp_gemm", "hp_gemm_atb", "hp_mf_gemm", "hp_mf_gemm_atb"):
        getattr(L, name).argtypes = [_dp, _dp, _dp, _sz, _sz, _sz]
    L.hp_dev_gemm.argtypes = [C.c_int, _vp, _vp, _vp, _vp, _vp, _vp, _sz, _sz, _sz, C.c_int, C.c_int, _vp]
    L.hp_dev_mf_gemm.argtypes = [C.c_int, _vp, _vp, _vp, _sz, _sz, _sz, _vp]
    L.hp_cholesky.argtypes = [_dp, _sz, C.POINTER(C.c_size_t), C.POINTER(C.c_double)]
    L.hp_cholesky_blocked.argtypes = [_dp, _sz, _sz, C.POINTER(C.c_size_t), C.POINTER(C.c_double)]
    L.hp_dev_cholesky.argtypes = [_vp, _vp, _sz, _sz, C.c_int, C.POINTER(C.c_size_t), C.POINTER(C.c_double), _vp]
    for name in ("hp_forward_solve", "hp_backward_solve", "hp_solve_cholesky", "hp_matvec"):
        getattr(L, name).argtypes = [_dp, _dp, _dp, _sz]
    L.hp_cond_estimate.argtypes = [_dp, _sz]
    L.hp_cond_estimate.restype = C.c_double
    for name in ("hp_solve_spd", "hp_solve_spd_f128"):
        getattr(L, name).argtypes = [_dp, _dp, _sz, C.POINTER(Solution), _dp, _dp]
    L.hp_dev_solve_spd_refined.argtypes = [C.c_int, _vp, _vp, _vp, _vp, _sz, C.c_double, C.c_int, _vp, _vp,
                                           C.POINTER(RefineStats), _vp]
    L.hp_solve_spd_refined.argtypes = [C.c_int, _dp, _dp, _sz, C.c_double, C.c_int, _dp, C.POINTER(RefineStats)]
    L.hp_set_split_factor_panel.argtypes = [C.c_int]
    L.hp_set_split_factor_panel.restype = C.c_int
    _ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
    L.hp_dev_batch_cholesky_solve.argtypes = [_ip, _sz, C.c_int] + [_vp] * 9 + [_vp]
    L.hp_batch_graph_create.argtypes = [_ip, _sz, C.c_int] + [_vp] * 9 + [_vp]
    L.hp_batch_graph_create.restype = _vp
    L.hp_batch_graph_launch.argtypes = [_vp, _vp]
    L.hp_batch_graph_destroy.argtypes = [_vp]
    L.hp_batch_set_schedule.argtypes = [C.c_int]
    L.hp_batch_graph_schedule.argtypes = [_vp]
    L.hp_batch_graph_schedule.restype = C.c_int
    L.hp_batch_graph_destroy.restype = None
    L.hp_assemble_overlap_mixed.argtypes = [_dp, np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS"),
                                            np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS"),
                                            _dp, _dp, _sz, _dp, _dp, _sz, C.POINTER(C.c_size_t)]
    L.hp_dev_assemble_overlap_mixed.argtypes = [_dp, np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS"),
                                                np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS"),
                                                _dp, _dp, _sz, _dp, _vp, _sz, C.POINTER(C.c_size_t), _vp]
    L.hp_jacobi_eigen.argtypes = [_dp, _sz, _sz, _dp, _dp]
    L.hp_dev_jacobi_eigen.argtypes = [_vp, _vp, _sz, _sz, C.c_int, _vp, _vp, _vp, _vp, C.c_int,
                                      C.POINTER(C.c_size_t), _vp]
    L.hp_canonical_orthogonalize.argtypes = [_dp, _sz, C.c_double, _dp, C.POINTER(C.c_size_t)]
    L.hp_solve_gen_eigen.argtypes = [_dp, _dp, _sz, C.c_double, _dp, _dp, C.POINTER(C.c_size_t)]
    L.hp_dev_solve_gen_eigen.argtypes = [_vp, _vp, _vp, _vp, _sz, C.c_double, _vp, _vp, _vp, _vp,
                                         C.POINTER(C.c_size_t), C.POINTER(C.c_size_t), _vp]
    L.hp_dot.argtypes = [_dp, _dp, _sz, _dp]
    L.hp_dev_dot.argtypes = [_vp] * 4 + [_sz, _vp, _vp, _vp]
    L.hp_dev_dot_noalloc.argtypes = [_vp] * 4 + [_sz, _vp, _vp, _vp]
    L.hp_gemv.argtypes = [_dp, _sz, _sz, _dp, _dp]
    L.hp_dev_gemv.argtypes = [_vp, _vp, _sz, _sz, _vp, _vp, _vp, _vp, _vp]
    for name in ("hp_neumaier_sum_f64", "hp_neumaier_sum", "hp_par_neumaier_sum_f64", "hp_par_neumaier_sum"):
        getattr(L, name).argtypes = [_dp, _sz, _dp]
    _op = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")
    _u8 = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")
    L.hp_format_f64.argtypes = [_dp, _sz, C.c_int, C.c_int, C.c_char_p, _sz, _op]
    for name in ("hp_dd_to_string", "hp_mf4_to_string", "hp_dd_serialize_str", "hp_mf4_serialize_str"):
        getattr(L, name).argtypes = [_dp, _sz, C.c_char_p, _sz, _op]
    for name in ("hp_parse_f64", "hp_dd_from_str", "hp_mf4_from_str", "hp_dd_deserialize_str",
                 "hp_mf4_deserialize_str"):
        getattr(L, name).argtypes = [C.c_char_p, _op, _sz, _dp, _u8]
    L.hp_dev_format.argtypes = [C.c_int, _vp, _sz, C.c_int, C.c_int, _vp, _sz, _vp, C.POINTER(C.c_size_t), _vp]
    L.hp_dev_parse.argtypes = [C.c_int, _vp, _vp, _sz, _vp, _vp, C.POINTER(C.c_size_t), _vp]
    L.hp_mf_cholesky.argtypes = [_dp, _sz, C.POINTER(C.c_size_t)]
    L.hp_mf_cholesky_solve.argtypes = [_dp, _dp, _dp, _sz]
    L.hp_mf_gemv.argtypes = [_dp, _dp, _dp, _sz, _sz]
    L.hp_mf4_elementwise.argtypes = [C.c_int, _dp, _dp, _dp, _sz]
    L.hp_mf4_from_dd.argtypes = [_dp, _dp, _sz]
    L.hp_eigen_set_block_min.argtypes = [_sz]
    for name in ("hp_mfn_gemm", "hp_mfn_gemm_atb"):
        getattr(L, name).argtypes = [C.c_int, _dp, _dp, _dp, _sz, _sz, _sz]
    L.hp_dev_mfn_gemm.argtypes = [C.c_int, C.c_int, _vp, _vp, _vp, _sz, _sz, _sz, _vp]
    L.hp_mfn_gemv.argtypes = [C.c_int, _dp, _dp, _dp, _sz, _sz]
    L.hp_mfn_dot.argtypes = [C.c_int, _dp, _dp, _sz, _dp]
    L.hp_mfn_cholesky.argtypes = [C.c_int, _dp, _sz, C.POINTER(C.c_size_t)]
    L.hp_mfn_cholesky_solve.argtypes = [C.c_int, _dp, _dp, _dp, _sz]
    L.hp_mfn_elementwise.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp, _sz]
    L.hp_mfn_convert.argtypes = [C.c_int, C.c_int, _dp, _dp, _sz]
    L.hp_dd_from_mf4.argtypes = [_dp, _dp, _sz]
    L.hp_overlap_s.argtypes = [_dp, _dp, _sz, _dp]
    L.hp_dev_overlap_s.argtypes = [_vp, _vp, _sz, _vp, _vp, _vp]
    L.hp_dev_overlap_forward.argtypes = [_vp, _vp, _vp, _vp, _sz, C.c_int, C.POINTER(_vp), C.POINTER(AdjointInfo),
                                         _vp]
    L.hp_dev_overlap_backward.argtypes = [_vp, _dp, C.c_int] + [_vp] * 6 + [_vp]
    L.hp_dev_overlap_gradients.argtypes = [_vp, C.c_int] + [_vp] * 4 + [_vp]
    L.hp_adjoint_get_info.argtypes = [_vp, C.POINTER(AdjointInfo)]
    L.hp_adjoint_destroy.argtypes = [_vp]
    L.hp_adjoint_destroy.restype = None
    L.hp_overlap_adjoint.argtypes = [_dp, _dp, _dp, _sz, C.c_int, C.c_int, C.POINTER(AdjointInfo), _dp, _dp]
    L.hp_gen_overlap_basis.argtypes = [C.c_uint64, _sz, C.c_double, C.c_double, C.c_double, C.c_double, _dp, _dp]
    L.hp_gen_overlap_basis.restype = None
    L.hp_split_gemm_variant.argtypes = [_sz, _sz, _sz, C.c_int]
    L.hp_split_gemm_variant.restype = C.c_int
    L.hp_split_gemm_set_variant.argtypes = [_sz, _sz, _sz, C.c_int, C.c_int]
    L.hp_split_gemm_set_variant.restype = C.c_int
    L.hp_split_slices_for_bits.argtypes = [_sz, C.c_int]
    L.hp_split_slices_for_bits.restype = C.c_int
    L.hp_gemm_plan_get.argtypes = [C.c_int, _sz, _sz, _sz, C.c_int, C.POINTER(GemmPlan)]
    L.hp_gemm_plan_get.restype = C.c_int
    L.hp_set_phase_timing.argtypes = [C.c_int]
    L.hp_set_phase_timing.restype = None
    L.hp_get_phase_times.argtypes = [_dp, C.c_int]
    L.hp_get_phase_times.restype = C.c_int
    L.hp_collect_phase_times.argtypes = [_dp, C.c_int, C.POINTER(C.c_int)]
    L.hp_collect_phase_times.restype = C.c_int
    L.hp_gen_dd_normal.argtypes = [C.c_uint64, _sz, _dp]
    L.hp_gen_dd_normal.restype = None
    L.hp_gen_mf4_normal.argtypes = [C.c_uint64, _sz, _dp]
    L.hp_gen_mf4_normal.restype = None
    _lib = L
    return L


def _check(rc, what):
    if rc == HP_OK:
        return
    msg = lib().hp_last_error().decode(errors="replace")
    if rc == HP_ESHAPE:
        raise ShapeError(f"{what}: {msg}")
    raise HyperprecError(f"{what}: status {rc}: {msg}")


def _arr(x, limbs, count, name):
    a = np.ascontiguousarray(x, dtype=np.float64)
    if a.size != limbs * count:
        raise ShapeError(f"{name}: expected {count} values of {limbs} limbs, got {a.size // limbs}")
    return a


def _out(out, limbs, count, name="out"):
    """A caller-supplied output: float64, C-contiguous, writable and exactly count values of
    limbs doubles (the library writes that many), else ShapeError."""
    if out is None:
        return np.zeros((count, limbs))
    if not isinstance(out, np.ndarray) or out.dtype != np.float64 or not out.flags.c_contiguous \
            or not out.flags.writeable or out.size != limbs * count:
        raise ShapeError(f"{name}: expected a writable C-contiguous float64 array of {limbs * count} doubles")
    return out


# ---------------------------------------------------------------- GEMM (linalg.rs:150-182)
def gemm(a, b, m, n, k, out=None):
    """C = A*B in double-double (crates/quad/src/linalg.rs:150-165). Returns C (m*n, 2)."""
    a = _arr(a, 2, m * k, "a")
    b = _arr(b, 2, k * n, "b")
    c = _out(out, 2, m * n)
    _check(lib().hp_gemm(a, b, c, m, n, k), "gemm")
    return c


def gemm_atb(a, b, m, n, k, out=None):
    """C = A^T*B with A stored k x m (linalg.rs:168-182)."""
    a = _arr(a, 2, k * m, "a")
    b = _arr(b, 2, k * n, "b")
    c = _out(out, 2, m * n)
    _check(lib().hp_gemm_atb(a, b, c, m, n, k), "gemm_atb")
    return c


def mf_gemm(a, b, m, n, k, out=None):
    """MultiFloat<4> C = A*B (crates/quad/src/multifloat/linalg.rs:39-61)."""
    a = _arr(a, 4, m * k, "a")
    b = _arr(b, 4, k * n, "b")
    c = _out(out, 4, m * n)
    _check(lib().hp_mf_gemm(a, b, c, m, n, k), "mf_gemm")
    return c


def mf_gemm_atb(a, b, m, n, k, out=None):
    """MultiFloat<4> C = A^T*B, A stored k x m (multifloat/linalg.rs:64-85)."""
    a = _arr(a, 4, k * m, "a")
    b = _arr(b, 4, k * n, "b")
    c = _out(out, 4, m * n)
    _check(lib().hp_mf_gemm_atb(a, b, c, m, n, k), "mf_gemm_atb")
    return c


# ---------------------------------------------------------------- Cholesky and solves
def _chol(fn, a, n, *extra):
    a = _arr(a, 2, n * n, "a").copy()
    pv, vv = C.c_size_t(0), C.c_double(0.0)
    rc = fn(a, n, *extra, C.byref(pv), C.byref(vv))
    if rc == HP_ENOTSPD:
        raise CholeskyError(pv.value, vv.value)
    _check(rc, "cholesky")
    return a


def cholesky(a, n):
    """Unblocked dd Cholesky (crates/quad/src/linalg.rs:42-96); returns the factored copy
    (lower triangle = L, upper keeps A).  Raises CholeskyError like the reference."""
    return _chol(lib().hp_cholesky, a, n)


def cholesky_blocked(a, n, nb=0):
    """Blocked dd Cholesky (linalg.rs:238-293), block size nb (0 -> 64)."""
    return _chol(lib().hp_cholesky_blocked, a, n, nb)


def _vec(name, l, b, n):
    l = _arr(l, 2, n * n, "l")
    b = _arr(b, 2, n, "b")
    out = np.zeros((n, 2))
    _check(getattr(lib(), name)(l, b, out, n), name)
    return out


def forward_solve(l, b, n):
    """L y = b (linalg.rs:99-108)."""
    return _vec("hp_forward_solve", l, b, n)


def backward_solve(l, y, n):
    """L^T x = y (linalg.rs:111-120)."""
    return _vec("hp_backward_solve", l, y, n)


def solve_cholesky(l, b, n):
    """L L^T x = b (linalg.rs:123-126)."""
    return _vec("hp_solve_cholesky", l, b, n)


def matvec(a, x, n):
    """y = A x (linalg.rs:129-147)."""
    return _vec("hp_matvec", a, x, n)


def cond_estimate(l, n):
    """(max|L_ii|/min|L_ii|)^2 (linalg.rs:185-194)."""
    return lib().hp_cond_estimate(_arr(l, 2, n * n, "l"), n)


def _solve(name, a, b, n, limbs):
    a = _arr(a, limbs, n * n, "a")
    b = _arr(b, limbs, n, "b")
    sol = Solution()
    x = np.zeros((n, 2))
    xf = np.zeros(n)
    rc = getattr(lib(), name)(a, b, n, C.byref(sol), x, xf)
    if rc == HP_ENOTSPD:
        raise SolveError(sol.pivot)
    _check(rc, name)
    sol.x = x
    sol.x_f64 = xf
    return sol


def solve_spd(a, b, n):
    """solve_spd (crates/solver/src/lib.rs:38-68): f64 A (n*n) and b (n)."""
    return _solve("hp_solve_spd", a, b, n, 1)


def solve_spd_f128(a, b, n):
    """solve_spd_f128 (lib.rs:71-73): dd A and b."""
    return _solve("hp_solve_spd_f128", a, b, n, 2)


def solve_spd_refined(s, b, n, factor_kind=HP_FACTOR_DD_SPLIT, tol=1e-30, max_iter=20):
    """Mixed-precision SPD solve with quad-double residual refinement (dd S and b).
    Returns (x as (n, 2) dd, RefineStats)."""
    s = _arr(s, 2, n * n, "s")
    b = _arr(b, 2, n, "b")
    x = np.zeros((n, 2))
    st = RefineStats()
    rc = lib().hp_solve_spd_refined(factor_kind, s, b, n, tol, max_iter, x, C.byref(st))
    if rc == HP_ENOTSPD:
        raise SolveError(st.pivot)
    _check(rc, "solve_spd_refined")
    return x, st


def set_split_factor_panel(nb):
    """Panel width (64, 128, 256, 512) of the split-dd factorisation."""
    _check(lib().hp_set_split_factor_panel(nb), "set_split_factor_panel")


# ---------------------------------------------------------------- overlap
# ---------------------------------------------------------------- contracted overlap
SCREEN_DEFAULT = (0.1, 500.0, 1e-14, 12.0)  # ScreenParams::default() (crates/overlap/src/screen.rs:38-46)


class TierCounts(tuple):
    """TierCounts {f64_pairs, f128_pairs, skipped_pairs} (crates/overlap/src/assemble.rs:17-22)."""

    f64_pairs = property(lambda self: self[0])
    f128_pairs = property(lambda self: self[1])
    skipped_pairs = property(lambda self: self[2])


def assemble_overlap_mixed(shells, params=SCREEN_DEFAULT):
    """assemble_overlap_mixed (crates/overlap/src/assemble.rs:35-85) on the GPU.  shells: list
    of (center (3,), l in 0..3, [(alpha, coeff), ...]).  Returns (S (n_ao, n_ao) f64, TierCounts)."""
    centers = np.ascontiguousarray([c for c, _, _ in shells], dtype=np.float64).ravel()
    ls = np.ascontiguousarray([l for _, l, _ in shells], dtype=np.int32)
    off = np.zeros(len(shells) + 1, dtype=np.int64)
    for i, (_, _, prims) in enumerate(shells):
        off[i + 1] = off[i] + len(prims)
    alpha = np.ascontiguousarray([a for _, _, prims in shells for a, _ in prims], dtype=np.float64)
    coeff = np.ascontiguousarray([c for _, _, prims in shells for _, c in prims], dtype=np.float64)
    n_ao = int(sum((l + 1) * (l + 2) // 2 for l in ls))
    s = np.zeros(max(n_ao * n_ao, 1))
    counts = (C.c_size_t * 3)()
    _check(lib().hp_assemble_overlap_mixed(centers, ls, off, alpha if alpha.size else np.zeros(1),
                                           coeff if coeff.size else np.zeros(1), len(shells),
                                           np.ascontiguousarray(params, dtype=np.float64), s, n_ao, counts),
           "assemble_overlap_mixed")
    return s[:n_ao * n_ao].reshape(n_ao, n_ao), TierCounts(tuple(counts))


# ---------------------------------------------------------------- eigensolvers
class NotPositiveDefinite(HyperprecError):
    """mf_cholesky's Err("Cholesky failed: matrix is not positive definite")
    (crates/quad/src/multifloat/linalg.rs:103-105); .pivot names the failing column."""
    pivot = None


class NotConverged(HyperprecError):
    """jacobi_eigen's Err("Jacobi eigenvalue solver did not converge ...") (linalg.rs:360)."""


def jacobi_eigen(a, n, max_iter):
    """Symmetric dd eigendecomposition (crates/quad/src/linalg.rs:299-362): (eigenvalues (n, 2)
    in diagonal order, eigenvectors (n*n, 2) with row i = eigenvector i)."""
    a = _arr(a, 2, n * n, "a")
    ev = np.zeros((max(n, 1), 2))
    vec = np.zeros((max(n * n, 1), 2))
    rc = lib().hp_jacobi_eigen(a, n, max_iter, ev, vec)
    if rc == HP_ENOCONV:
        raise NotConverged(lib().hp_last_error().decode(errors="replace"))
    _check(rc, "jacobi_eigen")
    return ev[:n], vec[:n * n]


def canonical_orthogonalize(s, n, threshold):
    """Canonical orthogonalization (crates/solver/src/ortho.rs:158-193): (X (n*n_orth, 2), n_orth)."""
    s = _arr(s, 2, n * n, "s")
    x = np.zeros((max(n * n, 1), 2))
    m = C.c_size_t(0)
    _check(lib().hp_canonical_orthogonalize(s, n, threshold, x, C.byref(m)), "canonical_orthogonalize")
    return x[:n * m.value], m.value


def solve_gen_eigen(f, s, n, threshold):
    """F C = S C eps (crates/solver/src/ortho.rs:201-248): (eigenvalues (m, 2) ascending,
    C (n*m, 2) with column k = eigenvector k, m)."""
    f = _arr(f, 2, n * n, "f")
    s = _arr(s, 2, n * n, "s")
    ev = np.zeros((max(n, 1), 2))
    c = np.zeros((max(n * n, 1), 2))
    m = C.c_size_t(0)
    _check(lib().hp_solve_gen_eigen(f, s, n, threshold, ev, c, C.byref(m)), "solve_gen_eigen")
    return ev[:m.value], c[:n * m.value], m.value


def eigen_set_block_min(n):
    """From n = min_n on the Jacobi eigensolvers run the block form (0: the default, 768;
    a huge value disables it).  Process-wide."""
    _check(lib().hp_eigen_set_block_min(int(n)), "eigen_set_block_min")


# ---------------------------------------------------------------- MultiFloat<4> linear algebra
def mf_cholesky(a, n):
    """mf_cholesky::<4> (crates/quad/src/multifloat/linalg.rs:89-141): the factored copy (n*n, 4).
    Raises NotPositiveDefinite (with .pivot) like the reference's Err."""
    a = _arr(a, 4, n * n, "a").copy().reshape(-1, 4)
    piv = C.c_size_t(0)
    rc = lib().hp_mf_cholesky(a if a.size else np.zeros(4), n, C.byref(piv))
    if rc == HP_ENOTSPD:
        e = NotPositiveDefinite(lib().hp_last_error().decode(errors="replace"))
        e.pivot = piv.value
        raise e
    _check(rc, "mf_cholesky")
    return a


def mf_cholesky_solve(l, b, n):
    """mf_cholesky_solve::<4> (multifloat/linalg.rs:145-172): x (n, 4)."""
    l = _arr(l, 4, n * n, "l")
    b = _arr(b, 4, n, "b")
    x = np.zeros((max(n, 1), 4))
    _check(lib().hp_mf_cholesky_solve(l if l.size else np.zeros(4), b if b.size else np.zeros(4), x, n),
           "mf_cholesky_solve")
    return x[:n]


def mf_gemv(a, x, n, m):
    """mf_gemv::<4> (multifloat/linalg.rs:17-36): y = A x, A n x m; (n, 4)."""
    a = _arr(a, 4, n * m, "a")
    x = _arr(x, 4, m, "x")
    y = np.zeros((max(n, 1), 4))
    _check(lib().hp_mf_gemv(a if a.size else np.zeros(4), x if x.size else np.zeros(4), y, n, m), "mf_gemv")
    return y[:n]


def mf4_elementwise(op, a, b=None):
    """MF4 recip / sqrt / div over arrays (multifloat/div.rs:14-77, math.rs:22-40): op in
    {"recip", "sqrt", "div"}; a (and b) (n, 4)."""
    code = {"recip": 0, "sqrt": 1, "div": 2}[op]
    a = np.ascontiguousarray(a, dtype=np.float64).reshape(-1, 4)
    n = a.shape[0]
    bb = np.ascontiguousarray(b, dtype=np.float64).reshape(-1, 4) if b is not None else np.zeros((1, 4))
    out = np.zeros((max(n, 1), 4))
    _check(lib().hp_mf4_elementwise(code, a if n else np.zeros(4), bb, out, n), "mf4_elementwise")
    return out[:n]


def mf4_from_dd(v):
    """MultiFloat::<4>::from_f128 (multifloat/conv.rs:120-125) of each dd: (n, 4)."""
    v = np.ascontiguousarray(v, dtype=np.float64).reshape(-1, 2)
    out = np.zeros((max(len(v), 1), 4))
    _check(lib().hp_mf4_from_dd(v if len(v) else np.zeros(2), out, len(v)), "mf4_from_dd")
    return out[:len(v)]


def dd_from_mf4(v):
    """MultiFloat::<4>::to_f128 (multifloat/conv.rs:127-129) of each MF4: (n, 2)."""
    v = np.ascontiguousarray(v, dtype=np.float64).reshape(-1, 4)
    out = np.zeros((max(len(v), 1), 2))
    _check(lib().hp_dd_from_mf4(v if len(v) else np.zeros(4), out, len(v)), "dd_from_mf4")
    return out[:len(v)]


# ---------------------------------------------------------------- MultiFloat<N>, N = 1..8
MFN_OPS = {"add": 0, "sub": 1, "mul": 2, "div": 3, "recip": 4, "sqrt": 5, "renormalize": 6}


def _mfn(v, nl, count, name):
    v = np.ascontiguousarray(v, dtype=np.float64).reshape(-1, nl)
    if v.shape[0] != count:
        raise AssertionError(f"{name}: {v.shape[0]} values, expected {count}")
    return v if count else np.zeros((1, nl))


def mfn_gemm(nl, a, b, m, n, k, trans_a=False):
    """mf_gemm::<nl> (multifloat/linalg.rs:39-61), or mf_gemm_atb::<nl> (64-85, A stored
    k x m) when trans_a: C (m*n, nl)."""
    a, b = _mfn(a, nl, m * k, "a"), _mfn(b, nl, k * n, "b")
    c = np.zeros((max(m * n, 1), nl))
    name = "hp_mfn_gemm_atb" if trans_a else "hp_mfn_gemm"
    _check(getattr(lib(), name)(nl, a, b, c, m, n, k), name)
    return c[:m * n]


def mfn_gemv(nl, a, x, n, m):
    """mf_gemv::<nl> (multifloat/linalg.rs:17-36): y = A x, A n x m; (n, nl)."""
    a, x = _mfn(a, nl, n * m, "a"), _mfn(x, nl, m, "x")
    y = np.zeros((max(n, 1), nl))
    _check(lib().hp_mfn_gemv(nl, a, x, y, n, m), "mfn_gemv")
    return y[:n]


def mfn_dot(nl, a, b):
    """mf_dot::<nl> (multifloat/linalg.rs:7-14): (nl,)."""
    a = np.ascontiguousarray(a, dtype=np.float64).reshape(-1, nl)
    b = _mfn(b, nl, a.shape[0], "b")
    out = np.zeros(nl)
    _check(lib().hp_mfn_dot(nl, a if a.size else np.zeros(nl), b, a.shape[0], out), "mfn_dot")
    return out


def mfn_cholesky(nl, a, n):
    """mf_cholesky::<nl> (multifloat/linalg.rs:89-141): the factored copy (n*n, nl); raises
    NotPositiveDefinite (with .pivot) like the reference's Err."""
    a = _mfn(a, nl, n * n, "a").copy()
    piv = C.c_size_t(0)
    rc = lib().hp_mfn_cholesky(nl, a, n, C.byref(piv))
    if rc == HP_ENOTSPD:
        e = NotPositiveDefinite(lib().hp_last_error().decode(errors="replace"))
        e.pivot = piv.value
        raise e
    _check(rc, "mfn_cholesky")
    return a[:n * n]


def mfn_cholesky_solve(nl, l, b, n):
    """mf_cholesky_solve::<nl> (multifloat/linalg.rs:145-172): x (n, nl)."""
    l, b = _mfn(l, nl, n * n, "l"), _mfn(b, nl, n, "b")
    x = np.zeros((max(n, 1), nl))
    _check(lib().hp_mfn_cholesky_solve(nl, l, b, x, n), "mfn_cholesky_solve")
    return x[:n]


def mfn_elementwise(nl, op, a, b=None):
    """op(a_i, b_i) over (count, nl) arrays: "add", "sub", "mul", "div" (add.rs, mul.rs,
    div.rs), "recip", "sqrt", "renormalize" (b unused)."""
    a = np.ascontiguousarray(a, dtype=np.float64).reshape(-1, nl)
    cnt = a.shape[0]
    bb = _mfn(b, nl, cnt, "b") if b is not None else np.zeros((1, nl))
    out = np.zeros((max(cnt, 1), nl))
    _check(lib().hp_mfn_elementwise(nl, MFN_OPS[op], a if cnt else np.zeros(nl), bb, out, cnt), "mfn_elementwise")
    return out[:cnt]


def mfn_convert(nl_from, nl_to, v):
    """MultiFloat<nl_to> from MultiFloat<nl_from> (multifloat/conv.rs:119-157): leading limbs
    kept, the rest zero; (count, nl_to)."""
    v = np.ascontiguousarray(v, dtype=np.float64).reshape(-1, nl_from)
    out = np.zeros((max(len(v), 1), nl_to))
    _check(lib().hp_mfn_convert(nl_from, nl_to, v if len(v) else np.zeros(nl_from), out, len(v)), "mfn_convert")
    return out[:len(v)]


# ---------------------------------------------------------------- reductions
def dot(a, b):
    """dd dot product (crates/quad/src/linalg.rs:7-18), as a (2,) array."""
    a = np.ascontiguousarray(a, dtype=np.float64)
    n = a.size // 2
    b = _arr(b, 2, n, "b")
    out = np.zeros(2)
    _check(lib().hp_dot(a if n else np.zeros(2), b if n else np.zeros(2), n, out), "dot")
    return out


def gemv(a, rows, cols, x):
    """y = A x, A row-major rows x cols (linalg.rs:29-40), as (rows, 2)."""
    a = _arr(a, 2, rows * cols, "a")
    x = _arr(x, 2, cols, "x")
    y = np.zeros((max(rows, 1), 2))
    _check(lib().hp_gemv(a if a.size else np.zeros(2), rows, cols, x if x.size else np.zeros(2), y), "gemv")
    return y[:rows]


def neumaier_sum_f64(values, parallel=False):
    """Neumaier sum of f64 values (crates/compensated-sum/src/lib.rs:6-21; parallel: 43-53)."""
    v = np.ascontiguousarray(values, dtype=np.float64).ravel()
    out = np.zeros(1)
    fn = lib().hp_par_neumaier_sum_f64 if parallel else lib().hp_neumaier_sum_f64
    _check(fn(v if v.size else np.zeros(1), v.size, out), "neumaier_sum_f64")
    return float(out[0])


def neumaier_sum(values, parallel=False):
    """Neumaier sum of dd values (lib.rs:24-39; parallel: 56-66), as a (2,) array."""
    v = np.ascontiguousarray(values, dtype=np.float64)
    n = v.size // 2
    out = np.zeros(2)
    fn = lib().hp_par_neumaier_sum if parallel else lib().hp_neumaier_sum
    _check(fn(v if n else np.zeros(2), n, out), "neumaier_sum")
    return out


# ---------------------------------------------------------------- text wire formats
class ParseError(HyperprecError):
    """A literal the reference rejects (ParseF128Error / ParseMultiFloatError / serde's
    invalid_value), or one whose rounding needs more than 40 significant digits."""


def pack_strings(strings):
    """list of str/bytes -> (bytes, offsets uint64 (n+1)), the Arrow layout of the text entry points."""
    bs = [x.encode() if isinstance(x, str) else bytes(x) for x in strings]
    off = np.zeros(len(bs) + 1, dtype=np.uint64)
    off[1:] = np.cumsum([len(b) for b in bs]) if bs else []
    return b"".join(bs), off


def unpack_strings(text, offsets):
    return [text[int(offsets[i]):int(offsets[i + 1])].decode() for i in range(len(offsets) - 1)]


def _format(fn, what, v, limbs, max_len, *extra, raw=False):
    v = np.ascontiguousarray(v, dtype=np.float64)
    n = v.size // limbs
    cap = max(n * max_len, 1)
    buf = C.create_string_buffer(cap)
    off = np.zeros(n + 1, dtype=np.uint64)
    _check(fn(v if n else np.zeros(limbs), n, *extra, buf, cap, off), what)
    text = buf.raw[:int(off[n])]
    return (text, off) if raw else unpack_strings(text, off)


def format_f64(values, precision, force_sign=False, raw=False):
    """Rust's format!("{:.P$e}") / ("{:+.P$e}") of each value: list of str (raw: (bytes, offsets))."""
    if not 0 <= precision <= 64:
        raise ValueError("precision must be in 0..64")
    return _format(lib().hp_format_f64, "format_f64", values, 1, precision + 8, int(precision), int(force_sign),
                   raw=raw)


def dd_to_string(v, raw=False):
    """Display of f128 values (dd.rs:337-341): "{:.32e}" of hi + lo."""
    return _format(lib().hp_dd_to_string, "dd_to_string", v, 2, 40, raw=raw)


def mf4_to_string(v, raw=False):
    """Display of MultiFloat<4> values (multifloat/fmt.rs:7-17): "{:.64e}" of limb0 + limb1."""
    return _format(lib().hp_mf4_to_string, "mf4_to_string", v, 4, 72, raw=raw)


def dd_serialize_str(v, raw=False):
    """Human-readable serde of f128 values (dd.rs:446-450): "{:+.18e}+{:+.18e}"."""
    return _format(lib().hp_dd_serialize_str, "dd_serialize_str", v, 2, 53, raw=raw)


def mf4_serialize_str(v, raw=False):
    """Human-readable serde of MultiFloat<4> values (traits.rs:14-24): '+'-joined "{:+.14e}" limbs."""
    return _format(lib().hp_mf4_serialize_str, "mf4_serialize_str", v, 4, 91, raw=raw)


def _parse(fn, what, strings, limbs, strict):
    text, off = strings if isinstance(strings, tuple) else pack_strings(strings)
    n = len(off) - 1
    out = np.zeros((max(n, 1), limbs))
    status = np.zeros(max(n, 1), dtype=np.uint8)
    rc = fn(text if text else b"\0", off, n, out, status)
    if rc == HP_EINVAL and not strict:
        pass
    elif rc == HP_EINVAL:
        raise ParseError(f"{what}: " + lib().hp_last_error().decode(errors="replace"))
    else:
        _check(rc, what)
    out = out[:n, 0] if limbs == 1 else out[:n]
    return (out, status[:n]) if not strict else out


def parse_f64(strings, strict=True):
    """Rust's str::parse::<f64> of each string.  strict: raise ParseError on any rejected
    element; else return (values, status) with status 0 ok / 1 rejected / 2 undecided."""
    return _parse(lib().hp_parse_f64, "parse_f64", strings, 1, strict)


def dd_from_str(strings, strict=True):
    """FromStr of f128 (dd.rs:371-435): (n, 2)."""
    return _parse(lib().hp_dd_from_str, "dd_from_str", strings, 2, strict)


def mf4_from_str(strings, strict=True):
    """FromStr of MultiFloat<4> (multifloat/fmt.rs:32-114): (n, 4)."""
    return _parse(lib().hp_mf4_from_str, "mf4_from_str", strings, 4, strict)


def dd_deserialize_str(strings, strict=True):
    """serde string form of f128 (dd.rs:496-511): (n, 2)."""
    return _parse(lib().hp_dd_deserialize_str, "dd_deserialize_str", strings, 2, strict)


def mf4_deserialize_str(strings, strict=True):
    """serde string form of MultiFloat<4> (traits.rs:104-148): (n, 4)."""
    return _parse(lib().hp_mf4_deserialize_str, "mf4_deserialize_str", strings, 4, strict)


def overlap_s(centers, alpha):
    """Normalised s-type Gaussian overlap matrix in dd, (n*n, 2)."""
    alpha = np.ascontiguousarray(alpha, dtype=np.float64)
    n = alpha.size
    centers = _arr(centers, 3, n, "centers")
    s = np.zeros((n * n, 2))
    _check(lib().hp_overlap_s(centers, alpha, n, s), "overlap_s")
    return s


def overlap_adjoint(centers, alpha, b, factor_kind=HP_FACTOR_DD, gemm_mode=HP_GEMM_PLAIN):
    """Gradients of log det S and b^T S^-1 b (S = overlap_s(centers, alpha), b dd (n, 2))
    over the 4n parameters (3n centre coordinates, Gaussian-major, then n exponents).
    Returns (g_logdet (4n, 2), g_quad (4n, 2), AdjointInfo); raises SolveError if S is not
    positive definite."""
    alpha = np.ascontiguousarray(alpha, dtype=np.float64)
    n = alpha.size
    centers = _arr(centers, 3, n, "centers")
    b = _arr(b, 2, n, "b")
    gl = np.zeros((4 * n, 2))
    gq = np.zeros((4 * n, 2))
    info = AdjointInfo()
    rc = lib().hp_overlap_adjoint(centers, alpha, b, n, factor_kind, gemm_mode, C.byref(info), gl, gq)
    if rc == HP_ENOTSPD:
        raise SolveError(info.pivot)
    _check(rc, "overlap_adjoint")
    return gl, gq, info


def gen_overlap_basis(seed, n, alpha_lo=0.01, alpha_hi=10.0, spacing=1.4, jitter=0.05):
    """Lattice Gaussians: first n points of an L^3 grid (L = ceil(n^(1/3))), spacing in
    bohr plus U(+-jitter), exponents log-uniform in [alpha_lo, alpha_hi]."""
    c = np.empty((n, 3))
    a = np.empty(n)
    lib().hp_gen_overlap_basis(seed, n, alpha_lo, alpha_hi, spacing, jitter, c, a)
    return c, a


# ---------------------------------------------------------------- measurement
def set_phase_timing(on):
    """0 / False: off; 1 / True: each split GEMM waits for its phase events (phase_times);
    2: calls accumulate without waiting (collect_phase_times)."""
    lib().hp_set_phase_timing(2 if on == 2 else int(bool(on)))


def split_gemm_variant(m, n, k, slices=0):
    """Int8 kernel the split GEMM autotuner chose for this shape class (0 = untuned)."""
    return lib().hp_split_gemm_variant(m, n, k, slices)


def split_slices_for_bits(k, bits):
    """The fewest split moduli whose integer width reaches bits + 1 at inner dimension k."""
    return lib().hp_split_slices_for_bits(k, bits)


def gemm_plan(m, n, k, bits=0, trans_a=False):
    """The GemmPlan HP_GEMM_AUTO chose for this shape class and accuracy, or None if untuned."""
    p = GemmPlan()
    return p if lib().hp_gemm_plan_get(int(trans_a), m, n, k, bits, C.byref(p)) else None


def split_gemm_set_variant(m, n, k, variant, slices=0):
    """Pin the split GEMM's int8 kernel for a shape class (m, n, k and the moduli count, slices
    = 0 meaning the dd-accuracy default the GEMM itself takes): 32, 64 or 128 (K stages of that
    depth staged through registers), 3064 (direct-to-LDS ring of 64-deep stages), 4064 / 5064
    (that ring software-pipelined, 8 / 4 wavefronts), 6064 (register-staged, three stages
    ahead) or 7064 (4064 as a persistent kernel); 0 re-tunes."""
    _check(lib().hp_split_gemm_set_variant(m, n, k, slices, variant), "split_gemm_set_variant")


def phase_times():
    """Milliseconds of each phase of the last multi-kernel call on this thread."""
    buf = np.zeros(8)
    n = lib().hp_get_phase_times(buf, 8)
    return list(buf[:n])


def collect_phase_times():
    """(per-phase milliseconds summed over the calls since the last collect, number of
    calls) in accumulate mode (set_phase_timing(2)); waits for the last call's events."""
    buf = np.zeros(8)
    calls = C.c_int(0)
    n = lib().hp_collect_phase_times(buf, 8, C.byref(calls))
    _check(n if n < 0 else 0, "collect_phase_times")
    return list(buf[:n]), calls.value


# ---------------------------------------------------------------- generators
def gen_dd_normal(seed, count):
    out = np.empty((count, 2))
    lib().hp_gen_dd_normal(seed, count, out)
    return out


def gen_mf4_normal(seed, count):
    out = np.empty((count, 4))
    lib().hp_gen_mf4_normal(seed, count, out)
    return out


def gen_mfn_normal(nl, seed, count):
    """(count, nl) MultiFloat<nl> values: limb 0 N(0,1), each further limb a uniform fraction
    of half an ulp of the one above (non-overlapping, every limb populated)."""
    rng = np.random.default_rng(seed)
    out = np.empty((count, nl))
    out[:, 0] = rng.standard_normal(count)
    for i in range(1, nl):
        out[:, i] = out[:, i - 1] * 2.0 ** -53 * rng.uniform(-0.5, 0.5, count)
    return out


from . import dev  # noqa: E402,F401  (device entry points on torch tensors)
