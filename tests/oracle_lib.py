"""ctypes bindings of the oracle (oracle/libhpo.so) for the test suite and bench's
CPU-baseline leg.  The oracle is the checker: a CPU restatement, op for op, of the
hyperprec reference algorithms.  Product code never imports this module.

Array conventions (numpy, float64):
  dd arrays  : shape (..., 2)  [hi, lo]         (Rust &[f128] memory layout)
  MF4 arrays : shape (..., 4)  limbs, most significant first
"""
import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
LIB_PATH = os.path.join(ORACLE_DIR, "libhpo.so")

_lib = None
_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_sz = C.c_size_t


class DD(C.Structure):
    _fields_ = [("hi", C.c_double), ("lo", C.c_double)]


class MF4(C.Structure):
    _fields_ = [("l", C.c_double * 4)]


class Solution(C.Structure):
    _fields_ = [("kappa", C.c_double), ("residual_norm", C.c_double), ("strategy", C.c_int),
                ("refinement_iters", C.c_int), ("error", C.c_int), ("pivot", C.c_size_t)]


def build():
    subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        for name in ("hpo_dd_add", "hpo_dd_sub", "hpo_dd_mul", "hpo_dd_div"):
            getattr(L, name).restype = DD
            getattr(L, name).argtypes = [DD, DD]
        for name in ("hpo_dd_add_f64", "hpo_dd_mul_f64", "hpo_dd_div_f64"):
            getattr(L, name).restype = DD
            getattr(L, name).argtypes = [DD, C.c_double]
        for name in ("hpo_dd_sqrt", "hpo_dd_recip", "hpo_dd_exp", "hpo_dd_ln"):
            getattr(L, name).restype = DD
            getattr(L, name).argtypes = [DD]
        L.hpo_dd_binop_n.argtypes = [C.c_int, _dp, _dp, _dp, _sz]
        L.hpo_dd_unop_n.argtypes = [C.c_int, _dp, _dp, _sz]
        L.hpo_mf4_binop_n.argtypes = [C.c_int, _dp, _dp, _dp, _sz]
        L.hpo_mf4_unop_n.argtypes = [C.c_int, _dp, _dp, _sz]
        L.hpo_mf4_renormalize.argtypes = [_dp]
        L.hpo_mf4_renormalize_from8.argtypes = [_dp, _dp]
        for name in ("hpo_gemm", "hpo_gemm_atb", "hpo_mf4_gemm", "hpo_mf4_gemm_atb"):
            getattr(L, name).argtypes = [_dp, _dp, _dp, _sz, _sz, _sz]
        for name in ("hpo_gemm_rows", "hpo_gemm_atb_rows", "hpo_mf4_gemm_rows"):
            getattr(L, name).argtypes = [_dp, _dp, _dp, _sz, _sz, _sz, _sz, _sz]
        L.hpo_cholesky.argtypes = [_dp, _sz, C.POINTER(C.c_size_t), C.POINTER(C.c_double)]
        L.hpo_cholesky_blocked.argtypes = [_dp, _sz, _sz, C.POINTER(C.c_size_t), C.POINTER(C.c_double)]
        for name in ("hpo_forward_solve", "hpo_backward_solve", "hpo_solve_cholesky", "hpo_matvec",
                     "hpo_mf4_cholesky_solve"):
            getattr(L, name).argtypes = [_dp, _dp, _dp, _sz]
        L.hpo_cond_estimate.argtypes = [_dp, _sz]
        L.hpo_cond_estimate.restype = C.c_double
        L.hpo_mf4_cholesky.argtypes = [_dp, _sz]
        L.hpo_mf4_gemv.argtypes = [_dp, _dp, _dp, _sz, _sz]
        for name in ("hpo_solve_spd", "hpo_solve_spd_f128"):
            getattr(L, name).argtypes = [_dp, _dp, _sz, C.POINTER(Solution), _dp, _dp]
        L.hpo_norm_f128.argtypes = [_dp, _sz]
        L.hpo_norm_f128.restype = C.c_double
        L.hpo_overlap_primitive_f64.argtypes = [C.c_double, _dp, C.c_double, _dp]
        L.hpo_overlap_primitive_f64.restype = C.c_double
        L.hpo_overlap_primitive_f128.argtypes = [C.c_double, _dp, C.c_double, _dp]
        L.hpo_overlap_primitive_f128.restype = DD
        L.hpo_screen_s_pair.argtypes = [C.c_double, _dp, C.c_double, _dp] + [C.c_double] * 4
        L.hpo_screen_s_pair.restype = C.c_int
        L.hpo_assemble_overlap_s.argtypes = [_dp, _dp, _dp, _sz, _dp, _dp,
                                             np.ctypeslib.ndpointer(dtype=np.uintp, flags="C_CONTIGUOUS")]
        L.hpo_overlap_s_dd_entry.argtypes = [C.c_double, _dp, C.c_double, _dp]
        L.hpo_overlap_s_dd_entry.restype = DD
        L.hpo_overlap_s_dd.argtypes = [_dp, _dp, _sz, _dp]
        L.hpo_overlap_adjoint_dd.argtypes = [_dp, _dp, _dp, _sz, _dp, _dp, _dp, _dp, C.c_void_p]
        L.hpo_overlap_adjoint_dd.restype = C.c_int
        L.hpo_overlap_adjoint_dd_planted.argtypes = [_dp, _dp, _dp, _sz, _dp, _dp, _dp, _dp, C.c_void_p, C.c_int]
        L.hpo_overlap_adjoint_dd_planted.restype = C.c_int
        L.hpo_overlap_fd_f256.argtypes = [_dp, _dp, _dp, _sz,
                                          np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS"),
                                          _sz, C.c_double, _dp, _dp]
        L.hpo_overlap_fd_f256.restype = C.c_int
        L.hpo_count_dd_mac.restype = C.c_long
        L.hpo_count_mf4_mac.argtypes = [_dp, _dp, _dp]
        L.hpo_count_mf4_mac.restype = C.c_long
        L.hpo_count_overlap_entry.argtypes = [C.c_double, _dp, C.c_double, _dp]
        L.hpo_count_overlap_entry.restype = C.c_long
        L.hpo_num_threads.restype = C.c_int
        L.hpo_assemble_overlap_mixed.argtypes = [_dp, np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS"),
                                                 np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS"),
                                                 _dp, _dp, _sz, _dp, _dp, C.POINTER(C.c_size_t)]
        L.hpo_assemble_overlap_mixed.restype = C.c_size_t
        L.hpo_mfn_elementwise.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp, _sz]
        L.hpo_mfn_gemm_rows.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp, _sz, _sz, _sz, _sz, _sz]
        L.hpo_mfn_gemv.argtypes = [C.c_int, _dp, _dp, _dp, _sz, _sz]
        L.hpo_mfn_cholesky.argtypes = [C.c_int, _dp, _sz]
        L.hpo_mfn_cholesky.restype = C.c_long
        L.hpo_mfn_cholesky_solve.argtypes = [C.c_int, _dp, _dp, _dp, _sz]
        L.hpo_count_mfn_mac.argtypes = [C.c_int, _dp, _dp, _dp]
        L.hpo_count_mfn_mac.restype = C.c_long
        L.hpo_assemble_overlap_envelope.argtypes = L.hpo_assemble_overlap_mixed.argtypes[:7] + [_dp]
        L.hpo_assemble_overlap_envelope.restype = None
        L.hpo_set_exp_ulp_shift.argtypes = [C.c_int]
        L.hpo_dot.argtypes = [_dp, _dp, _sz]
        L.hpo_dot.restype = DD
        L.hpo_gemv.argtypes = [_dp, _sz, _sz, _dp, _dp]
        for name in ("hpo_neumaier_sum_f64", "hpo_par_neumaier_sum_f64"):
            getattr(L, name).argtypes = [_dp, _sz]
            getattr(L, name).restype = C.c_double
        for name in ("hpo_neumaier_sum", "hpo_par_neumaier_sum"):
            getattr(L, name).argtypes = [_dp, _sz]
            getattr(L, name).restype = DD
        L.hpo_jacobi_eigen.argtypes = [_dp, _sz, _sz, _dp, _dp]
        L.hpo_jacobi_eigen.restype = C.c_int
        L.hpo_jacobi_eigen_planted.argtypes = [_dp, _sz, _sz, _dp, _dp, C.c_int]
        L.hpo_jacobi_eigen_planted.restype = C.c_int
        L.hpo_jacobi_eigen_maxpivot.argtypes = [_dp, _sz, _dp, _dp]
        L.hpo_canonical_orthogonalize.argtypes = [_dp, _sz, C.c_double, _dp]
        L.hpo_canonical_orthogonalize.restype = C.c_size_t
        L.hpo_solve_gen_eigen.argtypes = [_dp, _dp, _sz, C.c_double, _dp, _dp]
        L.hpo_solve_gen_eigen.restype = C.c_size_t
        L.hpo_format_exp.argtypes = [C.c_double, C.c_int, C.c_int, C.c_char_p]
        for name in ("hpo_dd_to_string", "hpo_mf4_to_string", "hpo_dd_serialize_str", "hpo_mf4_serialize_str"):
            getattr(L, name).argtypes = [_dp, C.c_char_p]
        L.hpo_parse_f64.argtypes = [C.c_char_p, _sz, C.POINTER(C.c_double)]
        for name in ("hpo_dd_from_str", "hpo_mf4_from_str", "hpo_dd_deserialize_str", "hpo_mf4_deserialize_str"):
            getattr(L, name).argtypes = [C.c_char_p, _sz, _dp]
        _op = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")
        L.hpo_dd_serialize_n.argtypes = [_dp, _sz, C.c_char_p, _op]
        L.hpo_dd_serialize_n.restype = _sz
        L.hpo_dd_deserialize_n.argtypes = [C.c_char_p, _op, _sz, _dp]
        L.hpo_dd_deserialize_n.restype = _sz
        _lib = L
    return _lib


def _c(x):
    return np.ascontiguousarray(x, dtype=np.float64)


# ---------------------------------------------------------------- dd scalar / array ops
_BIN = {"add": 0, "sub": 1, "mul": 2, "div": 3, "add_f64": 4, "mul_f64": 5, "div_f64": 6}
_UN = {"sqrt": 0, "recip": 1, "exp": 2, "ln": 3, "neg": 4, "abs": 5}


def dd_binop(op, a, b):
    a, b = _c(a), _c(b)
    a, b = np.broadcast_arrays(a, b)
    a, b = _c(a), _c(b)
    out = np.empty_like(a)
    lib().hpo_dd_binop_n(_BIN[op], a, b, out, a.size // 2)
    return out


def dd_unop(op, a):
    a = _c(a)
    out = np.empty_like(a)
    lib().hpo_dd_unop_n(_UN[op], a, out, a.size // 2)
    return out


_MBIN = {"add": 0, "sub": 1, "mul": 2, "div": 3, "mul_f64": 4}
_MUN = {"sqrt": 0, "recip": 1, "exp": 2, "renormalize": 3}


def mf4_binop(op, a, b):
    a, b = np.broadcast_arrays(_c(a), _c(b))
    a, b = _c(a), _c(b)
    out = np.empty_like(a)
    lib().hpo_mf4_binop_n(_MBIN[op], a, b, out, a.size // 4)
    return out


def mf4_unop(op, a):
    a = _c(a)
    out = np.empty_like(a)
    lib().hpo_mf4_unop_n(_MUN[op], a, out, a.size // 4)
    return out


# ------------------------------------------------- MultiFloat<N>, N = 1..8 (hpo_mfn.cpp)
_MFN_OP = {"add": 0, "sub": 1, "mul": 2, "div": 3, "recip": 4, "sqrt": 5, "renormalize": 6}


def mfn_op(nl, op, a, b=None):
    """Elementwise MultiFloat<nl> op on (count, nl) arrays."""
    a = _c(np.asarray(a, dtype=np.float64).reshape(-1, nl))
    b = a if b is None else _c(np.broadcast_to(np.asarray(b, dtype=np.float64).reshape(-1, nl), a.shape))
    out = np.empty_like(a)
    assert lib().hpo_mfn_elementwise(nl, _MFN_OP[op], a, b, out, a.shape[0]) == 0
    return out


def mfn_gemm(nl, a, b, m, n, k, trans_a=False, row0=0, nrows=None):
    """mf_gemm::<nl> (mf_gemm_atb when trans_a: A stored k x m); rows row0 .. row0+nrows-1."""
    c = np.zeros((m * n, nl))
    nrows = m - row0 if nrows is None else nrows
    assert lib().hpo_mfn_gemm_rows(nl, int(trans_a), _c(a), _c(b), c, m, n, k, row0, nrows) == 0
    return c


def mfn_gemv(nl, a, x, n, m):
    y = np.zeros((n, nl))
    assert lib().hpo_mfn_gemv(nl, _c(a), _c(x), y, n, m) == 0
    return y


def mfn_dot(nl, a, b):
    """mf_dot (linalg.rs:7-14): the one-row gemv."""
    return mfn_gemv(nl, a, b, 1, len(np.asarray(a).reshape(-1, nl)))[0]


def count_mfn_mac(nl, s, a, b):
    """FP64 operations of one MultiFloat<nl> MAC s + a * b on these values."""
    return lib().hpo_count_mfn_mac(nl, _c(s), _c(a), _c(b))


def mfn_cholesky(nl, a, n):
    """mf_cholesky::<nl> in place on a copy; raises ValueError at a non-positive pivot."""
    a = _c(a).copy()
    r = lib().hpo_mfn_cholesky(nl, a, n)
    if r:
        raise ValueError(f"Cholesky failed: matrix is not positive definite (column {r - 1})")
    return a


def mfn_cholesky_solve(nl, l, b, n):
    x = np.zeros((n, nl))
    assert lib().hpo_mfn_cholesky_solve(nl, _c(l), _c(b), x, n) == 0
    return x


# ---------------------------------------------------------------- linalg
def gemm(a, b, m, n, k):
    c = np.zeros((m * n, 2))
    lib().hpo_gemm(_c(a), _c(b), c, m, n, k)
    return c


def gemm_rows(a, b, m, n, k, row0, nrows):
    c = np.zeros((m * n, 2))
    lib().hpo_gemm_rows(_c(a), _c(b), c, m, n, k, row0, nrows)
    return c


def gemm_atb(a, b, m, n, k):
    c = np.zeros((m * n, 2))
    lib().hpo_gemm_atb(_c(a), _c(b), c, m, n, k)
    return c


def gemm_atb_rows(a, b, m, n, k, row0, nrows):
    c = np.zeros((m * n, 2))
    lib().hpo_gemm_atb_rows(_c(a), _c(b), c, m, n, k, row0, nrows)
    return c


def mf4_gemm(a, b, m, n, k):
    c = np.zeros((m * n, 4))
    lib().hpo_mf4_gemm(_c(a), _c(b), c, m, n, k)
    return c


def mf4_gemm_rows(a, b, m, n, k, row0, nrows):
    c = np.zeros((m * n, 4))
    lib().hpo_mf4_gemm_rows(_c(a), _c(b), c, m, n, k, row0, nrows)
    return c


def mf4_gemm_atb(a, b, m, n, k):
    c = np.zeros((m * n, 4))
    lib().hpo_mf4_gemm_atb(_c(a), _c(b), c, m, n, k)
    return c


class CholeskyError(Exception):
    def __init__(self, pivot, value):
        super().__init__(f"Cholesky failed at pivot {pivot} (value={value:.4e})")
        self.pivot, self.value = pivot, value


def cholesky(a, n):
    """Returns the factored copy; raises CholeskyError like linalg.rs:54-57."""
    a = _c(a).copy()
    pv, vv = C.c_size_t(0), C.c_double(0)
    if lib().hpo_cholesky(a, n, C.byref(pv), C.byref(vv)):
        raise CholeskyError(pv.value, vv.value)
    return a


def cholesky_blocked(a, n, nb):
    a = _c(a).copy()
    pv, vv = C.c_size_t(0), C.c_double(0)
    if lib().hpo_cholesky_blocked(a, n, nb, C.byref(pv), C.byref(vv)):
        raise CholeskyError(pv.value, vv.value)
    return a


def _vec_op(name, l, b, n):
    out = np.zeros((n, 2))
    getattr(lib(), name)(_c(l), _c(b), out, n)
    return out


def forward_solve(l, b, n):
    return _vec_op("hpo_forward_solve", l, b, n)


def backward_solve(l, y, n):
    return _vec_op("hpo_backward_solve", l, y, n)


def solve_cholesky(l, b, n):
    return _vec_op("hpo_solve_cholesky", l, b, n)


def matvec(a, x, n):
    return _vec_op("hpo_matvec", a, x, n)


def cond_estimate(l, n):
    return lib().hpo_cond_estimate(_c(l), n)


def mf4_cholesky(a, n):
    a = _c(a).copy()
    if lib().hpo_mf4_cholesky(a, n):
        raise ValueError("Cholesky failed: matrix is not positive definite")
    return a


def mf4_cholesky_solve(l, b, n):
    out = np.zeros((n, 4))
    lib().hpo_mf4_cholesky_solve(_c(l), _c(b), out, n)
    return out


def mf4_gemv(a, x, n, m):
    out = np.zeros((n, 4))
    lib().hpo_mf4_gemv(_c(a), _c(x), out, n, m)
    return out


def _solve(name, a, b, n):
    sol = Solution()
    x = np.zeros((n, 2))
    xf = np.zeros(n)
    getattr(lib(), name)(_c(a), _c(b), n, C.byref(sol), x, xf)
    return sol, x, xf


def solve_spd(a, b, n):
    return _solve("hpo_solve_spd", a, b, n)


def solve_spd_f128(a, b, n):
    return _solve("hpo_solve_spd_f128", a, b, n)


def norm_f128(x):
    x = _c(x)
    return lib().hpo_norm_f128(x, x.size // 2)


# ---------------------------------------------------------------- overlap
def overlap_s_dd(centers, alpha):
    n = len(alpha)
    s = np.zeros((n * n, 2))
    lib().hpo_overlap_s_dd(_c(centers), _c(alpha), n, s)
    return s


def overlap_s_dd_entry(ai, ci, aj, cj):
    r = lib().hpo_overlap_s_dd_entry(ai, _c(ci), aj, _c(cj))
    return np.array([r.hi, r.lo])


def overlap_primitive_f64(ai, ci, aj, cj):
    return lib().hpo_overlap_primitive_f64(ai, _c(ci), aj, _c(cj))


def overlap_primitive_f128(ai, ci, aj, cj):
    r = lib().hpo_overlap_primitive_f128(ai, _c(ci), aj, _c(cj))
    return np.array([r.hi, r.lo])


def assemble_overlap_s(centers, alpha, coeff, params=(0.1, 500.0, 1e-14, 12.0)):
    n = len(alpha)
    s = np.zeros(n * n)
    counts = np.zeros(3, dtype=np.uintp)
    lib().hpo_assemble_overlap_s(_c(centers), _c(alpha), _c(coeff), n, _c(params), s, counts)
    return s, counts


def overlap_adjoint_dd(centers, alpha, b, mistake=0):
    """Reference gradient of (log det S, b^T S^-1 b) from dd ops: returns
    (logdet (2,), quad (2,), g_ld (4n, 2), g_q (4n, 2), x (n, 2)); parameters are the 3n
    centre coordinates (Gaussian-major) then the n exponents.  mistake > 0 plants a known
    error (see hpo.h) for tests of the checks."""
    n = len(alpha)
    ld, q = np.zeros(2), np.zeros(2)
    gl, gq, x = np.zeros((4 * n, 2)), np.zeros((4 * n, 2)), np.zeros((n, 2))
    rc = lib().hpo_overlap_adjoint_dd_planted(_c(centers), _c(alpha), _c(b), n, ld, q, gl, gq, x.ctypes.data,
                                              mistake)
    if rc:
        raise CholeskyError(-1, 0.0)
    return ld, q, gl, gq, x


def overlap_fd_f256(centers, alpha, b, coords, h=1e-15):
    """f256 central differences at steps h and 2h: arrays (ncoords, 2, 4) for log det S and
    for b^T S^-1 b."""
    n = len(alpha)
    coords = np.ascontiguousarray(coords, dtype=np.int64)
    fl = np.zeros((coords.size, 2, 4))
    fq = np.zeros((coords.size, 2, 4))
    rc = lib().hpo_overlap_fd_f256(_c(centers), _c(alpha), _c(b), n, coords, coords.size, h, fl, fq)
    if rc:
        raise CholeskyError(-1, 0.0)
    return fl, fq


def count_dd_mac():
    return lib().hpo_count_dd_mac()


def count_mf4_mac(s, a, b):
    return lib().hpo_count_mf4_mac(_c(s), _c(a), _c(b))


def count_overlap_entry(ai, ci, aj, cj):
    return lib().hpo_count_overlap_entry(ai, _c(ci), aj, _c(cj))


# ---------------------------------------------------------------- eigensolvers
class NotConverged(Exception):
    pass


def jacobi_eigen(a, n, max_iter, mistake=0):
    """linalg.rs:299-362: (evals (n, 2) in diagonal order, evecs (n*n, 2), row i = eigenvector i)."""
    ev = np.zeros((n, 2))
    vec = np.zeros((n * n, 2))
    if lib().hpo_jacobi_eigen_planted(_c(a), n, max_iter, ev, vec, mistake):
        raise NotConverged("Jacobi eigenvalue solver did not converge within max iterations")
    return ev, vec


def jacobi_eigen_maxpivot(a, n):
    """ortho.rs:61-152: (evals (n, 2) in diagonal order, evecs (n*n, 2), column j = eigenvector j)."""
    ev = np.zeros((max(n, 1), 2))
    vec = np.zeros((max(n * n, 1), 2))
    lib().hpo_jacobi_eigen_maxpivot(_c(a), n, ev, vec)
    return ev[:n], vec[:n * n]


def canonical_orthogonalize(s, n, threshold):
    """ortho.rs:158-193: X (n * n_orth, 2) row-major, and n_orth."""
    x = np.zeros((max(n * n, 1), 2))
    m = lib().hpo_canonical_orthogonalize(_c(s), n, threshold, x)
    return x[:n * m], m


def solve_gen_eigen(f, s, n, threshold):
    """ortho.rs:201-248: (eigenvalues (m, 2) ascending, C (n * m, 2), m)."""
    ev = np.zeros((max(n, 1), 2))
    c = np.zeros((max(n * n, 1), 2))
    m = lib().hpo_solve_gen_eigen(_c(f), _c(s), n, threshold, ev, c)
    return ev[:m], c[:n * m], m


# ---------------------------------------------------------------- contracted overlap
DEFAULT_SCREEN = (0.1, 500.0, 1e-14, 12.0)  # ScreenParams::default() (screen.rs:38-46)


def flatten_basis(shells):
    """shells: list of (center (3,), l, [(alpha, coeff), ...]) -> flat arrays."""
    centers = np.array([c for c, _, _ in shells], dtype=np.float64).ravel()
    ls = np.array([l for _, l, _ in shells], dtype=np.int32)
    off = np.zeros(len(shells) + 1, dtype=np.int64)
    for i, (_, _, prims) in enumerate(shells):
        off[i + 1] = off[i] + len(prims)
    alpha = np.array([a for _, _, prims in shells for a, _ in prims], dtype=np.float64)
    coeff = np.array([c for _, _, prims in shells for _, c in prims], dtype=np.float64)
    return centers, ls, off, alpha, coeff


def assemble_overlap_mixed(shells, params=DEFAULT_SCREEN):
    """assemble.rs:35-85: (S (n_ao, n_ao) f64, (f64_pairs, f128_pairs, skipped_pairs))."""
    centers, ls, off, alpha, coeff = flatten_basis(shells)
    n_ao = int(sum((l + 1) * (l + 2) // 2 for l in ls))
    s = np.zeros(max(n_ao * n_ao, 1))
    counts = (C.c_size_t * 3)()
    lib().hpo_assemble_overlap_mixed(centers, ls, off, alpha, coeff, len(shells), _c(params), s, counts)
    return s[:n_ao * n_ao].reshape(n_ao, n_ao), tuple(counts)


def assemble_overlap_envelope(shells, params=DEFAULT_SCREEN):
    """Per entry, the sum of |term| over the primitive pairs whose signed sum is S_uv."""
    centers, ls, off, alpha, coeff = flatten_basis(shells)
    n_ao = int(sum((l + 1) * (l + 2) // 2 for l in ls))
    env = np.zeros(max(n_ao * n_ao, 1))
    lib().hpo_assemble_overlap_envelope(centers, ls, off, alpha, coeff, len(shells), _c(params), env)
    return env[:n_ao * n_ao].reshape(n_ao, n_ao)


def set_exp_ulp_shift(ulps):
    """Moves the overlap integrals' exp by up to +-ulps (deterministic); 0 is libm's exp."""
    lib().hpo_set_exp_ulp_shift(int(ulps))


# ---------------------------------------------------------------- reductions
def dot(a, b):
    r = lib().hpo_dot(_c(a), _c(b), len(a))
    return np.array([r.hi, r.lo])


def gemv(a, rows, cols, x):
    y = np.zeros((rows, 2))
    lib().hpo_gemv(_c(a), rows, cols, _c(x), y)
    return y


def neumaier_sum_f64(v, parallel=False):
    v = _c(v)
    return (lib().hpo_par_neumaier_sum_f64 if parallel else lib().hpo_neumaier_sum_f64)(v, v.size)


def neumaier_sum(v, parallel=False):
    v = _c(v)
    r = (lib().hpo_par_neumaier_sum if parallel else lib().hpo_neumaier_sum)(v, v.size // 2)
    return np.array([r.hi, r.lo])


# ---------------------------------------------------------------- text wire formats
def _bytes(t):
    return t.encode() if isinstance(t, str) else bytes(t)


def format_exp(x, prec, plus=False):
    """Rust's format!("{:.prec$e}") (or "{:+.prec$e}") of one f64."""
    buf = C.create_string_buffer(prec + 32)
    n = lib().hpo_format_exp(float(x), prec, int(plus), buf)
    return buf.raw[:n].decode()


def _fmt_value(name, v, limbs):
    buf = C.create_string_buffer(512)
    n = getattr(lib(), name)(np.ascontiguousarray(v, dtype=np.float64).reshape(limbs), buf)
    return buf.raw[:n].decode()


def dd_to_string(v):
    return _fmt_value("hpo_dd_to_string", v, 2)


def mf4_to_string(v):
    return _fmt_value("hpo_mf4_to_string", v, 4)


def dd_serialize_str(v):
    return _fmt_value("hpo_dd_serialize_str", v, 2)


def mf4_serialize_str(v):
    return _fmt_value("hpo_mf4_serialize_str", v, 4)


def parse_f64(text):
    """Rust's str::parse::<f64>: the value, or None when the literal is rejected."""
    b = _bytes(text)
    out = C.c_double(0.0)
    return None if lib().hpo_parse_f64(b, len(b), C.byref(out)) else out.value


def _parse_value(name, text, limbs):
    b = _bytes(text)
    out = np.zeros(limbs)
    return None if getattr(lib(), name)(b, len(b), out) else out


def dd_from_str(text):
    return _parse_value("hpo_dd_from_str", text, 2)


def mf4_from_str(text):
    return _parse_value("hpo_mf4_from_str", text, 4)


def dd_deserialize_str(text):
    return _parse_value("hpo_dd_deserialize_str", text, 2)


def mf4_deserialize_str(text):
    return _parse_value("hpo_mf4_deserialize_str", text, 4)


def dd_serialize_n(v):
    """serde strings of n dd values in one C loop: (bytes, offsets)."""
    v = _c(v)
    n = v.size // 2
    buf = C.create_string_buffer(max(53 * n, 1))
    off = np.zeros(n + 1, dtype=np.uint64)
    total = lib().hpo_dd_serialize_n(v, n, buf, off)
    return buf.raw[:total], off


def dd_deserialize_n(text, off):
    n = len(off) - 1
    out = np.zeros((max(n, 1), 2))
    bad = lib().hpo_dd_deserialize_n(text, np.ascontiguousarray(off, dtype=np.uint64), n, out)
    return out[:n], bad
