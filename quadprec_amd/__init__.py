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

HP_OK, HP_ESHAPE, HP_ENOTSPD, HP_EDEVICE, HP_EINVAL, HP_ENOMEM = range(6)
HP_GEMM_PLAIN = 0

_lib = None
_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_sz = C.c_size_t
_vp = C.c_void_p


class HyperprecError(RuntimeError):
    pass


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


# ---------------------------------------------------------------- GEMM (linalg.rs:150-182)
def gemm(a, b, m, n, k, out=None):
    """C = A*B in double-double (crates/quad/src/linalg.rs:150-165). Returns C (m*n, 2)."""
    a = _arr(a, 2, m * k, "a")
    b = _arr(b, 2, k * n, "b")
    c = np.zeros((m * n, 2)) if out is None else out
    _check(lib().hp_gemm(a, b, c, m, n, k), "gemm")
    return c


def gemm_atb(a, b, m, n, k, out=None):
    """C = A^T*B with A stored k x m (linalg.rs:168-182)."""
    a = _arr(a, 2, k * m, "a")
    b = _arr(b, 2, k * n, "b")
    c = np.zeros((m * n, 2)) if out is None else out
    _check(lib().hp_gemm_atb(a, b, c, m, n, k), "gemm_atb")
    return c


def mf_gemm(a, b, m, n, k, out=None):
    """MultiFloat<4> C = A*B (crates/quad/src/multifloat/linalg.rs:39-61)."""
    a = _arr(a, 4, m * k, "a")
    b = _arr(b, 4, k * n, "b")
    c = np.zeros((m * n, 4)) if out is None else out
    _check(lib().hp_mf_gemm(a, b, c, m, n, k), "mf_gemm")
    return c


def mf_gemm_atb(a, b, m, n, k, out=None):
    """MultiFloat<4> C = A^T*B, A stored k x m (multifloat/linalg.rs:64-85)."""
    a = _arr(a, 4, k * m, "a")
    b = _arr(b, 4, k * n, "b")
    c = np.zeros((m * n, 4)) if out is None else out
    _check(lib().hp_mf_gemm_atb(a, b, c, m, n, k), "mf_gemm_atb")
    return c


# ---------------------------------------------------------------- generators
def gen_dd_normal(seed, count):
    out = np.empty((count, 2))
    lib().hp_gen_dd_normal(seed, count, out)
    return out


def gen_mf4_normal(seed, count):
    out = np.empty((count, 4))
    lib().hp_gen_mf4_normal(seed, count, out)
    return out


from . import dev  # noqa: E402,F401  (device entry points on torch tensors)
