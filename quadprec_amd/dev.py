"""Device entry points: operands resident in HBM as torch float64 tensors laid out as
structure-of-arrays planes (dd: hi and lo planes; MF4: 4 limb planes stacked on dim 0).

Kernels are enqueued on torch's current HIP stream.  torch is used only for device
memory and streams.
"""
import numpy as np

from . import HP_GEMM_PLAIN, _check, lib


def _ptr(t):
    return t.data_ptr()


def _stream():
    import torch
    return torch.cuda.current_stream().cuda_stream


def _need(t, numel, name):
    import torch
    if not isinstance(t, torch.Tensor) or t.dtype != torch.float64 or not t.is_cuda or not t.is_contiguous():
        raise TypeError(f"{name}: expected a contiguous float64 CUDA/HIP tensor")
    if t.numel() != numel:
        raise ValueError(f"{name}: expected {numel} elements, got {t.numel()}")


def to_planes(x_aos, limbs, device="cuda"):
    """Host AoS numpy array (count, limbs) -> device tensor (limbs, count)."""
    import torch
    a = np.ascontiguousarray(np.asarray(x_aos, dtype=np.float64).reshape(-1, limbs).T)
    return torch.from_numpy(a).to(device)


def from_planes(t):
    """Device tensor (limbs, count) -> host AoS numpy array (count, limbs)."""
    return np.ascontiguousarray(t.detach().cpu().numpy().T)


def gemm(a, b, c, m, n, k, trans_a=False, mode=HP_GEMM_PLAIN, slices=0):
    """C = op(A) B on dd planes: a (2, m*k), b (2, k*n), c (2, m*n)."""
    _need(a, 2 * m * k, "a")
    _need(b, 2 * k * n, "b")
    _need(c, 2 * m * n, "c")
    rc = lib().hp_dev_gemm(int(trans_a), _ptr(a[0]), _ptr(a[1]), _ptr(b[0]), _ptr(b[1]), _ptr(c[0]), _ptr(c[1]),
                           m, n, k, mode, slices, _stream())
    _check(rc, "dev.gemm")
    return c


def solve_spd_refined(s, b, x, n, factor_kind, tol=1e-30, max_iter=20):
    """Mixed-precision SPD solve on planes: s (2, n*n), b (2, n), x (2, n) written."""
    from . import RefineStats, SolveError, HP_ENOTSPD
    import ctypes as C
    _need(s, 2 * n * n, "s")
    _need(b, 2 * n, "b")
    _need(x, 2 * n, "x")
    st = RefineStats()
    rc = lib().hp_dev_solve_spd_refined(factor_kind, _ptr(s[0]), _ptr(s[1]), _ptr(b[0]), _ptr(b[1]), n, tol, max_iter,
                                        _ptr(x[0]), _ptr(x[1]), C.byref(st), _stream())
    if rc == HP_ENOTSPD:
        raise SolveError(st.pivot)
    _check(rc, "dev.solve_spd_refined")
    return st


def overlap_s(centers, alpha, s, n):
    """dd overlap matrix into planes s (2, n*n); centers (n*3,), alpha (n,) device tensors."""
    _check(lib().hp_dev_overlap_s(_ptr(centers), _ptr(alpha), n, _ptr(s[0]), _ptr(s[1]), _stream()), "dev.overlap_s")
    return s


def mf_gemm(a, b, c, m, n, k, trans_a=False):
    """MF4 C = op(A) B on limb planes: a (4, m*k), b (4, k*n), c (4, m*n)."""
    _need(a, 4 * m * k, "a")
    _need(b, 4 * k * n, "b")
    _need(c, 4 * m * n, "c")
    _check(lib().hp_dev_mf_gemm(int(trans_a), _ptr(a), _ptr(b), _ptr(c), m, n, k, _stream()), "dev.mf_gemm")
    return c
