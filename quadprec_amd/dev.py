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


def batch_cholesky_solve(sizes, nrhs, s, l, b, x, status, graph=False):
    """Batched factor + solve on packed planes (see hp_dev_batch_cholesky_solve).
    s, l: (2, sum n_i^2); b, x: (2, nrhs * sum n_i); status: int32 (count,).  With graph=True
    returns a BatchGraph whose launch() replays the captured batch."""
    import torch
    sizes = np.ascontiguousarray(sizes, dtype=np.int32)
    tot = int((sizes.astype(np.int64) ** 2).sum())
    tr = int(sizes.astype(np.int64).sum()) * nrhs
    for t, num, name in ((s, 2 * tot, "s"), (l, 2 * tot, "l"), (b, 2 * tr, "b"), (x, 2 * tr, "x")):
        _need(t, num, name)
    if status.dtype != torch.int32 or status.numel() != sizes.size:
        raise TypeError("status: int32 tensor with one entry per instance")
    args = (sizes, sizes.size, nrhs, _ptr(s[0]), _ptr(s[1]), _ptr(l[0]), _ptr(l[1]), _ptr(b[0]), _ptr(b[1]),
            _ptr(x[0]), _ptr(x[1]), _ptr(status), _stream())
    if not graph:
        _check(lib().hp_dev_batch_cholesky_solve(*args), "dev.batch_cholesky_solve")
        return None
    h = lib().hp_batch_graph_create(*args)
    if not h:
        _check(4, "dev.batch_cholesky_solve(graph)")
    return BatchGraph(h)


class BatchGraph:
    def __init__(self, h):
        self.h = h

    def launch(self):
        _check(lib().hp_batch_graph_launch(self.h, _stream()), "batch graph launch")

    def __del__(self):
        if getattr(self, "h", None):
            lib().hp_batch_graph_destroy(self.h)
            self.h = None


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
