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


def set_batch_schedule(schedule):
    """0: the autotuner picks the batched schedule per batch class; HP_BATCH_PER_INSTANCE /
    HP_BATCH_PANEL_STEPS: pinned for later batches."""
    _check(lib().hp_batch_set_schedule(int(schedule)), "set_batch_schedule")


class BatchGraph:
    def __init__(self, h):
        self.h = h

    @property
    def schedule(self):
        return lib().hp_batch_graph_schedule(self.h)

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


def mfn_gemm(nl, a, b, c, m, n, k, trans_a=False):
    """MultiFloat<nl> C = op(A) B on limb planes: a (nl, m*k), b (nl, k*n), c (nl, m*n)."""
    _need(a, nl * m * k, "a")
    _need(b, nl * k * n, "b")
    _need(c, nl * m * n, "c")
    _check(lib().hp_dev_mfn_gemm(nl, int(trans_a), _ptr(a), _ptr(b), _ptr(c), m, n, k, _stream()), "dev.mfn_gemm")
    return c


# ---------------------------------------------------------------- adjoints
class OverlapAdjoint:
    """Saved state of one forward evaluation of F1 = log det S and F2 = b^T S^-1 b for the
    overlap matrix S(centers, alpha) (hp_dev_overlap_forward): the factor L and x = S^-1 b.
    centers (n, 3) / (3n,), alpha (n,) float64 device tensors; b dd planes (2, n)."""

    def __init__(self, centers, alpha, b, factor_kind=None):
        import ctypes as C
        import torch
        from . import AdjointInfo, HP_FACTOR_DD, HP_ENOTSPD, SolveError
        n = alpha.numel()
        _need(centers, 3 * n, "centers")
        _need(alpha, n, "alpha")
        _need(b, 2 * n, "b")
        self.n = n
        self.h = None
        self.info = AdjointInfo()
        h = C.c_void_p()
        kind = HP_FACTOR_DD if factor_kind is None else factor_kind
        rc = lib().hp_dev_overlap_forward(_ptr(centers), _ptr(alpha), _ptr(b[0]), _ptr(b[1]), n, kind, C.byref(h),
                                          C.byref(self.info), _stream())
        if rc == HP_ENOTSPD:
            raise SolveError(self.info.pivot)
        _check(rc, "dev.overlap_forward")
        self.h = h.value
        self.device = alpha.device

    @property
    def logdet(self):
        return np.array(self.info.logdet[:])

    @property
    def quad(self):
        return np.array(self.info.quad[:])

    def gradients(self, gemm_mode=HP_GEMM_PLAIN, logdet=True):
        """Both gradients as dd planes (2, 4n): (g_logdet or None, g_quad)."""
        import torch
        gq = torch.empty((2, 4 * self.n), dtype=torch.float64, device=self.device)
        gl = torch.empty_like(gq) if logdet else None
        rc = lib().hp_dev_overlap_gradients(self.h, gemm_mode, _ptr(gl[0]) if logdet else None,
                                            _ptr(gl[1]) if logdet else None, _ptr(gq[0]), _ptr(gq[1]), _stream())
        _check(rc, "dev.overlap_gradients")
        return gl, gq

    def backward(self, cot_logdet, cot_quad, gemm_mode=HP_GEMM_PLAIN, grad_b=True):
        """Vector-Jacobian product with dd cotangents (hi, lo) of (F1, F2): returns dd planes
        (grad_centers (2, 3n), grad_alpha (2, n), grad_b (2, n) or None)."""
        import torch
        cot = np.array([cot_logdet[0], cot_logdet[1], cot_quad[0], cot_quad[1]], dtype=np.float64)
        gc = torch.empty((2, 3 * self.n), dtype=torch.float64, device=self.device)
        ga = torch.empty((2, self.n), dtype=torch.float64, device=self.device)
        gb = torch.empty((2, self.n), dtype=torch.float64, device=self.device) if grad_b else None
        rc = lib().hp_dev_overlap_backward(self.h, cot, gemm_mode, _ptr(gc[0]), _ptr(gc[1]), _ptr(ga[0]), _ptr(ga[1]),
                                           _ptr(gb[0]) if grad_b else None, _ptr(gb[1]) if grad_b else None,
                                           _stream())
        _check(rc, "dev.overlap_backward")
        return gc, ga, gb

    def refresh_info(self):
        import ctypes as C
        _check(lib().hp_adjoint_get_info(self.h, C.byref(self.info)), "adjoint info")
        return self.info

    def close(self):
        if self.h:
            lib().hp_adjoint_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


def _overlap_function():
    import torch

    class OverlapLogdetQuad(torch.autograd.Function):
        """(log det S, b^T S^-1 b) of the overlap matrix as a torch.autograd.Function on
        binary64 leaves: centers (n, 3), alpha (n,), b (n,) (b enters as an exact dd with a
        zero low part).  The values are returned rounded to binary64; the dd values stay on
        ctx.state.  backward saves nothing but the factor and x (inside ctx.state) and returns
        the dd vector-Jacobian product rounded to binary64."""

        @staticmethod
        def forward(ctx, centers, alpha, b, gemm_mode=HP_GEMM_PLAIN):
            bp = torch.stack([b.detach().contiguous(), torch.zeros_like(b)])
            st = OverlapAdjoint(centers.detach().contiguous().reshape(-1), alpha.detach().contiguous(), bp)
            ctx.state = st
            ctx.gemm_mode = gemm_mode
            ld, q = st.logdet, st.quad
            return (torch.tensor(ld[0] + ld[1], dtype=torch.float64, device=alpha.device),
                    torch.tensor(q[0] + q[1], dtype=torch.float64, device=alpha.device))

        @staticmethod
        def backward(ctx, g_ld, g_q):
            c1 = (float(g_ld), 0.0) if g_ld is not None else (0.0, 0.0)
            c2 = (float(g_q), 0.0) if g_q is not None else (0.0, 0.0)
            gc, ga, gb = ctx.state.backward(c1, c2, ctx.gemm_mode, grad_b=ctx.needs_input_grad[2])
            gcen = (gc[0] + gc[1]).reshape(-1, 3)
            galp = ga[0] + ga[1]
            gbv = (gb[0] + gb[1]) if gb is not None else None
            return gcen, galp, gbv, None

    return OverlapLogdetQuad


_OVERLAP_FN = None


def overlap_logdet_quad(centers, alpha, b, gemm_mode=HP_GEMM_PLAIN):
    """Differentiable (log det S, b^T S^-1 b) for torch float64 tensors on the device."""
    global _OVERLAP_FN
    if _OVERLAP_FN is None:
        _OVERLAP_FN = _overlap_function()
    return _OVERLAP_FN.apply(centers, alpha, b, gemm_mode)
