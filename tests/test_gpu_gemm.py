"""GPU parity of the dd and MF4 GEMM kernels against the oracle (bit-exact: the kernels
keep the reference's per-element operation order, crates/quad/src/linalg.rs:150-182 and
crates/quad/src/multifloat/linalg.rs:39-85)."""
import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def qp():
    import quadprec_amd
    quadprec_amd.lib()
    return quadprec_amd


def dd_of(vals):
    a = np.zeros((len(vals), 2))
    a[:, 0] = vals
    return a


def test_gemm_2x2_identity_atb(qp):  # linalg.rs:422-462
    c = qp.gemm(dd_of([1, 2, 3, 4]), dd_of([5, 6, 7, 8]), 2, 2, 2)
    assert list(c[:, 0]) == [19, 22, 43, 50] and not c[:, 1].any()
    eye = dd_of([1, 0, 0, 0, 1, 0, 0, 0, 1])
    a = dd_of(list(range(1, 10)))
    assert np.array_equal(qp.gemm(eye, a, 3, 3, 3), a)
    at = dd_of([1, 3, 5, 2, 4, 6])
    b = dd_of(list(range(1, 13)))
    assert np.array_equal(qp.gemm(at, b, 2, 4, 3), qp.gemm_atb(dd_of([1, 2, 3, 4, 5, 6]), b, 2, 4, 3))


@pytest.mark.parametrize("m,n,k", [(1, 1, 1), (37, 53, 29), (64, 64, 16), (65, 63, 17), (130, 70, 200), (3, 129, 1)])
def test_gemm_bitexact_ragged(qp, m, n, k):
    a = qp.gen_dd_normal(11 + m, m * k)
    b = qp.gen_dd_normal(12 + n, k * n)
    ref = O.gemm(a, b, m, n, k)
    got = qp.gemm(a, b, m, n, k)
    assert np.array_equal(got.view(np.uint64), ref.view(np.uint64))
    at = qp.gen_dd_normal(13 + k, k * m)
    assert np.array_equal(qp.gemm_atb(at, b, m, n, k).view(np.uint64), O.gemm_atb(at, b, m, n, k).view(np.uint64))


def test_gemm_k_zero_gives_zero(qp):
    c = qp.gemm(np.zeros((0, 2)), np.zeros((0, 2)), 3, 4, 0)
    assert c.shape == (12, 2) and not c.any()


def test_gemm_bitexact_256(qp):
    m = n = k = 256
    a = qp.gen_dd_normal(1, m * k)
    b = qp.gen_dd_normal(2, k * n)
    assert np.array_equal(qp.gemm(a, b, m, n, k).view(np.uint64), O.gemm(a, b, m, n, k).view(np.uint64))


def test_gemm_shape_error(qp):
    with pytest.raises(AssertionError):
        qp.gemm(np.zeros((5, 2)), np.zeros((4, 2)), 2, 2, 2)


@pytest.mark.parametrize("m,n,k", [(2, 2, 2), (17, 19, 23), (40, 33, 48)])
def test_mf_gemm_bitexact(qp, m, n, k):
    a = qp.gen_mf4_normal(21 + m, m * k)
    b = qp.gen_mf4_normal(22 + n, k * n)
    assert np.array_equal(qp.mf_gemm(a, b, m, n, k).view(np.uint64), O.mf4_gemm(a, b, m, n, k).view(np.uint64))
    at = qp.gen_mf4_normal(23 + k, k * m)
    assert np.array_equal(qp.mf_gemm_atb(at, b, m, n, k).view(np.uint64),
                          O.mf4_gemm_atb(at, b, m, n, k).view(np.uint64))


def test_mf_gemm_special_inputs(qp):
    # zeros, exact cancellation and mixed magnitudes exercise the data-dependent
    # compaction of multifloat/add.rs:63-95 and core.rs:77-90
    m, n, k = 5, 6, 7
    a = qp.gen_mf4_normal(5, m * k)
    b = qp.gen_mf4_normal(6, k * n)
    a[::3] = 0.0
    b[1::4] = -b[1::4]
    a[2] = [1.0, 0, 0, 0]
    b[5] = [1e-300, 0, 0, 0]
    b[7] = [1e300, 0, 0, 0]
    assert np.array_equal(qp.mf_gemm(a, b, m, n, k).view(np.uint64), O.mf4_gemm(a, b, m, n, k).view(np.uint64))


def test_mf_gemm_fast_and_general_multiply(qp):
    """mf_mul takes a pruned cascade when every limb is below 2^500 (nothing can overflow)
    and the reference's full cascade otherwise: limbs just below, at and above 2^500, an
    overflowing product, inf and nan, mixed in one wavefront, all equal to the oracle (NaNs
    compared as NaNs: the host's and the device's default NaN differ in sign)."""
    m, n, k = 8, 8, 16
    a = qp.gen_mf4_normal(31, m * k)
    b = qp.gen_mf4_normal(32, k * n)
    a[3] *= 2.0 ** 499.5
    a[10] *= 2.0 ** 500
    a[17] = [2.0 ** 500, 2.0 ** 440, 0, 0]
    a[25] = [1e200, 1e183, 0, 0]
    b[4] *= 2.0 ** 520
    b[33] = [1e200, 0, 0, 0]
    b[9] = [np.inf, 0, 0, 0]
    b[20] = [np.nan, 0, 0, 0]
    b[50] = [-0.0, 0, 0, 0]
    got, ref = qp.mf_gemm(a, b, m, n, k), O.mf4_gemm(a, b, m, n, k)
    same = (got.view(np.uint64) == ref.view(np.uint64)) | (np.isnan(got) & np.isnan(ref))
    assert same.all(), np.argwhere(~same)[:4]
    assert np.isfinite(got).sum() > 0 and np.isnan(got).sum() > 0


def test_dev_gemm_full_size_rows(qp):
    """Full-size 4096^3 plain GEMM on device planes: a sample of rows must be
    bit-identical to the oracle (each element is independent of the others)."""
    import torch
    m = n = k = 4096
    a = qp.gen_dd_normal(1, m * k)
    b = qp.gen_dd_normal(2, k * n)
    da, db = qp.dev.to_planes(a, 2), qp.dev.to_planes(b, 2)
    dc = torch.empty((2, m * n), dtype=torch.float64, device="cuda")
    qp.dev.gemm(da, db, dc, m, n, k)
    torch.cuda.synchronize()
    c = qp.dev.from_planes(dc)
    for row in (0, 1234, 4095):
        ref = O.gemm(a[row * k:(row + 1) * k], b, 1, n, k)  # one row of C = that row of A times B
        assert np.array_equal(c[row * n:(row + 1) * n].view(np.uint64), ref.view(np.uint64))
