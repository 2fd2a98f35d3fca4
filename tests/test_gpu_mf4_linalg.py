"""GPU MultiFloat<4> reciprocal / division / square root (crates/quad/src/multifloat/div.rs,
math.rs:22-40) and mf_cholesky / mf_cholesky_solve / mf_gemv (multifloat/linalg.rs:17-172)
against the oracle, bit for bit: every element follows the reference's operation order.
Also the reference's own f256 Cholesky tests (multifloat/tests.rs)."""
import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def qp():
    import quadprec_amd
    quadprec_amd.lib()
    return quadprec_amd


def mf4_random(rng, n, spread=3):
    hi = rng.standard_normal(n) * 10.0 ** rng.integers(-spread, spread + 1, n)
    v = np.zeros((n, 4))
    v[:, 0] = hi
    v[:, 1] = hi * 2.0 ** -53 * rng.uniform(-0.5, 0.5, n)
    v[:, 2] = v[:, 1] * 2.0 ** -53 * rng.uniform(-0.5, 0.5, n)
    v[:, 3] = v[:, 2] * 2.0 ** -53 * rng.uniform(-0.5, 0.5, n)
    return v


def same(a, b):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    return np.array_equal(a.view(np.uint64), b.view(np.uint64)) or (
        np.array_equal(np.isnan(a), np.isnan(b)) and np.array_equal(a[~np.isnan(a)], b[~np.isnan(b)]))


def spd_mf4(rng, n):
    """B B^T + n I in MF4, B ~ N(0,1) MF4 (the products formed with the oracle's MF4 GEMM)"""
    b = mf4_random(rng, n * n, 0)
    bt = b.reshape(n, n, 4).transpose(1, 0, 2).reshape(n * n, 4).copy()
    a = O.mf4_gemm(b, bt, n, n, n).reshape(n, n, 4)
    for i in range(n):
        a[i, i] = O.mf4_binop("add", a[i, i].reshape(1, 4), np.array([[float(n), 0, 0, 0]]))[0]
    return a.reshape(n * n, 4)


@pytest.mark.parametrize("op", ["recip", "sqrt", "div"])
def test_elementwise_bit_exact(qp, op):
    rng = np.random.default_rng(len(op))
    a = mf4_random(rng, 3000, 8)
    b = mf4_random(rng, 3000, 8)
    specials = np.array([[0.0, 0, 0, 0], [-0.0, 0, 0, 0], [np.inf, 0, 0, 0], [-np.inf, 0, 0, 0], [-4.0, 0, 0, 0],
                         [np.nan, 0, 0, 0], [1e-300, 0, 0, 0], [1e300, 0, 0, 0]])
    a = np.concatenate([a, specials, specials])
    b = np.concatenate([b, specials, specials[::-1]])
    if op == "sqrt":
        a[: len(a) // 2, 0] = np.abs(a[: len(a) // 2, 0])
    got = qp.mf4_elementwise(op, a, b if op == "div" else None)
    ref = O.mf4_binop("div", a, b) if op == "div" else O.mf4_unop(op, a)
    bad = [i for i in range(len(a)) if not same(got[i], ref[i])]
    print(f"mf4 {op}: {len(a)} values, {len(bad)} differ", flush=True)
    assert not bad, [(a[i], got[i], ref[i]) for i in bad[:3]]


@pytest.mark.parametrize("n", [1, 2, 5, 33, 64, 130])
def test_mf_cholesky_and_solve_bit_exact(qp, n):
    rng = np.random.default_rng(n)
    a = spd_mf4(rng, n)
    L = qp.mf_cholesky(a, n)
    ref = O.mf4_cholesky(a, n)
    print(f"mf_cholesky n={n}: identical {np.mean(np.all(L == ref, axis=1)):.3f}", flush=True)
    assert same(L, ref)
    b = mf4_random(rng, n, 1)
    x = qp.mf_cholesky_solve(L, b, n)
    xr = O.mf4_cholesky_solve(ref, b, n)
    assert same(x, xr)


def test_mf_cholesky_not_spd(qp):
    n = 6
    rng = np.random.default_rng(1)
    a = spd_mf4(rng, n).reshape(n, n, 4)
    a[3, 3] = [-50.0, 0, 0, 0]
    a = a.reshape(n * n, 4)
    with pytest.raises(qp.NotPositiveDefinite) as ei:
        qp.mf_cholesky(a, n)
    assert ei.value.pivot == 3
    with pytest.raises(ValueError):
        O.mf4_cholesky(a, n)


def test_f256_cholesky_solve_simple(qp):
    """multifloat/tests.rs:618-639: A = [[4, 2], [2, 3]], b = [1, 1]; the residual of A x
    through mf_gemv below F256_EPS = 1e-63 (exact rational arithmetic on the limbs)"""
    from fractions import Fraction
    a = np.zeros((4, 4))
    a[:, 0] = [4.0, 2.0, 2.0, 3.0]
    L = qp.mf_cholesky(a, 2)
    b = np.array([[1.0, 0, 0, 0], [1.0, 0, 0, 0]])
    x = qp.mf_cholesky_solve(L, b, 2)
    ax = qp.mf_gemv(a, x, 2, 2)
    for i in range(2):
        err = abs(sum(Fraction(float(v)) for v in ax[i]) - 1)
        assert err < Fraction(1, 10 ** 63)
    assert abs(x[0].sum() - 0.125) < 1e-15 and abs(x[1].sum() - 0.25) < 1e-15


@pytest.mark.parametrize("n,m", [(1, 1), (7, 13), (64, 64), (65, 130), (200, 17), (3, 0)])
def test_mf_gemv_bit_exact(qp, n, m):
    rng = np.random.default_rng(n * 100 + m)
    a = mf4_random(rng, n * m, 2)
    x = mf4_random(rng, m, 2)
    y = qp.mf_gemv(a, x, n, m)
    ref = O.mf4_gemv(a, x, n, m) if m else np.zeros((n, 4))
    assert same(y, ref)


def test_dd_mf4_conversions(qp):  # multifloat/conv.rs:120-131
    rng = np.random.default_rng(3)
    m = mf4_random(rng, 1000, 5)
    d = qp.dd_from_mf4(m)
    assert np.array_equal(d, m[:, :2])
    back = qp.mf4_from_dd(d)
    assert np.array_equal(back[:, :2], d) and not back[:, 2:].any()
