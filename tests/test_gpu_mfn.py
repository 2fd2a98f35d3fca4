"""MultiFloat<N> for N = 1..8 on the GPU (hp_mfn_*, hp_dev_mfn_gemm; f512 = MultiFloat<8>)
against the oracle's generic restatement (oracle/hpo_mfn.cpp), bit for bit: the element
operations on random and special values, mf_gemm / mf_gemm_atb on ragged shapes,
mf_gemv / mf_dot, mf_cholesky (and its failing pivot) and mf_cholesky_solve, the limb
conversions, and the reference's f512 unit tests (multifloat/tests.rs) run through the
device.  N = 4 goes to the f256 kernels and must give the generic restatement's bits too."""
import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu

ALL_N = [1, 2, 3, 4, 5, 6, 7, 8]


@pytest.fixture(scope="module")
def qp():
    import quadprec_amd
    quadprec_amd.lib()
    return quadprec_amd


def rand_mf(rng, count, nl, scale=1.0):
    x = np.zeros((count, nl))
    x[:, 0] = rng.standard_normal(count) * scale
    for i in range(1, nl):
        x[:, i] = x[:, i - 1] * 2.0 ** -53 * rng.uniform(-0.5, 0.5, count)
    return x


def same(a, b):
    """bit for bit, every NaN taken as one value"""
    a, b = np.array(a, dtype=np.float64), np.array(b, dtype=np.float64)
    if a.shape != b.shape:
        return False
    a[np.isnan(a)] = np.nan
    b[np.isnan(b)] = np.nan
    return np.array_equal(a.view(np.uint64), b.view(np.uint64))


def specials(nl):
    rows = [[0.0], [-0.0], [np.inf], [-np.inf], [np.nan], [1.0], [-2.0, 1e-17], [1e300, 1e283], [3.0, -1e-16, 1e-33],
            [5e-324], [1e-300, 1e-317]]
    out = np.zeros((len(rows), nl))
    for i, r in enumerate(rows):
        out[i, :min(nl, len(r))] = r[:nl]
    return out


@pytest.mark.parametrize("nl", ALL_N)
def test_elementwise_bit_exact(qp, nl):
    rng = np.random.default_rng(100 + nl)
    a, b = rand_mf(rng, 700, nl), rand_mf(rng, 700, nl, 1e3)
    b[::9] = a[::9]
    b[::13] *= -1.0
    sp = specials(nl)
    sa, sb = np.repeat(sp, len(sp), axis=0), np.tile(sp, (len(sp), 1))
    a, b = np.vstack([a, sa]), np.vstack([b, sb])
    for op in ("add", "sub", "mul", "div"):
        g, r = qp.mfn_elementwise(nl, op, a, b), O.mfn_op(nl, op, a, b)
        bad = np.where(np.any(g != r, axis=1) & ~(np.isnan(g).all(axis=1) & np.isnan(r).all(axis=1)))[0]
        assert same(g, r), (nl, op, bad[:5], a[bad[:1]], b[bad[:1]], g[bad[:1]], r[bad[:1]])
    for op in ("recip", "sqrt", "renormalize"):
        x = a * np.sign(a[:, :1]) if op == "sqrt" else a
        x = np.nan_to_num(x, nan=0.0) if op == "sqrt" else x
        assert same(qp.mfn_elementwise(nl, op, x), O.mfn_op(nl, op, x)), (nl, op)
    print(f"N={nl}: {len(a)} values x 7 ops bit-identical", flush=True)


@pytest.mark.parametrize("nl", ALL_N)
@pytest.mark.parametrize("m,n,k", [(1, 1, 1), (13, 9, 21), (33, 17, 5), (16, 16, 16), (7, 40, 3)])
def test_gemm_bit_exact(qp, nl, m, n, k):
    rng = np.random.default_rng(m * 100 + n * 10 + k + nl)
    a, b = rand_mf(rng, m * k, nl), rand_mf(rng, k * n, nl, 2.0)
    assert same(qp.mfn_gemm(nl, a, b, m, n, k), O.mfn_gemm(nl, a, b, m, n, k))
    assert same(qp.mfn_gemm(nl, a, b, m, n, k, trans_a=True), O.mfn_gemm(nl, a, b, m, n, k, trans_a=True))


def test_f512_gemm_larger_sampled_rows(qp):
    """MultiFloat<8> mf_gemm at 160 x 144 x 130 through the device entry point; 12 rows
    bit-exact against the oracle"""
    import torch
    nl, m, n, k = 8, 160, 144, 130
    rng = np.random.default_rng(7)
    a, b = rand_mf(rng, m * k, nl), rand_mf(rng, k * n, nl)
    da = torch.from_numpy(np.ascontiguousarray(a.T)).cuda()
    db = torch.from_numpy(np.ascontiguousarray(b.T)).cuda()
    dc = torch.zeros((nl, m * n), dtype=torch.float64, device="cuda")
    qp.dev.mfn_gemm(nl, da, db, dc, m, n, k)
    torch.cuda.synchronize()
    c = dc.cpu().numpy().T
    rows = [0, 1, 31, 32, 63, 64, 95, 100, 127, 128, 158, 159]
    for r in rows:
        ref = O.mfn_gemm(nl, a, b, m, n, k, row0=r, nrows=1)
        assert same(c[r * n:(r + 1) * n], ref[r * n:(r + 1) * n]), r


@pytest.mark.parametrize("nl", [2, 3, 5, 8])
def test_gemv_and_dot(qp, nl):
    rng = np.random.default_rng(200 + nl)
    for n, m in ((1, 1), (13, 37), (64, 64), (65, 130), (3, 1000)):
        a, x = rand_mf(rng, n * m, nl), rand_mf(rng, m, nl)
        assert same(qp.mfn_gemv(nl, a, x, n, m), O.mfn_gemv(nl, a, x, n, m)), (n, m)
    a, b = rand_mf(rng, 5000, nl), rand_mf(rng, 5000, nl)
    assert same(qp.mfn_dot(nl, a, b), O.mfn_dot(nl, a, b))


def spd(rng, n, nl):
    g = rng.standard_normal((n, n))
    out = np.zeros((n * n, nl))
    out[:, 0] = (g @ g.T + n * np.eye(n)).ravel()
    return out


@pytest.mark.parametrize("nl", [2, 3, 5, 8])
@pytest.mark.parametrize("n", [1, 7, 40])
def test_cholesky_and_solve(qp, nl, n):
    rng = np.random.default_rng(300 + nl + n)
    a = spd(rng, n, nl)
    l = qp.mfn_cholesky(nl, a, n)
    assert same(l, O.mfn_cholesky(nl, a, n))
    rhs = rand_mf(rng, n, nl)
    assert same(qp.mfn_cholesky_solve(nl, l, rhs, n), O.mfn_cholesky_solve(nl, l, rhs, n))


@pytest.mark.parametrize("nl", [2, 8])
def test_cholesky_failing_pivot(qp, nl):
    rng = np.random.default_rng(9)
    a = spd(rng, 12, nl)
    a[5 * 12 + 5] = 0.0
    a[5 * 12 + 5, 0] = -1.0
    with pytest.raises(qp.NotPositiveDefinite) as e:
        qp.mfn_cholesky(nl, a, 12)
    assert e.value.pivot == 5
    with pytest.raises(ValueError, match="column 5"):
        O.mfn_cholesky(nl, a, 12)


def test_conversions(qp):  # multifloat/conv.rs:119-157
    rng = np.random.default_rng(11)
    v4 = rand_mf(rng, 50, 4)
    w = qp.mfn_convert(4, 8, v4)
    assert np.array_equal(w[:, :4], v4) and np.all(w[:, 4:] == 0.0)
    assert np.array_equal(qp.mfn_convert(8, 4, w), v4)
    dd = rand_mf(rng, 50, 2)
    w8 = qp.mfn_convert(2, 8, dd)
    assert np.array_equal(w8[:, :2], dd) and np.all(w8[:, 2:] == 0.0)
    assert np.array_equal(qp.mfn_convert(8, 2, w8), dd)


def test_f512_reference_cases_on_device(qp):  # multifloat/tests.rs (f512 cases)
    def one8(v):
        r = np.zeros((1, 8))
        r[0, 0] = v
        return r

    def op(name, x, y=None):
        return qp.mfn_elementwise(8, name, x, y)

    def err(x, y):
        return abs(op("sub", x, y)[0, 0])

    one = one8(1.0)
    assert err(op("sub", op("add", one, one8(1e-80)), one), one8(1e-80)) < 1e-127
    assert err(op("mul", op("div", one, one8(3.0)), one8(3.0)), one) < 1e-127
    assert err(op("mul", op("div", one, one8(7.0)), one8(7.0)), one) < 1e-127
    s = op("sqrt", one8(2.0))
    assert err(op("mul", s, s), one8(2.0)) < 1e-127
    ones = np.zeros((20_000, 8))  # the reference sums 100000 ones; one device thread runs this chain
    ones[:, 0] = 1.0
    assert qp.mfn_dot(8, ones, ones)[0] == 20_000.0


def test_bad_limb_count(qp):
    with pytest.raises(qp.HyperprecError):
        qp.mfn_elementwise(9, "add", np.zeros((1, 9)), np.zeros((1, 9)))
    with pytest.raises(qp.HyperprecError):
        qp.mfn_gemm(9, np.zeros((1, 9)), np.zeros((1, 9)), 1, 1, 1)
