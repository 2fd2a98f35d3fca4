"""The oracle's MultiFloat<N> for N = 1..8 (oracle/hpo_mfn.cpp, a restatement of the
reference's generic multifloat/{core,add,mul,div,math,linalg}.rs), pinned by:

  * the separately written MultiFloat<4> oracle (hpo_core.hpp): at N = 4 every operation
    and every linear-algebra routine must agree bit for bit;
  * the reference's own f512 (= MultiFloat<8>) unit tests (multifloat/tests.rs), ported;
  * mpmath at 250 digits: for N = 2, 3, 5, 8 the operations' relative errors stay within
    a few units of 2^(2 - 53 N), the linear algebra within its sums' bounds;
  * invariants: gemm_atb is gemm of the transpose, gemv's rows are dots, L L^T = A."""
import numpy as np
import pytest

import oracle_lib as O

mp = pytest.importorskip("mpmath")


def rand_mf(rng, count, nl, scale=1.0):
    """non-overlapping expansions: limb i+1 below half an ulp of limb i"""
    x = np.zeros((count, nl))
    x[:, 0] = rng.standard_normal(count) * scale
    for i in range(1, nl):
        x[:, i] = x[:, i - 1] * 2.0 ** -53 * rng.uniform(-0.5, 0.5, count)
    return x


def bits_equal(a, b):
    """bit for bit, every NaN taken as one value (a NaN's sign and payload carry nothing)"""
    a, b = np.array(a, dtype=np.float64), np.array(b, dtype=np.float64)
    a[np.isnan(a)] = np.nan
    b[np.isnan(b)] = np.nan
    return np.array_equal(a.view(np.uint64), b.view(np.uint64))


def to_mp(v):
    return mp.fsum([mp.mpf(float(x)) for x in v])


# ---------------------------------------------------------------- N = 4 cross-check
def test_n4_equals_the_mf4_oracle_elementwise():
    rng = np.random.default_rng(1)
    a, b = rand_mf(rng, 600, 4), rand_mf(rng, 600, 4, 1e3)
    b[::7] = a[::7]  # exact cancellation in sub
    b[::11] *= -1.0
    for op in ("add", "sub", "mul", "div"):
        assert bits_equal(O.mfn_op(4, op, a, b), O.mf4_binop(op, a, b)), op
    for op in ("recip", "sqrt", "renormalize"):
        x = a * np.sign(a[:, :1]) if op == "sqrt" else a
        assert bits_equal(O.mfn_op(4, op, x), O.mf4_unop(op, x)), op
    specials = np.array([[0.0, 0, 0, 0], [-0.0, 0, 0, 0], [np.inf, 0, 0, 0], [-np.inf, 0, 0, 0], [np.nan, 0, 0, 0],
                         [1.0, 0, 0, 0], [-2.0, 1e-17, 0, 0], [1e300, 1e283, 0, 0]])
    sa, sb = np.repeat(specials, len(specials), axis=0), np.tile(specials, (len(specials), 1))
    for op in ("add", "sub", "mul", "div"):
        assert bits_equal(O.mfn_op(4, op, sa, sb), O.mf4_binop(op, sa, sb)), op
    for op in ("recip", "sqrt"):
        assert bits_equal(O.mfn_op(4, op, specials), O.mf4_unop(op, specials)), op


def test_n4_equals_the_mf4_oracle_linalg():
    rng = np.random.default_rng(2)
    m, n, k = 13, 9, 11
    a, b = rand_mf(rng, m * k, 4), rand_mf(rng, k * n, 4)
    assert bits_equal(O.mfn_gemm(4, a, b, m, n, k), O.mf4_gemm(a, b, m, n, k))
    assert bits_equal(O.mfn_gemm(4, a, b, m, n, k, trans_a=True), O.mf4_gemm_atb(a, b, m, n, k))
    x = rand_mf(rng, k, 4)
    assert bits_equal(O.mfn_gemv(4, a, x, m, k), O.mf4_gemv(a, x, m, k))
    nn = 17
    g = rng.standard_normal((nn, nn))
    spd = np.zeros((nn * nn, 4))
    spd[:, 0] = (g @ g.T + nn * np.eye(nn)).ravel()
    l4 = O.mfn_cholesky(4, spd, nn)
    assert bits_equal(l4, O.mf4_cholesky(spd, nn))
    rhs = rand_mf(rng, nn, 4)
    assert bits_equal(O.mfn_cholesky_solve(4, l4, rhs, nn), O.mf4_cholesky_solve(l4, rhs, nn))


# ---------------------------------------------------------------- reference f512 tests
F512_EPS = 1e-127


def one8(v):
    r = np.zeros((1, 8))
    r[0, 0] = v
    return r


def err8(a, b):  # (a - b).abs().limbs[0].abs()
    return abs(O.mfn_op(8, "sub", a, b)[0, 0])


def test_f512_reference_cases():  # multifloat/tests.rs:38-210, 323-350, 560-565
    one = one8(1.0)
    tiny = one8(1e-80)
    assert err8(O.mfn_op(8, "sub", O.mfn_op(8, "add", one, tiny), one), tiny) < F512_EPS
    third = O.mfn_op(8, "div", one, one8(3.0))
    assert err8(O.mfn_op(8, "mul", third, one8(3.0)), one) < F512_EPS
    eps = 1e-80
    val = O.mfn_op(8, "add", one, one8(eps))
    diff = O.mfn_op(8, "sub", O.mfn_op(8, "mul", val, val), one)
    assert err8(diff, one8(2.0 * eps)) < eps * eps * 10.0
    assert err8(O.mfn_op(8, "mul", one8(1e100), one8(1e-100)), one) < 1e-15
    seventh = O.mfn_op(8, "div", one, one8(7.0))
    assert err8(O.mfn_op(8, "mul", seventh, one8(7.0)), one) < F512_EPS
    pi = one8(np.pi)
    assert err8(O.mfn_op(8, "div", pi, pi), one) < F512_EPS
    s = O.mfn_op(8, "sqrt", one8(2.0))
    assert err8(O.mfn_op(8, "mul", s, s), one8(2.0)) < F512_EPS
    s = O.mfn_op(8, "sqrt", one8(1e100))
    assert err8(O.mfn_op(8, "mul", s, s), one8(1e100)) / 1e100 < F512_EPS
    assert O.mfn_op(8, "sqrt", one8(0.0))[0, 0] == 0.0
    assert np.isnan(O.mfn_op(8, "sqrt", one8(-1.0))[0, 0])
    assert np.isinf(O.mfn_op(8, "sqrt", one8(np.inf))[0, 0])
    a, b = one8(1e50), one8(1e-70)
    assert err8(O.mfn_op(8, "sub", O.mfn_op(8, "add", a, b), a), b) < F512_EPS
    # kahan_sum_100k: ZERO + ONE, 100000 times, as one row of ONE * ONE products (exact)
    ones = np.zeros((100_000, 8))
    ones[:, 0] = 1.0
    assert O.mfn_dot(8, ones, ones)[0] == 100_000.0
    e1, e2 = np.zeros((3, 8)), np.zeros((3, 8))
    e1[0, 0] = e2[1, 0] = 1.0
    assert O.mfn_dot(8, e1, e2)[0] == 0.0


# ---------------------------------------------------------------- mpmath accuracy
@pytest.mark.parametrize("nl", [2, 3, 5, 8])
def test_ops_against_mpmath(nl):
    mp.mp.dps = 250
    rng = np.random.default_rng(10 + nl)
    cnt = 150
    a, b = rand_mf(rng, cnt, nl), rand_mf(rng, cnt, nl, 10.0)
    u = mp.mpf(2) ** (2 - 53 * nl)
    res = {op: O.mfn_op(nl, op, a, b) for op in ("add", "sub", "mul", "div")}
    rec, sq = O.mfn_op(nl, "recip", a), O.mfn_op(nl, "sqrt", a * np.sign(a[:, :1]))  # |a|: every limb flipped
    worst = {}
    for i in range(cnt):
        x, y = to_mp(a[i]), to_mp(b[i])
        for op, exact, scale, c in (("add", x + y, abs(x) + abs(y), 4), ("sub", x - y, abs(x) + abs(y), 4),
                                    ("mul", x * y, abs(x * y), 8), ("div", x / y, abs(x / y), 32)):
            e = abs(to_mp(res[op][i]) - exact) / scale
            worst[op] = max(worst.get(op, 0), float(e / u))
            assert e <= c * u, (nl, op, i, float(e / u))
        e = abs(to_mp(rec[i]) - 1 / x) / abs(1 / x)
        assert e <= 32 * u, (nl, "recip", float(e / u))
        e = abs(to_mp(sq[i]) - mp.sqrt(abs(x))) / mp.sqrt(abs(x))
        assert e <= 32 * u, (nl, "sqrt", float(e / u))
    print(nl, {k: round(v, 3) for k, v in worst.items()})


@pytest.mark.parametrize("nl", [2, 3, 8])
def test_linalg_against_mpmath_and_invariants(nl):
    mp.mp.dps = 250
    rng = np.random.default_rng(20 + nl)
    m, n, k = 6, 5, 7
    a, b = rand_mf(rng, m * k, nl), rand_mf(rng, k * n, nl)
    c = O.mfn_gemm(nl, a, b, m, n, k)
    u = mp.mpf(2) ** (2 - 53 * nl)
    for i in range(m):
        for j in range(n):
            terms = [to_mp(a[i * k + p]) * to_mp(b[p * n + j]) for p in range(k)]
            bound = 8 * (k + 1) * u * mp.fsum([abs(t) for t in terms])
            assert abs(to_mp(c[i * n + j]) - mp.fsum(terms)) <= bound
    at = a.reshape(m, k, nl).transpose(1, 0, 2).reshape(k * m, nl).copy()
    assert bits_equal(O.mfn_gemm(nl, at, b, m, n, k, trans_a=True), c)
    x = rand_mf(rng, k, nl)
    y = O.mfn_gemv(nl, a, x, m, k)
    for i in range(m):
        assert bits_equal(y[i], O.mfn_dot(nl, a[i * k:(i + 1) * k], x))
    nn = 9
    g = rng.standard_normal((nn, nn))
    spd = np.zeros((nn * nn, nl))
    spd[:, 0] = (g @ g.T + nn * np.eye(nn)).ravel()
    lf = O.mfn_cholesky(nl, spd, nn)
    L = [[to_mp(lf[i * nn + j]) if j <= i else mp.mpf(0) for j in range(nn)] for i in range(nn)]
    for i in range(nn):
        for j in range(i + 1):
            s = mp.fsum([L[i][q] * L[j][q] for q in range(nn)])
            assert abs(s - spd[i * nn + j, 0]) <= 64 * nn * u * (abs(spd[i * nn + i, 0]) + abs(spd[j * nn + j, 0]))
    rhs = rand_mf(rng, nn, nl)
    xs = O.mfn_cholesky_solve(nl, lf, rhs, nn)
    A = mp.matrix(nn, nn)
    for i in range(nn):
        for j in range(nn):
            A[i, j] = spd[i * nn + j, 0]
    r = A * mp.matrix([to_mp(v) for v in xs]) - mp.matrix([to_mp(v) for v in rhs])
    assert max(abs(v) for v in r) <= 1e4 * nn * u * mp.norm(A, 1) * max(abs(to_mp(v)) for v in xs)
    bad = spd.copy()
    bad[4 * nn + 4, 0] = -1.0
    with pytest.raises(ValueError, match="column 4"):
        O.mfn_cholesky(nl, bad, nn)


def test_n1_and_renormalize_shapes():
    """N = 1 is plain f64 arithmetic through the same code (ceil(log2 1) = 0 steps plus the
    reference's extra ones); renormalize makes overlapping limbs non-overlapping and keeps
    the sum."""
    rng = np.random.default_rng(5)
    a, b = rng.standard_normal(50), rng.standard_normal(50)
    assert np.array_equal(O.mfn_op(1, "add", a, b)[:, 0], a + b)
    assert np.array_equal(O.mfn_op(1, "mul", a, b)[:, 0], a * b)
    mp.mp.dps = 250
    x = rng.standard_normal((40, 8))  # heavily overlapping limbs
    r = O.mfn_op(8, "renormalize", x)
    for i in range(40):
        assert abs(to_mp(r[i]) - to_mp(x[i])) <= mp.mpf(2) ** -400 * mp.fsum([abs(v) for v in x[i]])
        nz = r[i][r[i] != 0]
        assert np.all(np.abs(nz[1:]) <= np.abs(nz[:-1]) * 2.0 ** -52)


def test_mac_op_count_matches_the_mf4_counter():
    """The FP64 operation count of one MAC (the GEMM's work figure) agrees at N = 4 with the
    MultiFloat<4> oracle's counting-double instantiation, value for value."""
    rng = np.random.default_rng(8)
    s, a, b = rand_mf(rng, 40, 4), rand_mf(rng, 40, 4), rand_mf(rng, 40, 4)
    for i in range(40):
        assert O.count_mfn_mac(4, s[i], a[i], b[i]) == O.lib().hpo_count_mf4_mac(O._c(s[i]), O._c(a[i]), O._c(b[i]))
    counts = {nl: np.mean([O.count_mfn_mac(nl, *rand_mf(rng, 3, nl)) for _ in range(20)]) for nl in (2, 3, 8)}
    assert counts[2] < counts[3] < 515 < counts[8]


def test_f512_mac_op_count_matches_the_bench_constant():
    """bench.py prices the f512 512^3 gemm at MFN8_OPS_PER_MAC_MEAN FP64 operations per MAC:
    re-derive that mean from the op-counting restatement on the same inputs (four
    accumulation chains of 512 MACs)."""
    import os
    import re
    import quadprec_amd as qp
    src = open(os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "bench.py")).read()
    const = int(re.search(r"MFN8_OPS_PER_MAC_MEAN = (\d+)", src).group(1))
    n = 512
    a, b = qp.gen_mfn_normal(8, 1, n * n), qp.gen_mfn_normal(8, 2, n * n)
    tot = cnt = 0
    for i, j in ((0, 0), (17, 400), (255, 3), (511, 511)):
        s = np.zeros(8)
        for p in range(n):
            ai, bj = a[i * n + p], b[p * n + j]
            tot += O.count_mfn_mac(8, s, ai, bj)
            cnt += 1
            s = O.mfn_op(8, "add", s, O.mfn_op(8, "mul", ai, bj))[0]
    assert abs(tot / cnt - const) <= 0.01 * const, tot / cnt
