"""Pins the oracle's MultiFloat<4> (f256) arithmetic (crates/quad/src/multifloat).

The reference's own unit tests (core.rs:296-478, mul.rs:387-478, div.rs:122-187,
add.rs:132-166) are ported, and mpmath golden vectors bound every op's error near
2^-200: a mis-ordered renormalisation pass, a lost cascade term in mul.rs:31-42 or a
skipped grow-expansion level of add.rs:63-95 shows up as errors near 2^-100 or worse.
"""
import math

import numpy as np
import pytest

import oracle_lib as O
from golden_util import exact_sum, fx, load, rel_err


def mf(*limbs):
    v = list(limbs) + [0.0] * (4 - len(limbs))
    return np.array(v, dtype=np.float64)


def non_overlapping(l):
    for i in range(3):
        if l[i] == 0.0:
            continue
        assert abs(l[i + 1]) <= math.ulp(l[i]) / 2 + 1e-300, l


def renorm(l):
    x = np.array(l, dtype=np.float64)
    O.lib().hpo_mf4_renormalize(x)
    return x


# ---------------------------------------------------------------- reference unit tests
def test_renorm_identity():  # core.rs:296-305
    l = [1.0, 1e-20, 1e-40, 0.0]
    assert list(renorm(l)) == l


def test_renorm_overlapping():  # core.rs:307-316
    assert list(renorm([1.0, 1.0, 0.0, 0.0])) == [2.0, 0.0, 0.0, 0.0]


def test_renorm_from_larger():  # core.rs:318-327
    out = np.zeros(4)
    O.lib().hpo_mf4_renormalize_from8(np.array([1.0, 1e-17, 1e-34, 1e-51, 1e-68, 0.0, 0.0, 0.0]), out)
    assert out[0] == 1.0
    assert abs(out[1] - 1e-17) < 1e-32


def test_renorm_invariants():  # core.rs:329-348, 403-430
    non_overlapping(renorm([3.14159, 2.71828, 1.41421, 1.73205]))
    r = renorm([1e10, 1e10, 1e10, 1e10])
    assert list(r) == [4e10, 0.0, 0.0, 0.0]
    l = [1.0, 0.5, -0.25, 0.125]
    r = renorm(l)
    assert abs(sum(r) - sum(l)) < 1e-15
    non_overlapping(r)


def test_mul_1_plus_tiny_squared_full_precision():  # core.rs:432-478
    tiny = 1e-20
    a = renorm([1.0, tiny, 0.0, 0.0])
    c = O.mf4_binop("mul", a, a)
    s = sum(c)
    assert abs(s - (1.0 + 2 * tiny + tiny * tiny)) < 1e-55
    assert c[0] == 1.0
    assert abs(c[1] - 2 * tiny) < 1e-35
    assert abs(c[2] - tiny * tiny) < 1e-55
    non_overlapping(c)


def test_mul_third_roundtrip():  # mul.rs:464-478
    third = renorm([3.333333333333333e-1, 1.8503717077085942e-17, -5.164681284890277e-34, 2.8687459674023424e-50])
    r = O.mf4_binop("mul", third, mf(3.0))
    assert abs(sum(float(v) for v in r) - 1.0) < 1e-60  # the reference sums the limbs in f64
    # the limbs given in the test are 1/3 to ~3e-33 only, which the exact product shows
    assert abs(float(exact_sum(*r) - 1)) < 1e-32


def test_div_one_third_and_self():  # div.rs:130-146
    third = O.mf4_binop("div", mf(1.0), mf(3.0))
    back = O.mf4_binop("mul", third, mf(3.0))
    assert abs(O.mf4_binop("sub", back, mf(1.0))[0]) < 1e-60
    x = mf(math.pi)
    assert abs(O.mf4_binop("sub", O.mf4_binop("div", x, x), mf(1.0))[0]) < 1e-60


def test_div_special_values():  # div.rs:175-187
    assert math.isinf(O.mf4_unop("recip", mf(0.0))[0])
    assert math.isnan(O.mf4_binop("div", mf(0.0), mf(0.0))[0])


def test_add_small_cancellation():  # add.rs:139-146
    c = O.mf4_binop("add", mf(1.0), mf(1e-20))
    d = O.mf4_binop("sub", c, mf(1.0))
    assert abs((d[0] + d[1]) - 1e-20) < 1e-35
    assert (O.mf4_binop("sub", mf(42.0), mf(42.0))[0]) == 0.0


def test_sqrt_two_squared():  # tests.rs:161 style
    s = O.mf4_unop("sqrt", mf(2.0))
    sq = O.mf4_binop("mul", s, s)
    assert abs(float(exact_sum(*sq) - 2)) < 1e-62


# ---------------------------------------------------------------- mpmath golden vectors
G = load("mf4_ops")
A = np.array([[fx(v) for v in c["a"]] for c in G["binary"]])
B = np.array([[fx(v) for v in c["b"]] for c in G["binary"]])


@pytest.mark.parametrize("op,bound", [("add", 2.0 ** -203), ("mul", 2.0 ** -201), ("div", 2.0 ** -198)])
def test_binary_vs_mpmath(op, bound):
    out = O.mf4_binop(op, A, B)
    for i, c in enumerate(G["binary"]):
        assert rel_err(out[i], c[op]) <= bound, (op, i, rel_err(out[i], c[op]))
        non_overlapping(out[i])


def test_recip_sqrt_vs_mpmath():
    rc = O.mf4_unop("recip", A)
    sq = O.mf4_unop("sqrt", A)
    for i, c in enumerate(G["binary"]):
        assert rel_err(rc[i], c["recip"]) <= 2.0 ** -198, i
        if "sqrt" in c:
            assert rel_err(sq[i], c["sqrt"]) <= 2.0 ** -198, i


def test_exp_vs_mpmath():
    X = np.array([[fx(v) for v in c["a"]] for c in G["exp"]])
    out = O.mf4_unop("exp", X)
    for i, c in enumerate(G["exp"]):
        k = abs(round(X[i, 0] / 0.6931471805599453))
        assert rel_err(out[i], c["exp"]) <= (400 + k) * 2.0 ** -208, (i, rel_err(out[i], c["exp"]))


def test_mf4_mac_op_count_matches_the_bench_constant():
    """bench.py prices the MF4 1024^3 gemm at MF4_OPS_PER_MAC_MEAN FP64 operations per MAC:
    re-derive that mean from the op-counting restatement on the same inputs (4 accumulation
    chains of 1024 MACs)."""
    import re
    import os
    import quadprec_amd as qp
    src = open(os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "bench.py")).read()
    const = int(re.search(r"MF4_OPS_PER_MAC_MEAN = (\d+)", src).group(1))
    n = 1024
    a = qp.gen_mf4_normal(1, n * n)
    b = qp.gen_mf4_normal(1 + 0x5EED, n * n)
    tot = cnt = 0
    for i, j in ((0, 0), (17, 900), (511, 3), (1023, 1023)):
        s = np.zeros(4)
        for p in range(n):
            ai, bj = a[i * n + p], b[p * n + j]
            tot += O.count_mf4_mac(s, ai, bj)
            cnt += 1
            s = O.mf4_binop("add", s.reshape(1, 4), O.mf4_binop("mul", ai.reshape(1, 4), bj.reshape(1, 4)))[0]
    assert abs(tot / cnt - const) <= 0.01 * const, tot / cnt
