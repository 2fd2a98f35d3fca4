"""Pins the oracle's double-double arithmetic (restating crates/quad/src/dd.rs).

Two kinds of check:
  * the reference's own unit tests for dd.rs (dd.rs:522-669), ported;
  * mpmath golden vectors (tests/golden/dd_ops.json): every op's error against the
    400-bit exact result must sit inside the error bound of the reference's algorithm.
    A dropped error term (e.g. the lo*hi cross product of dd.rs:180 or e2 of dd.rs:155)
    or a sign slip gives errors near 2^-53 and fails these bounds by ~15 orders.
"""
import math

import numpy as np
import pytest

import oracle_lib as O
from golden_util import exact_sum, frac, fx, load, rel_err

U2 = 2.0 ** -106


def dd(hi, lo=0.0):
    return np.array([hi, lo], dtype=np.float64)


def val(x):
    return x[..., 0] + x[..., 1]


# ---------------------------------------------------------------- reference unit tests
def test_add_exact():  # dd.rs:522-530
    c = O.dd_binop("add", dd(1.0), dd(1e-20))
    d = O.dd_binop("sub", c, dd(1.0))
    assert abs(val(d) - 1e-20) < 1e-35


def test_mul_precision():  # dd.rs:532-542
    small = np.finfo(float).eps * 0.5
    a = dd(1.0, small)
    c = O.dd_binop("mul", a, a)
    expected = 2.0 * small + small * small
    assert abs(val(O.dd_binop("sub", c, dd(1.0))) - expected) < 1e-31


def test_sqrt_precision():  # dd.rs:544-551
    s = O.dd_unop("sqrt", dd(2.0))
    err = O.dd_unop("abs", O.dd_binop("sub", O.dd_binop("mul", s, s), dd(2.0)))
    assert val(err) < 1e-30


def test_sqrt_edge_cases():  # dd.rs:553-559
    assert math.isnan(O.dd_unop("sqrt", dd(-1.0))[0])
    assert val(O.dd_unop("sqrt", dd(0.0))) == 0.0
    assert math.isinf(val(O.dd_unop("sqrt", dd(math.inf))))
    assert math.isnan(O.dd_unop("sqrt", dd(math.nan))[0])


def test_div_precision():  # dd.rs:561-568
    third = O.dd_binop("div", dd(1.0), dd(3.0))
    r = O.dd_binop("mul", third, dd(3.0))
    assert abs(val(O.dd_binop("sub", r, dd(1.0)))) < 1e-31


def test_catastrophic_cancellation():  # dd.rs:570-577
    r = O.dd_binop("sub", O.dd_binop("add", dd(1e16), dd(1.0)), dd(1e16))
    assert abs(val(r) - 1.0) < 1e-15


def test_mixed_f64_ops():  # dd.rs:599-608
    b = O.dd_binop("add_f64", dd(1.0), dd(1e-20))
    c = O.dd_binop("add_f64", b, dd(-1e-20))
    assert abs(val(c) - 1.0) < 1e-31
    assert val(O.dd_binop("mul_f64", dd(3.0), dd(2.0))) == 6.0


def test_abs_negative_zero():  # dd.rs:610-615
    r = O.dd_unop("abs", dd(-0.0, 0.0))
    assert not math.copysign(1.0, r[0]) < 0


def test_exp_zero_is_one():  # dd.rs:626-629
    assert list(O.dd_unop("exp", dd(0.0))) == [1.0, 0.0]


def test_exp_one_is_e():  # dd.rs:631-640
    e = O.dd_unop("exp", dd(1.0))
    assert abs(val(O.dd_binop("sub", e, dd(2.718281828459045)))) < 1e-15
    rt = O.dd_binop("mul", e, O.dd_unop("exp", dd(-1.0)))
    assert abs(val(O.dd_binop("sub", rt, dd(1.0)))) < 1e-30


def test_ln_e_is_one_and_roundtrip():  # dd.rs:642-661
    e = O.dd_unop("exp", dd(1.0))
    assert abs(val(O.dd_binop("sub", O.dd_unop("ln", e), dd(1.0)))) < 1e-30
    x = dd(3.14159265358979)
    r = O.dd_unop("exp", O.dd_unop("ln", x))
    assert abs(val(O.dd_binop("sub", r, x))) < 1e-29


def test_kahan_sum_scenario():  # dd.rs:579-588 (100k ones)
    s = dd(0.0)
    ones = np.tile(dd(1.0), (1000, 1))
    for _ in range(100):
        for i in range(1000):
            s = O.dd_binop("add", s, ones[i])
    assert val(s) == 100000.0


# ---------------------------------------------------------------- mpmath golden vectors
G = load("dd_ops")
BIN = G["binary"]
A = np.array([[fx(c["a"][0]), fx(c["a"][1])] for c in BIN])
B = np.array([[fx(c["b"][0]), fx(c["b"][1])] for c in BIN])


@pytest.mark.parametrize("op,bound", [("add", 3.5 * U2), ("mul", 8 * U2), ("div", 32 * U2)])
def test_binary_ops_vs_mpmath(op, bound):
    out = O.dd_binop(op, A, B)
    worst = 0.0
    for i, c in enumerate(BIN):
        e = rel_err(out[i], c[op])
        worst = max(worst, e)
        assert e <= bound, (op, i, e, c)
        # results are normalised: |lo| <= ulp(hi)/2
        hi, lo = out[i]
        assert abs(lo) <= math.ulp(hi) / 2 or hi == 0.0
    assert worst > 0.0  # the vectors exercise rounding, not only exact cases


def test_sqrt_recip_vs_mpmath():
    sq = O.dd_unop("sqrt", A)
    rc = O.dd_unop("recip", A)
    for i, c in enumerate(BIN):
        if "sqrt" in c:
            assert rel_err(sq[i], c["sqrt"]) <= 16 * U2, i
        assert rel_err(rc[i], c["recip"]) <= 32 * U2, i


def test_exp_vs_mpmath():
    X = np.array([[fx(c["a"][0]), fx(c["a"][1])] for c in G["exp"]])
    out = O.dd_unop("exp", X)
    for i, c in enumerate(G["exp"]):
        k = abs(round(X[i, 0] / 0.6931471805599453))
        # argument reduction with a 106-bit LN2 (dd.rs:40,111) costs ~|k| 2^-107 relative
        assert rel_err(out[i], c["exp"]) <= (40 + k) * U2, (i, X[i], rel_err(out[i], c["exp"]))


def test_add_sub_are_exact_on_disjoint_magnitudes():
    # a + b with |b| < ulp(a.lo)/2 style inputs: closed form is the TwoSum split
    a = dd(1.0, 2.0 ** -60)
    b = dd(2.0 ** -120, 0.0)
    r = O.dd_binop("add", a, b)
    assert exact_sum(*r) == exact_sum(1.0, 2.0 ** -60) + exact_sum(2.0 ** -120) or \
        abs(float(exact_sum(*r) - exact_sum(1.0, 2.0 ** -60, 2.0 ** -120))) <= 2.0 ** -106


def test_scalar_entry_points_match_array_forms():
    lib = O.lib()
    for i in range(20):
        a = O.DD(A[i, 0], A[i, 1])
        b = O.DD(B[i, 0], B[i, 1])
        for name, op in (("hpo_dd_add", "add"), ("hpo_dd_mul", "mul"), ("hpo_dd_div", "div")):
            r = getattr(lib, name)(a, b)
            arr = O.dd_binop(op, A[i], B[i])
            assert (r.hi, r.lo) == (arr[0], arr[1])
