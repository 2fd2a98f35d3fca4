"""The oracle's text wire formats (crates/quad/src/dd.rs:331-514, multifloat/fmt.rs,
multifloat/traits.rs:7-150), pinned independently of its own code:
  * `{:.Pe}` formatting against exact decimal arithmetic (Python's Decimal holds every
    double exactly; ties to even) and against Python's own correctly rounded format();
  * Rust's f64 parse against Python's correctly rounded float() on literals of Rust's
    grammar, and the grammar itself against the cases Rust's documentation lists;
  * the dd / MF4 FromStr digit loops against mpmath at 300 digits (error bounds of the
    accumulation: a dropped term, a wrong power of ten or a misplaced digit fails them) and
    the reference's own parse tests (dd.rs:655-712, multifloat/tests.rs:357-422);
  * the serde string forms against strings built from the exact formatter, the split rules
    on hand-made cases, and round trips."""
import math
import struct
from decimal import Decimal, ROUND_HALF_EVEN, localcontext

import mpmath
import numpy as np
import pytest

import oracle_lib as O


def rust_exp(x, prec, plus=False):
    """format!("{:.prec$e}") from exact decimal arithmetic."""
    if x != x:
        return "NaN"
    sign = "-" if math.copysign(1.0, x) < 0 else ("+" if plus else "")
    if math.isinf(x):
        return sign + "inf"
    if x == 0:
        return sign + "0" + ("." + "0" * prec if prec else "") + "e0"
    with localcontext() as ctx:
        ctx.prec = 2000
        d = Decimal(abs(x))
        e10 = d.adjusted()
        q = d.scaleb(prec - e10).quantize(Decimal(1), rounding=ROUND_HALF_EVEN)
    digits = str(int(q))
    if len(digits) == prec + 2:  # rounded up to 10^(prec+1)
        e10 += 1
        digits = digits[:-1]
    assert len(digits) == prec + 1
    return f"{sign}{digits[0]}{'.' + digits[1:] if prec else ''}e{e10}"


def py_exp(x, prec, plus=False):
    """Python's correctly rounded format(), respelled with Rust's exponent."""
    s = format(x, ("+" if plus else "") + f".{prec}e")
    m, e = s.split("e")
    return f"{m}e{int(e)}"


def random_doubles(rng, n):
    bits = rng.integers(0, 2**63, n, dtype=np.uint64) | (rng.integers(0, 2, n, dtype=np.uint64) << np.uint64(63))
    x = bits.view(np.float64)
    return [float(v) for v in x if math.isfinite(v)]


SPECIAL = [0.0, -0.0, 1.0, -1.0, 0.1, 1 / 3, 2.0 ** -1074, 2.0 ** -1022, 2.0 ** -1022 - 2.0 ** -1074,
           1.7976931348623157e308, 9.999999999999999e22, 1e23, 0.5, 2.5, 1234567890123445.0, 1234567890123455.0,
           999999999999999.9, 9.5, 99.5, 2.0 ** 53, 2.0 ** 63, 123456789.0, 5e-324, 1e-300, 1e300]


@pytest.mark.parametrize("prec", [0, 1, 2, 5, 14, 16, 17, 18, 20, 32, 64])
def test_format_exp_is_exact(prec):
    rng = np.random.default_rng(prec)
    xs = SPECIAL + [-v for v in SPECIAL] + random_doubles(rng, 400) + list(rng.standard_normal(100))
    for x in xs:
        want = rust_exp(x, prec)
        assert O.format_exp(x, prec) == want, (x, prec)
        assert O.format_exp(x, prec, True) == rust_exp(x, prec, True), (x, prec)
        assert want == py_exp(x, prec)


def test_format_ties_round_to_even():
    # 1234567890123445 = 1.234567890123445e15 exactly: 15 significant digits + a 5
    assert O.format_exp(1234567890123445.0, 14) == "1.23456789012344e15"
    assert O.format_exp(1234567890123455.0, 14) == "1.23456789012346e15"
    assert O.format_exp(2.5, 0) == "2e0" and O.format_exp(3.5, 0) == "4e0"
    assert O.format_exp(9.5, 0) == "1e1" and O.format_exp(0.5, 0) == "5e-1"


def test_format_specials():
    assert O.format_exp(float("nan"), 3, True) == "NaN"
    assert O.format_exp(-float("nan"), 3) == "NaN"
    assert O.format_exp(float("inf"), 3, True) == "+inf"
    assert O.format_exp(-float("inf"), 3) == "-inf"
    assert O.format_exp(-0.0, 2, True) == "-0.00e0"
    assert O.format_exp(1e-7, 1) == "1.0e-7"


VALID = ["3.14", "-3.14", "2.5E10", "2.5e-10", "5.", ".5", "0", "-0", "+1", "1e+5", "1E-5", "inf", "-inf",
         "+infinity", "InFiNiTy", "NaN", "nan", "-nan", "000123.4500", "1e400", "1e-400", "-1e-400",
         "4.9406564584124654e-324", "2.4703282292062327e-324", "2.4703282292062328e-324",
         "1.7976931348623157e308", "1.7976931348623158e308", "1.797693134862315807e308", "0.000000001e9",
         "123456789012345678901234567890", "9007199254740993", "1e0000000000000000000000001"]
INVALID = ["", "+", "-", ".", "e5", "1e", "1e+", "1e-", " 1", "1 ", "1_0", "0x10", "1.2.3", "infinit", "nana",
           "+-1", "1e5.0", "1.0e5e5", "--1", "in f", "١"]


def test_parse_f64_grammar_and_values():
    for s in VALID:
        v = O.parse_f64(s)
        assert v is not None, s
        want = float(s)
        assert (v == want and math.copysign(1, v) == math.copysign(1, want)) or (v != v and want != want), s
    assert math.copysign(1, O.parse_f64("-nan")) < 0
    for s in INVALID:
        assert O.parse_f64(s) is None, s


def test_parse_f64_random_literals_round_correctly():
    rng = np.random.default_rng(7)
    xs = random_doubles(rng, 300) + SPECIAL
    for x in xs:
        for prec in (0, 3, 14, 16, 17, 25, 40):
            s = rust_exp(x, prec)
            v = O.parse_f64(s)
            assert v == float(s), s
        assert O.parse_f64(rust_exp(x, 17)) == x  # 17 significant digits round-trip
    for _ in range(500):  # plain decimals with long digit strings and leading zeros
        ip = "".join(rng.choice(list("0123456789"), rng.integers(0, 30)))
        fp = "".join(rng.choice(list("0123456789"), rng.integers(0, 40)))
        if not ip and not fp:
            continue
        s = ip + ("." + fp if fp or rng.integers(0, 2) else "")
        if rng.integers(0, 2):
            s += "e" + str(int(rng.integers(-340, 320)))
        assert O.parse_f64(s) == float(s), s


def dd_val(v):
    return mpmath.mpf(float(v[0])) + mpmath.mpf(float(v[1]))


def mf_val(v):
    return sum(mpmath.mpf(float(x)) for x in v)


def test_dd_from_str_reference_cases():  # dd.rs:655-712
    v = O.dd_from_str("1.0")
    assert v[0] == 1.0 and v[1] == 0.0
    pi = O.dd_from_str("3.14159265358979323846264338327950288")
    lit = 3.14159265358979323846264338327950288
    assert abs(pi[0] + pi[1] - lit) <= abs(math.pi - lit) and pi[1] != 0.0
    assert abs(O.dd_from_str("1.23e-10")[0] - 1.23e-10) < 1e-25
    assert abs(O.dd_from_str("5.0E3")[0] - 5000.0) < 1e-10
    assert abs(O.dd_from_str("-42.5")[0] + 42.5) < 1e-30
    for bad in ["", "abc", "1.2.3", "-", "1.0e"]:
        assert O.dd_from_str(bad) is None, bad


def test_mf4_from_str_reference_cases():  # multifloat/tests.rs:357-415
    pi = O.mf4_from_str("3.14159265358979323846264338327950288419716939937510")
    assert abs(pi[0] - math.pi) < 1e-15 and (pi[1] != 0.0 or pi[2] != 0.0)
    assert abs(O.mf4_from_str("1.23e-10")[0] - 1.23e-10) < 1e-25
    assert abs(O.mf4_from_str("5.0E3")[0] - 5000.0) < 1e-10
    assert abs(O.mf4_from_str("-42.5")[0] + 42.5) < 1e-63
    for bad in ["", "abc", "1.2.3", "-", "1.0e"]:
        assert O.mf4_from_str(bad) is None, bad


def random_literal(rng):
    nd = int(rng.integers(1, 45))
    digits = "".join(rng.choice(list("0123456789"), nd))
    cut = int(rng.integers(0, nd + 1))
    s = digits[:cut] + "." + digits[cut:] if rng.integers(0, 3) else digits
    if s in (".",):
        s = "1"
    e = int(rng.integers(-60, 60)) if rng.integers(0, 2) else 0
    return ("-" if rng.integers(0, 2) else "") + s + (f"e{e}" if e else ""), e


def test_from_str_accuracy_bounds():
    """|parsed - exact| <= c (digits + |exp| + 2) u |exact|: each step of the loops (a
    multiply-add per digit, a division per fraction digit or exponent step) adds a few units
    of the type's rounding; a dropped term or misplaced power of ten is off by far more."""
    mpmath.mp.dps = 300
    rng = np.random.default_rng(11)
    for _ in range(300):
        s, e = random_literal(rng)
        exact = mpmath.mpf(s)
        steps = len(s) + abs(e) + 2
        d = O.dd_from_str(s)
        m = O.mf4_from_str(s)
        assert abs(dd_val(d) - exact) <= 4 * steps * mpmath.mpf(2) ** -104 * abs(exact), s
        assert abs(mf_val(m) - exact) <= 4 * steps * mpmath.mpf(2) ** -205 * abs(exact), s


def test_from_str_trim_sign_exponent():
    one_five = O.dd_from_str("1.5")
    for s in ["  1.5\t\n", " 1.5　", " 1.5 ", "+1.5", "15e-1", "1.5e0", "1.5e+0"]:
        v = O.dd_from_str(s)
        assert v is not None and np.array_equal(v, one_five), repr(s)
    v = O.dd_from_str("0.15e+1")  # 0.15 is not a dd; the accumulated rounding stays ~2^-104
    assert v[0] == 1.5 and abs(v[1]) < 2.0 ** -100
    for s in ["​1.5", "- 1.5", "1.5 x", "1e2147483648", "1e+-1", "1e"]:
        assert O.dd_from_str(s) is None, repr(s)
    assert np.array_equal(O.dd_from_str("-2.25"), -O.dd_from_str("2.25"))
    # i32::MIN: `0..(-exp)` wraps to an empty range in a release build, so nothing divides
    assert np.array_equal(O.dd_from_str("7e-2147483648"), np.array([7.0, 0.0]))
    assert O.dd_from_str("inf") is None and O.dd_from_str("nan") is None


def test_serialize_strings():
    rng = np.random.default_rng(5)
    for _ in range(200):
        hi = float(rng.standard_normal() * 10.0 ** rng.integers(-300, 300))
        lo = hi * 2.0 ** -53 * float(rng.uniform(-0.5, 0.5))
        assert O.dd_serialize_str([hi, lo]) == rust_exp(hi, 18, True) + "+" + rust_exp(lo, 18, True)
        assert O.dd_to_string([hi, lo]) == rust_exp(hi + lo, 32)
        limbs = [hi, lo, lo * 2.0 ** -53, -lo * 2.0 ** -106]
        assert O.mf4_serialize_str(limbs) == "+".join(rust_exp(v, 14, True) for v in limbs)
        assert O.mf4_to_string(limbs) == rust_exp(hi + lo, 64)


def test_serde_round_trips():
    rng = np.random.default_rng(9)
    for _ in range(300):
        hi = float(rng.standard_normal() * 10.0 ** rng.integers(-300, 300))
        lo = hi * 2.0 ** -53 * float(rng.uniform(-0.5, 0.5))
        v = np.array([hi, lo])
        assert np.array_equal(O.dd_deserialize_str(O.dd_serialize_str(v)), v)  # 19 digits: lossless
        limbs = [hi, lo, lo * 2.0 ** -53, -lo * 2.0 ** -106]
        got = O.mf4_deserialize_str(O.mf4_serialize_str(limbs))
        # 15 significant digits per limb: the reference's MF text form is lossy
        assert np.array_equal(got, [float(rust_exp(x, 14)) for x in limbs])
    for special in ([0.0, -0.0], [float("inf"), 0.0], [-float("inf"), -0.0]):
        got = O.dd_deserialize_str(O.dd_serialize_str(special))
        assert [struct.pack("<d", x) for x in got] == [struct.pack("<d", x) for x in special]
    nan = O.dd_deserialize_str(O.dd_serialize_str([float("nan"), 0.0]))  # "NaN++0.0…e0"
    assert nan[0] != nan[0] and nan[1] == 0.0


def test_serde_split_rules():
    assert np.array_equal(O.dd_deserialize_str("+1.0e5+-2.0e-12"), [1.0e5, -2.0e-12])
    assert np.array_equal(O.dd_deserialize_str("1.5++2"), [1.5, 2.0])
    assert np.array_equal(O.dd_deserialize_str("1e+5++1"), [1e5, 1.0])  # the '+' after 'e' is skipped
    for bad in ["1e+5", "+1.0e0+1.0e0", "++1", "1.0", "1.0+", "+1+-x", "1 ++2"]:
        assert O.dd_deserialize_str(bad) is None, bad
    assert np.array_equal(O.mf4_deserialize_str("1++2++3+-4"), [1.0, 2.0, 3.0, -4.0])
    assert np.array_equal(O.mf4_deserialize_str("1e+1++2E+2++3+-4e-1"), [10.0, 200.0, 3.0, -0.4])
    for bad in ["1++2++3", "1++2++3++4++5", "1++2++3++x", "", "1+2+3+4"]:
        assert O.mf4_deserialize_str(bad) is None, bad
