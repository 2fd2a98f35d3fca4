"""GPU text wire formats against the oracle: every formatted string byte-identical, every
parsed value bit-identical with the same accept/reject decision (Display / serde of f128
and MultiFloat<4>, dd.rs:331-514, multifloat/fmt.rs, multifloat/traits.rs:7-150; Rust's f64
`{:.Pe}` and parse).  At one million elements the GPU's own serialize -> deserialize round
trip must return every dd bit for bit."""
import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def qp():
    import quadprec_amd
    quadprec_amd.lib()
    return quadprec_amd


def random_doubles(rng, n):
    bits = rng.integers(0, 2**64 - 1, n, dtype=np.uint64, endpoint=True)
    x = bits.view(np.float64).copy()
    return x


SPECIAL = np.array([0.0, -0.0, 1.0, -1.0, 0.1, 1 / 3, 2.0 ** -1074, 2.0 ** -1022, 1.7976931348623157e308,
                    9.999999999999999e22, 1e23, 0.5, 2.5, 1234567890123445.0, 1234567890123455.0, 9.5, 99.5,
                    2.0 ** 53, 123456789.0, float("inf"), -float("inf"), float("nan"), -float("nan"), 1e-300,
                    999999999999999.9, 0.30000000000000004])


def dd_values(rng, n):
    hi = rng.standard_normal(n) * 10.0 ** rng.integers(-300, 300, n)
    lo = hi * 2.0 ** -53 * rng.uniform(-0.5, 0.5, n)
    v = np.stack([hi, lo], axis=1)
    v[:len(SPECIAL), 0] = SPECIAL
    v[:len(SPECIAL), 1] = 0.0
    return v


def mf4_values(rng, n):
    d = dd_values(rng, n)
    return np.stack([d[:, 0], d[:, 1], d[:, 1] * 2.0 ** -53, -d[:, 1] * 2.0 ** -106], axis=1)


@pytest.mark.parametrize("prec", [0, 1, 5, 14, 16, 17, 18, 25, 32, 64])
@pytest.mark.parametrize("plus", [False, True])
def test_format_f64_byte_identical(qp, prec, plus):
    rng = np.random.default_rng(prec * 2 + plus)
    x = np.concatenate([SPECIAL, -SPECIAL, random_doubles(rng, 3000), rng.standard_normal(500)])
    got = qp.format_f64(x, prec, plus)
    want = [O.format_exp(float(v), prec, plus) for v in x]
    bad = [(float(x[i]), got[i], want[i]) for i in range(len(x)) if got[i] != want[i]]
    print(f"format_f64 prec={prec} plus={plus}: {len(x)} values, {len(bad)} differ", flush=True)
    assert not bad, bad[:5]


@pytest.mark.parametrize("name,limbs,gen", [("dd_to_string", 2, dd_values), ("dd_serialize_str", 2, dd_values),
                                            ("mf4_to_string", 4, mf4_values), ("mf4_serialize_str", 4, mf4_values)])
def test_value_formats_byte_identical(qp, name, limbs, gen):
    rng = np.random.default_rng(len(name))
    v = gen(rng, 4000)
    got = getattr(qp, name)(v)
    want = [getattr(O, name)(v[i]) for i in range(len(v))]
    bad = [(i, got[i], want[i]) for i in range(len(v)) if got[i] != want[i]]
    print(f"{name}: {len(v)} values, {len(bad)} differ", flush=True)
    assert not bad, bad[:5]


def test_format_sizes_and_empty(qp):
    assert qp.dd_to_string(np.zeros((0, 2))) == []
    text, off = qp.dd_serialize_str(np.array([[1.0, 0.0]]), raw=True)
    assert text == b"+1.000000000000000000e0++0.000000000000000000e0" and list(off) == [0, 47]


def same(a, b):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    return np.array_equal(a.view(np.uint64), b.view(np.uint64)) or (
        np.array_equal(np.isnan(a), np.isnan(b)) and np.array_equal(a[~np.isnan(a)].view(np.uint64),
                                                                    b[~np.isnan(b)].view(np.uint64)))


def literals(rng, n):
    out = []
    for _ in range(n):
        kind = rng.integers(0, 4)
        if kind == 0:
            x = float(random_doubles(rng, 1)[0])
            out.append(O.format_exp(x, int(rng.integers(0, 30))))
        elif kind == 1:
            ip = "".join(rng.choice(list("0123456789"), rng.integers(0, 25)))
            fp = "".join(rng.choice(list("0123456789"), rng.integers(0, 30)))
            s = (ip or "0") + ("." + fp if fp else "")
            if rng.integers(0, 2):
                s += ("e" if rng.integers(0, 2) else "E") + str(int(rng.integers(-345, 320)))
            out.append(("-" if rng.integers(0, 2) else "") + s)
        elif kind == 2:
            out.append(str(float(rng.standard_normal() * 10.0 ** rng.integers(-20, 20))))
        else:
            out.append(str(rng.choice(["", "+", ".", "e5", "1e", "1e+", " 1", "1 ", "inf", "-Infinity", "nan",
                                       "NaN", "1.2.3", "0x1", "5.", ".5", "-0", "1e400", "1e-400",
                                       "2.4703282292062327e-324", "2.4703282292062328e-324",
                                       "1.7976931348623158e308", "00012.500e-0001"])))
    return out


def test_parse_f64_bit_identical(qp):
    rng = np.random.default_rng(3)
    strs = literals(rng, 20000)
    vals, status = qp.parse_f64(strs, strict=False)
    for i, s in enumerate(strs):
        want = O.parse_f64(s)
        if want is None:
            assert status[i] == 1, (s, status[i])
        else:
            assert status[i] == 0, (s, status[i])
            assert same(vals[i], want), (s, vals[i], want)
    print(f"parse_f64: {len(strs)} literals, {int((status == 1).sum())} rejected", flush=True)


def test_parse_f64_long_digit_strings(qp):
    """More than 40 significant digits: exact whenever the first 40 settle the rounding; a
    string built to sit on a halfway point gets status 2, never a wrong value."""
    rng = np.random.default_rng(4)
    strs = []
    for _ in range(2000):
        d = "".join(rng.choice(list("0123456789"), rng.integers(41, 120)))
        strs.append(d[0] + "." + d[1:] + "e" + str(int(rng.integers(-320, 300))))
    # the halfway point between 0 and the smallest subnormal, 2^-1075, written out in full
    strs.append("2.4703282292062327208828439643411068618252990130716238221279284125033775363510437593264991818081799618989828234772285886546332835517796989819938739800539093906315035659515570226392290858392449105184435931802849936536152500319370457678249219365623669863658480757001585769269903706311928279558551332927834338409351978015531246597263579574622766465272827220056374006485499977096599470454020828166226237857393450736339007967761930577506740176324673600968951340535537458516661134223766678604162159680461914467291840300530057530849048765391711386591646239524912623653881879636239373280423891018672348497668235089863388587925628302755995657524455507255189313690836254779186948667994968324049705821028513185451396213837722826145437693412532098591327667236328125e-324")
    vals, status = qp.parse_f64(strs, strict=False)
    for i, s in enumerate(strs):
        want = O.parse_f64(s)
        assert status[i] in (0, 2), s
        if status[i] == 0:
            assert same(vals[i], want), s
    assert status[-1] == 2
    print(f"long literals: {int((status[:-1] == 2).sum())} of {len(strs) - 1} random ones undecided", flush=True)
    assert (status[:-1] == 2).sum() == 0


def from_str_literals(rng, n):
    out = []
    for _ in range(n):
        nd = int(rng.integers(1, 50))
        digits = "".join(rng.choice(list("0123456789"), nd))
        cut = int(rng.integers(0, nd + 1))
        s = digits[:cut] + "." + digits[cut:] if rng.integers(0, 3) else digits
        if rng.integers(0, 2):
            s += ("e" if rng.integers(0, 2) else "E") + str(int(rng.integers(-400, 400)))
        s = str(rng.choice(["", "-", "+"])) + s
        s = str(rng.choice(["", " ", "\t", " ", "　 "])) + s + str(rng.choice(["", " ", "\n", " "]))
        out.append(s)
    out += ["", "abc", "1.2.3", "-", "1.0e", "1e2147483648", "7e-2147483648", "1e-5000", "1e99999", "-1e-99999",
            "​1", "1.5 x", "inf", ".", "+.5", "5.", "1e+5", "1e+-5", "3.14159265358979323846264338327950288"]
    return out


@pytest.mark.parametrize("name,limbs", [("dd_from_str", 2), ("mf4_from_str", 4)])
def test_from_str_bit_identical(qp, name, limbs):
    rng = np.random.default_rng(limbs)
    strs = from_str_literals(rng, 3000)
    vals, status = getattr(qp, name)(strs, strict=False)
    for i, s in enumerate(strs):
        want = getattr(O, name)(s)
        if want is None:
            assert status[i] == 1, repr(s)
        else:
            assert status[i] == 0, repr(s)
            assert same(vals[i], want), (repr(s), vals[i], want)
    print(f"{name}: {len(strs)} literals, {int((status == 1).sum())} rejected", flush=True)


@pytest.mark.parametrize("name,fmt,limbs,gen", [("dd_deserialize_str", "dd_serialize_str", 2, dd_values),
                                                ("mf4_deserialize_str", "mf4_serialize_str", 4, mf4_values)])
def test_deserialize_bit_identical(qp, name, fmt, limbs, gen):
    rng = np.random.default_rng(limbs + 10)
    v = gen(rng, 3000)
    strs = [getattr(O, fmt)(v[i]) for i in range(len(v))]
    strs += ["+1.0e5+-2.0e-12", "1.5++2", "1e+5++1", "1e+5", "+1.0e0+1.0e0", "++1", "1.0", "1.0+", "1++2++3+-4",
             "1e+1++2E+2++3+-4e-1", "1++2++3", "1++2++3++4++5", "1++2++3++x", "", "NaN++0.0e0", "+inf+-inf"]
    vals, status = getattr(qp, name)(strs, strict=False)
    for i, s in enumerate(strs):
        want = getattr(O, name)(s)
        if want is None:
            assert status[i] == 1, s
        else:
            assert status[i] == 0, s
            assert same(vals[i], want), (s, vals[i], want)


def test_strict_mode_raises(qp):
    with pytest.raises(qp.ParseError):
        qp.dd_from_str(["1.5", "1.2.3"])
    assert np.array_equal(qp.dd_from_str(["1.5"]), [[1.5, 0.0]])


def test_round_trip_one_million(qp):
    """serialize -> deserialize on the GPU returns every dd bit for bit (19 digits per f64)."""
    rng = np.random.default_rng(1)
    n = 1 << 20
    hi = rng.standard_normal(n) * 10.0 ** rng.integers(-300, 300, n)
    lo = hi * 2.0 ** -53 * rng.uniform(-0.5, 0.5, n)
    v = np.stack([hi, lo], axis=1)
    text, off = qp.dd_serialize_str(v, raw=True)
    back = qp.dd_deserialize_str((text, off))
    print(f"round trip: {n} dd, {len(text)} bytes of text", flush=True)
    assert np.array_equal(back.view(np.uint64), v.view(np.uint64))
    idx = rng.integers(0, n, 500)
    strs = qp.unpack_strings(text, off)
    for i in idx:
        assert strs[i] == O.dd_serialize_str(v[i])


def test_device_entry_points(qp):
    """hp_dev_format / hp_dev_parse on device-resident planes, text and offsets."""
    import ctypes as C

    import torch
    rng = np.random.default_rng(12)
    v = dd_values(rng, 5000)
    dev = torch.device("cuda:0")
    hi = torch.tensor(v[:, 0], device=dev)
    lo = torch.tensor(v[:, 1], device=dev)
    planes = (C.c_void_p * 4)(hi.data_ptr(), lo.data_ptr(), 0, 0)
    n = len(v)
    off = torch.zeros(n + 1, dtype=torch.int64, device=dev)
    text = torch.zeros(n * 53, dtype=torch.uint8, device=dev)
    need = C.c_size_t(0)
    torch.cuda.synchronize()
    rc = qp.lib().hp_dev_format(3, C.cast(planes, C.c_void_p), n, 18, 1, C.c_void_p(text.data_ptr()), n * 53,
                                C.c_void_p(off.data_ptr()), C.byref(need), None)
    assert rc == 0, qp.lib().hp_last_error()
    host_off = off.cpu().numpy().astype(np.uint64)
    host_text = bytes(text[:need.value].cpu().numpy())
    assert qp.unpack_strings(host_text, host_off) == qp.dd_serialize_str(v)
    out_hi = torch.empty(n, dtype=torch.float64, device=dev)
    out_lo = torch.empty(n, dtype=torch.float64, device=dev)
    oplanes = (C.c_void_p * 4)(out_hi.data_ptr(), out_lo.data_ptr(), 0, 0)
    status = torch.empty(n, dtype=torch.uint8, device=dev)
    bad = C.c_size_t(99)
    rc = qp.lib().hp_dev_parse(3, C.c_void_p(text.data_ptr()), C.c_void_p(off.data_ptr()), n,
                               C.cast(oplanes, C.c_void_p), C.c_void_p(status.data_ptr()), C.byref(bad), None)
    assert rc == 0 and bad.value == 0
    back = np.stack([out_hi.cpu().numpy(), out_lo.cpu().numpy()], axis=1)
    assert same(back, v)
