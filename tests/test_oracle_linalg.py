"""Pins the oracle's dd / MF4 linear algebra and the solver crate restatements
(crates/quad/src/linalg.rs, crates/quad/src/multifloat/linalg.rs, crates/solver/src/lib.rs).

Ported reference tests plus mpmath-exact small systems (tests/golden/small_linalg.json),
the README Hilbert table (README.md:189-194) and structural invariants.
"""
import math

import numpy as np
import pytest

import oracle_lib as O
from golden_util import dd_arr, exact_sum, frac, fx, load, rel_err


def dd_of(vals):
    a = np.zeros((len(vals), 2))
    a[:, 0] = vals
    return a


def val(x):
    return x[..., 0] + x[..., 1]


# ---------------------------------------------------------------- linalg.rs tests
def test_cholesky_3x3():  # linalg.rs:376-403
    a = dd_of([4, 2, 0, 2, 5, 1, 0, 1, 6])
    l = O.cholesky(a, 3)
    orig = [[4, 2, 0], [2, 5, 1], [0, 1, 6]]
    for i in range(3):
        for j in range(3):
            s = np.zeros(2)
            for k in range(min(i, j) + 1):
                s = O.dd_binop("add", s, O.dd_binop("mul", l[i * 3 + k], l[j * 3 + k]))
            assert abs(val(s) - orig[i][j]) < 1e-28


def test_cholesky_upper_triangle_untouched():  # linalg.rs:42-96 writes only the lower triangle
    a = dd_of([4, 2, 0, 2, 5, 1, 0, 1, 6])
    l = O.cholesky(a, 3)
    assert l[1, 0] == 2 and l[2, 0] == 0 and l[5, 0] == 1


def test_cholesky_error_pivot():  # linalg.rs:54-57 and blocked pivot offset 251-254
    a = dd_of([1, 2, 2, 1])
    with pytest.raises(O.CholeskyError) as e:
        O.cholesky(a, 2)
    assert e.value.pivot == 1 and e.value.value == pytest.approx(-3.0)
    n = 10
    a = np.zeros((n * n, 2))
    for i in range(n):
        a[i * n + i, 0] = 1.0
    a[7 * n + 7, 0] = -1.0
    with pytest.raises(O.CholeskyError) as e:
        O.cholesky_blocked(a, n, 4)
    assert e.value.pivot == 7


def test_gemm_identity_and_2x2():  # linalg.rs:422-447
    eye = dd_of([1, 0, 0, 0, 1, 0, 0, 0, 1])
    a = dd_of(list(range(1, 10)))
    assert np.array_equal(O.gemm(eye, a, 3, 3, 3), a)
    c = O.gemm(dd_of([1, 2, 3, 4]), dd_of([5, 6, 7, 8]), 2, 2, 2)
    assert list(val(c)) == [19, 22, 43, 50]


def test_gemm_atb_matches_transpose():  # linalg.rs:449-462
    a = dd_of([1, 2, 3, 4, 5, 6])
    b = dd_of(list(range(1, 13)))
    at = dd_of([1, 3, 5, 2, 4, 6])
    assert np.array_equal(O.gemm(at, b, 2, 4, 3), O.gemm_atb(a, b, 2, 4, 3))


def make_spd(n):  # linalg.rs:466-475
    a = np.zeros((n * n, 2))
    for i in range(n):
        for j in range(n):
            a[i * n + j, 0] = 1.0 / (abs(i - j) + 1.0)
        a[i * n + i] = O.dd_binop("add", a[i * n + i], np.array([float(n), 0.0]))
    return a


def test_blocked_matches_unblocked():  # linalg.rs:477-491
    n = 20
    a = make_spd(n)
    le = O.cholesky(a, n)
    lb = O.cholesky_blocked(a, n, 8)
    for i in range(n):
        for j in range(i + 1):
            assert abs(val(O.dd_binop("sub", le[i * n + j], lb[i * n + j]))) < 1e-28


def test_blocked_with_nb_ge_n_is_unblocked_bitwise():
    # with one block, cholesky_blocked runs `cholesky` on the whole matrix (linalg.rs:250-255)
    n = 23
    a = make_spd(n)
    assert np.array_equal(O.cholesky(a, n), O.cholesky_blocked(a, n, 64))


def test_solve_roundtrip():  # linalg.rs:543-563
    n = 4
    a = np.zeros((n * n, 2))
    for i in range(n):
        for j in range(n):
            a[i * n + j, 0] = 1.0 / (i + j + 1)
        a[i * n + i] = O.dd_binop("add", a[i * n + i], np.array([1.0, 0.0]))
    l = O.cholesky(a, n)
    b = dd_of([1.0] * n)
    x = O.solve_cholesky(l, b, n)
    ax = O.matvec(a, x, n)
    for i in range(n):
        assert abs(val(O.dd_binop("sub", ax[i], b[i]))) < 1e-28


# ---------------------------------------------------------------- mpmath-exact small systems
SL = load("small_linalg")


def test_gemm_vs_exact():
    g = SL["gemm"]
    m, n, k = g["m"], g["n"], g["k"]
    c = O.gemm(dd_arr(g["a"]), dd_arr(g["b"]), m, n, k)
    a, b = dd_arr(g["a"]), dd_arr(g["b"])
    for i in range(m):
        for j in range(n):
            scale = sum(abs(val(a[i * k + p])) * abs(val(b[p * n + j])) for p in range(k))
            ex = frac(g["c"][i * n + j])
            err = float(abs(exact_sum(*c[i * n + j]) - ex))
            assert err <= 2 * k * 2.0 ** -104 * scale, (i, j, err / scale)


def test_cholesky_and_solve_vs_exact():
    s = SL["spd"]
    n = s["n"]
    a, b = dd_arr(s["a"]), dd_arr(s["b"])
    for l in (O.cholesky(a, n), O.cholesky_blocked(a, n, 3)):
        for i in range(n):
            for j in range(i + 1):
                assert rel_err(l[i * n + j], s["l"][i * n + j]) < 1e-29 or abs(float(frac(s["l"][i * n + j]))) < 1e-25
        x = O.solve_cholesky(l, b, n)
        for i in range(n):
            assert rel_err(x[i], s["x"][i]) < 1e-28


def test_forward_backward_compose_solve_cholesky():
    s = SL["spd"]
    n = s["n"]
    a, b = dd_arr(s["a"]), dd_arr(s["b"])
    l = O.cholesky(a, n)
    y = O.forward_solve(l, b, n)
    assert np.array_equal(O.backward_solve(l, y, n), O.solve_cholesky(l, b, n))


# ---------------------------------------------------------------- MF4 linalg tests
def mf_of(vals):
    a = np.zeros((len(vals), 4))
    a[:, 0] = vals
    return a


def test_mf4_gemm_2x2_and_atb():  # multifloat/linalg.rs:208-234
    c = O.mf4_gemm(mf_of([1, 2, 3, 4]), mf_of([5, 6, 7, 8]), 2, 2, 2)
    assert list(c[:, 0] + c[:, 1]) == [19, 22, 43, 50]
    a = mf_of([1, 2, 3, 4, 5, 6])
    b = mf_of(list(range(1, 13)))
    at = mf_of([1, 3, 5, 2, 4, 6])
    assert np.array_equal(O.mf4_gemm(at, b, 2, 4, 3), O.mf4_gemm_atb(a, b, 2, 4, 3))


def test_mf4_cholesky_solve_simple():  # multifloat/tests.rs:617-639
    a = mf_of([4, 2, 2, 3])
    l = O.mf4_cholesky(a, 2)
    x = O.mf4_cholesky_solve(l, mf_of([1, 1]), 2)
    assert abs(float(exact_sum(*x[0]) - frac("0.125"))) < 1e-62
    assert abs(float(exact_sum(*x[1]) - frac("0.25"))) < 1e-62
    ax = O.mf4_gemv(a, x, 2, 2)
    for i in range(2):
        assert abs(float(exact_sum(*ax[i]) - 1)) < 1e-62


def test_mf4_solve_vs_exact():
    s = SL["spd"]
    n = s["n"]
    a = np.zeros((n * n, 4))
    a[:, :2] = dd_arr(s["a"])
    b = np.zeros((n, 4))
    b[:, :2] = dd_arr(s["b"])
    x = O.mf4_cholesky_solve(O.mf4_cholesky(a, n), b, n)
    for i in range(n):
        assert rel_err(x[i], s["x"][i]) < 1e-60


# ---------------------------------------------------------------- solver crate
def hilbert(n):
    return np.array([1.0 / (i + j + 1) for i in range(n) for j in range(n)])


def test_solve_well_conditioned():  # solver/src/lib.rs:271-285
    n = 10
    a = hilbert(n).reshape(n, n) + np.eye(n)
    sol, x, xf = O.solve_spd(a.ravel(), np.ones(n), n)
    assert sol.strategy == 0 and sol.residual_norm < 1e-12 and sol.error == 0


def test_solve_hilbert_13():  # lib.rs:287-297
    n = 13
    sol, x, xf = O.solve_spd(hilbert(n), np.ones(n), n)
    assert sol.strategy in (1, 2)
    assert sol.kappa > 1e12


def test_solve_hilbert_12_mixed():  # lib.rs:299-317
    n = 12
    a = hilbert(n)
    b = np.zeros(n)
    for i in range(n):
        for j in range(n):
            b[i] += a[i * n + j]
    sol, x, xf = O.solve_spd(a, b, n)
    assert sol.strategy == 1
    assert sol.residual_norm < 1e-25


def test_solve_hilbert_12_f128_input():  # lib.rs:319-340
    n = 12
    a128 = dd_of(list(hilbert(n)))
    b128 = np.zeros((n, 2))
    for i in range(n):
        for j in range(n):
            b128[i] = O.dd_binop("add", b128[i], O.dd_binop("mul", a128[i * n + j], np.array([1.0, 0.0])))
    sol, x, xf = O.solve_spd_f128(a128, b128, n)
    err = math.sqrt(sum(val(O.dd_binop("sub", x[i], np.array([1.0, 0.0]))) ** 2 for i in range(n)))
    assert err < 1e-15


def test_hilbert_readme_table():
    """README.md:189-194: kappa_est reproduces to the printed digits.  The printed
    ||x-1|| digits come from an earlier revision of dd.rs (eight plausible variants of
    add/div/sqrt were tried; none reproduces them), so they pin only the magnitude:
    ours must sit within a factor 40 and the residual at the dd floor (< 2e-31)."""
    for row in load("hilbert_readme"):
        n = row["n"]
        one = np.array([1.0, 0.0])
        a = np.array([O.dd_binop("div", one, np.array([float(i + j + 1), 0.0]))
                      for i in range(n) for j in range(n)])
        b = np.zeros((n, 2))
        for i in range(n):
            for j in range(n):
                b[i] = O.dd_binop("add", b[i], O.dd_binop("mul", a[i * n + j], one))
        sol, x, xf = O.solve_spd_f128(a, b, n)
        err = math.sqrt(sum(val(O.dd_binop("sub", x[i], one)) ** 2 for i in range(n)))
        assert f"{sol.kappa:.1e}" == f"{row['kappa']:.1e}"
        assert row["err"] / 40 < err < row["err"] * 40
        assert sol.residual_norm < 2e-31


def test_not_positive_definite_error():  # lib.rs:129, 251-265
    n = 3
    a = np.array([1.0, 2, 0, 2, 1, 0, 0, 0, 1])
    sol, x, xf = O.solve_spd(a, np.ones(n), n)
    assert sol.error == 1 and sol.pivot == 1
