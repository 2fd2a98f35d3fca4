"""The oracle's reductions: dot and gemv (crates/quad/src/linalg.rs:7-40) and the Neumaier
compensated sums (crates/compensated-sum/src/lib.rs:6-66), pinned by the reference's own
tests (linalg.rs:368-374, lib.rs:70-102) and by exact rational arithmetic: every double
and every dd is a rational, so Python's Fraction gives the exact sums and dot products,
against which the algorithms' error bounds are checked (dd dot: 2 n 2^-104 sum |a_i b_i|;
Neumaier f64: 2u |S| + 4 n u^2 sum |v|; dd: 16u |S| + 8 n u^2 sum |v|).  dot's chunked branch is also
checked to be rayon's halving tree over its chunk sums (split at len / 2 down to single
chunks, each leaf ZERO + p_i, restated here in Python over the oracle's dd add), and a
planted mistake (the compensation branch swapped) fails the Neumaier bound."""
from fractions import Fraction

import numpy as np
import pytest

import oracle_lib as O

U = 2.0 ** -104


def dd_frac(v):
    return Fraction(float(v[0])) + Fraction(float(v[1]))


def rand_dd(rng, n, scale=1.0):
    x = np.zeros((n, 2))
    x[:, 0] = rng.standard_normal(n) * scale
    x[:, 1] = x[:, 0] * 2.0 ** -53 * rng.uniform(-0.5, 0.5, n)
    return x


def test_dot_basic():  # linalg.rs:368-374
    a = np.stack([np.arange(100.0), np.zeros(100)], axis=1)
    b = np.stack([np.ones(100), np.zeros(100)], axis=1)
    r = O.dot(a, b)
    assert abs(r[0] + r[1] - 4950.0) < 1e-10


@pytest.mark.parametrize("n", [1, 7, 1023, 1024, 1025, 5000])
def test_dot_error_bound(n):
    rng = np.random.default_rng(n)
    a, b = rand_dd(rng, n), rand_dd(rng, n)
    exact = sum(dd_frac(a[i]) * dd_frac(b[i]) for i in range(n))
    absum = sum(abs(dd_frac(a[i]) * dd_frac(b[i])) for i in range(n))
    r = O.dot(a, b)
    assert abs(dd_frac(r) - exact) <= 2 * n * U * absum
    if n >= 1024:  # the chunk sums, combined by the halving tree
        assert np.array_equal(r, halving_tree(chunk_sums(a, b)))


def chunk_sums(a, b):
    return [O.dot(a[c0:c0 + 1024], b[c0:c0 + 1024]).reshape(1, 2) for c0 in range(0, len(a), 1024)]


def halving_tree(p, lo=0, ln=None):
    """rayon's reduce split at len // 2 down to single items: leaf ZERO + p, node left + right"""
    ln = len(p) if ln is None else ln
    if ln == 1:
        return O.dd_binop("add", np.zeros((1, 2)), p[lo])[0]
    h = ln // 2
    return O.dd_binop("add", halving_tree(p, lo, h).reshape(1, 2), halving_tree(p, lo + h, ln - h).reshape(1, 2))[0]


@pytest.mark.parametrize("m", [2, 3, 5, 6, 7, 11, 13, 64, 100])
def test_dot_tree_shape(m):
    """Ragged chunk counts: the tree differs from both the left fold and the level-by-level
    pairing for some m (e.g. 3: p0 + (p1 + p2)); the oracle must take the halving tree."""
    rng = np.random.default_rng(100 + m)
    n = 1024 * m - (517 if m > 2 else 0)
    a, b = rand_dd(rng, n, 1e3), rand_dd(rng, n)
    b[::3] *= -1.0
    r = O.dot(a, b)
    p = chunk_sums(a, b)
    assert np.array_equal(r, halving_tree(p))
    if m == 3:  # p0 + (p1 + p2), not (p0 + p1) + p2
        right = O.dd_binop("add", p[1], p[2])
        assert np.array_equal(r, O.dd_binop("add", p[0], right)[0])
    if m in (3, 5, 11, 64, 100):  # inputs on which the left fold gives other bits
        acc = np.zeros((1, 2))
        for q in p:
            acc = O.dd_binop("add", acc, q)
        assert not np.array_equal(r, acc[0])


def test_gemv_rows_are_serial_dots():
    rng = np.random.default_rng(3)
    rows, cols = 13, 37
    a, x = rand_dd(rng, rows * cols), rand_dd(rng, cols)
    y = O.gemv(a, rows, cols, x)
    for i in range(rows):
        row = a[i * cols:(i + 1) * cols]
        exact = sum(dd_frac(row[j]) * dd_frac(x[j]) for j in range(cols))
        absum = sum(abs(dd_frac(row[j]) * dd_frac(x[j])) for j in range(cols))
        assert abs(dd_frac(y[i]) - exact) <= 2 * cols * U * absum
        assert np.array_equal(y[i], O.dot(row, x))  # below 1024: dot_serial


def test_neumaier_f64_precision():  # lib.rs:70-86 (10^6 terms here instead of 10^7)
    n = 1_000_000
    v = np.full(n, 1e-16)
    v[0] = 1.0
    naive = float(np.cumsum(v)[-1])
    comp = O.neumaier_sum_f64(v)
    exact = 1.0 + (n - 1) * 1e-16
    assert abs(comp - exact) < abs(naive - exact) or abs(comp - exact) < 1e-15


def test_par_sum_matches_serial():  # lib.rs:88-95
    v = 1.0 / (np.arange(100_000) + 1.0)
    assert abs(O.neumaier_sum_f64(v) - O.neumaier_sum_f64(v, parallel=True)) < 1e-14


def test_neumaier_f128_basic():  # lib.rs:97-102
    v = np.stack([1.0 / np.arange(1, 1001), np.zeros(1000)], axis=1)
    r = O.neumaier_sum(v)
    assert abs(r[0] + r[1] - 7.485470860550343) < 1e-12


@pytest.mark.parametrize("n,parallel", [(50, False), (9000, False), (9000, True), (20000, True)])
def test_neumaier_bounds(n, parallel):
    rng = np.random.default_rng(n)
    v = rng.standard_normal(n) * 10 ** rng.uniform(-8, 8, n)  # wide dynamic range, cancellation
    exact = sum(Fraction(float(x)) for x in v)
    absum = sum(abs(Fraction(float(x))) for x in v)
    u = Fraction(2) ** -53
    s = O.neumaier_sum_f64(v, parallel)
    assert abs(Fraction(s) - exact) <= 2 * u * abs(exact) + 4 * n * u * u * absum
    vd = rand_dd(rng, n, 1.0) * 10 ** rng.uniform(-8, 8, (n, 1))
    exact_d = sum(dd_frac(x) for x in vd)
    absum_d = sum(abs(dd_frac(x)) for x in vd)
    sd = O.neumaier_sum(vd, parallel)
    # dd additions are not error-free, so the compensation carries dd rounding of its own:
    # a few more units of |S| than the f64 bound (a lost compensation costs ~u sum |v|)
    ud = Fraction(2) ** -104
    assert abs(dd_frac(sd) - exact_d) <= 16 * ud * abs(exact_d) + 8 * n * ud * ud * absum_d


def test_swapped_compensation_is_caught():
    """The Neumaier branch taken the wrong way round (the plain Kahan update when |v| > |sum|)
    loses the low part of large terms; the exact-sum bound rejects it."""
    v = np.array([1.0, 1e100, 1.0, -1e100] * 50)
    exact = Fraction(100)
    good = O.neumaier_sum_f64(v)
    assert Fraction(good) == exact

    def swapped(vals):
        s = c = 0.0
        for x in vals:
            t = s + x
            c += (s - t) + x if abs(s) < abs(x) else (x - t) + s
            s = t
        return s + c
    assert abs(Fraction(swapped(v)) - exact) > 1
