"""GPU dot / gemv (crates/quad/src/linalg.rs:7-40) and Neumaier sums
(crates/compensated-sum/src/lib.rs:6-66) against the oracle, bit for bit: both run the same
dd operations in the same order (chunk sums, then rayon's halving tree or the Neumaier sum),
so every result must be identical, across the chunk boundaries (1023/1024/1025,
4095/4096/4097), at ragged chunk counts (3, 5, 1029: two tree passes), on empty inputs and at
a size of millions of elements."""
import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def qp():
    import quadprec_amd
    quadprec_amd.lib()
    return quadprec_amd


def rand_dd(rng, n, spread=0.0):
    x = np.zeros((n, 2))
    x[:, 0] = rng.standard_normal(n) * 10 ** rng.uniform(-spread, spread, n)
    x[:, 1] = x[:, 0] * 2.0 ** -53 * rng.uniform(-0.5, 0.5, n)
    return x


@pytest.mark.parametrize("n", [0, 1, 7, 1023, 1024, 1025, 2048, 2555, 3072, 5000, 1 << 20, (1 << 20) + 5000,
                               3 << 20])
def test_dot_bit_exact(qp, n):
    rng = np.random.default_rng(n + 11)
    a, b = rand_dd(rng, n, 4), rand_dd(rng, n, 4)
    g = qp.dot(a, b)
    r = O.dot(a, b) if n else np.zeros(2)
    print(f"dot n={n}: gpu {g}, oracle {r}", flush=True)
    assert np.array_equal(g, r)


def test_dot_reference_case(qp):  # linalg.rs:368-374
    a = np.stack([np.arange(100.0), np.zeros(100)], axis=1)
    b = np.stack([np.ones(100), np.zeros(100)], axis=1)
    r = qp.dot(a, b)
    assert abs(r[0] + r[1] - 4950.0) < 1e-10


@pytest.mark.parametrize("rows,cols", [(1, 1), (13, 37), (64, 64), (65, 130), (300, 1025), (2, 5000)])
def test_gemv_bit_exact(qp, rows, cols):
    rng = np.random.default_rng(rows * 7 + cols)
    a, x = rand_dd(rng, rows * cols, 3), rand_dd(rng, cols, 3)
    g = qp.gemv(a, rows, cols, x)
    r = O.gemv(a, rows, cols, x)
    print(f"gemv {rows}x{cols}: identical {np.mean(np.all(g == r, axis=1)):.3f}", flush=True)
    assert np.array_equal(g, r)


def test_gemv_zero_cols(qp):
    y = qp.gemv(np.zeros((0, 2)), 3, 0, np.zeros((0, 2)))
    assert np.array_equal(y, np.zeros((3, 2)))


@pytest.mark.parametrize("n", [0, 1, 50, 4095, 4096, 4097, 9000, 100_000, 1 << 21])
@pytest.mark.parametrize("parallel", [False, True])
def test_neumaier_f64_bit_exact(qp, n, parallel):
    rng = np.random.default_rng(n + 3)
    v = rng.standard_normal(n) * 10 ** rng.uniform(-8, 8, n)
    g = qp.neumaier_sum_f64(v, parallel)
    r = O.neumaier_sum_f64(v, parallel) if n else 0.0
    print(f"neumaier_f64 n={n} par={parallel}: {g!r} vs {r!r}", flush=True)
    assert g == r or (g != g and r != r)


@pytest.mark.parametrize("n", [0, 1, 50, 4095, 4096, 4097, 9000, 100_000, 1 << 20])
@pytest.mark.parametrize("parallel", [False, True])
def test_neumaier_dd_bit_exact(qp, n, parallel):
    rng = np.random.default_rng(n + 5)
    v = rand_dd(rng, n, 8)
    g = qp.neumaier_sum(v, parallel)
    r = O.neumaier_sum(v, parallel) if n else np.zeros(2)
    print(f"neumaier_dd n={n} par={parallel}: {g} vs {r}", flush=True)
    assert np.array_equal(g, r)


def test_neumaier_reference_cases(qp):  # lib.rs:70-102
    v = np.full(10_000_000, 1e-16)
    v[0] = 1.0
    exact = 1.0 + (v.size - 1) * 1e-16
    naive = float(np.cumsum(v)[-1])
    for par in (False, True):  # lib.rs:70-86 runs the serial form on these 10^7 terms
        comp = qp.neumaier_sum_f64(v, par)
        assert abs(comp - exact) < abs(naive - exact) or abs(comp - exact) < 1e-15
    h = 1.0 / (np.arange(100_000) + 1.0)
    assert abs(qp.neumaier_sum_f64(h) - qp.neumaier_sum_f64(h, True)) < 1e-14
    d = np.stack([1.0 / np.arange(1, 1001), np.zeros(1000)], axis=1)
    r = qp.neumaier_sum(d)
    assert abs(r[0] + r[1] - 7.485470860550343) < 1e-12
