"""GPU parity of the Cholesky factorisations, triangular solves, matvec, the solver
crate's SPD solves and the overlap build, against the oracle, bit for bit (the kernels
keep the reference's per-element operation order).  Inputs include the reference's own
test matrices (crates/quad/src/linalg.rs:376-563, crates/solver/src/lib.rs:271-351) and
the overlap workload of config 0 (N = 512 lattice Gaussians, alpha in [0.01, 10])."""
import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def qp():
    import quadprec_amd
    quadprec_amd.lib()
    return quadprec_amd


def bits(x):
    return np.ascontiguousarray(x).view(np.uint64)


def same(a, b):
    return np.array_equal(bits(a), bits(b))


def dd_of(vals):
    a = np.zeros((len(vals), 2))
    a[:, 0] = vals
    return a


def make_spd(n):  # linalg.rs:466-475
    a = np.zeros((n * n, 2))
    for i in range(n):
        for j in range(n):
            a[i * n + j, 0] = 1.0 / (abs(i - j) + 1.0)
        a[i * n + i] = O.dd_binop("add", a[i * n + i], np.array([float(n), 0.0]))
    return a


def overlap_matrix(qp, n, seed=0, lo=0.01, hi=10.0):
    c, a = qp.gen_overlap_basis(seed, n, lo, hi)
    return qp.overlap_s(c, a), c, a


def test_cholesky_3x3_and_upper_untouched(qp):  # linalg.rs:376-403
    a = dd_of([4, 2, 0, 2, 5, 1, 0, 1, 6])
    l = qp.cholesky(a, 3)
    assert same(l, O.cholesky(a, 3))
    assert l[1, 0] == 2 and l[2, 0] == 0 and l[5, 0] == 1


def test_cholesky_errors_match_reference(qp):
    a = dd_of([1, 2, 2, 1])
    with pytest.raises(qp.CholeskyError) as e:
        qp.cholesky(a, 2)
    with pytest.raises(O.CholeskyError) as r:
        O.cholesky(a, 2)
    assert (e.value.pivot, e.value.value) == (r.value.pivot, r.value.value)
    n = 150
    a = np.zeros((n * n, 2))
    for i in range(n):
        a[i * n + i, 0] = 1.0
    a[97 * n + 97, 0] = -2.0
    for nb in (8, 64, 0, 100):
        with pytest.raises(qp.CholeskyError) as e:
            qp.cholesky_blocked(a, n, nb)
        assert e.value.pivot == 97 and e.value.value == -2.0


@pytest.mark.parametrize("n", [1, 5, 63, 64, 65, 130, 200])
def test_cholesky_unblocked_bitexact(qp, n):
    a = make_spd(n)
    assert same(qp.cholesky(a, n), O.cholesky(a, n))


@pytest.mark.parametrize("n,nb", [(20, 8), (130, 16), (200, 64), (200, 0), (257, 100), (90, 128), (70, 70)])
def test_cholesky_blocked_bitexact(qp, n, nb):
    a = make_spd(n)
    assert same(qp.cholesky_blocked(a, n, nb), O.cholesky_blocked(a, n, nb))


def test_overlap_bitexact_and_structure(qp):
    n = 300
    s, c, al = overlap_matrix(qp, n, seed=3)
    ref = O.overlap_s_dd(c, al)
    assert same(s, ref)
    s3 = s.reshape(n, n, 2)
    assert np.array_equal(s3, s3.transpose(1, 0, 2))
    assert all(list(s3[i, i]) == [1.0, 0.0] for i in range(n))


@pytest.mark.parametrize("n", [7, 64, 200, 513])
def test_solves_and_matvec_bitexact(qp, n):
    s, _, _ = overlap_matrix(qp, n, seed=n, lo=0.5, hi=10.0)
    l = O.cholesky(s, n)
    b = qp.gen_dd_normal(5, n)
    assert same(qp.forward_solve(l, b, n), O.forward_solve(l, b, n))
    assert same(qp.backward_solve(l, b, n), O.backward_solve(l, b, n))
    assert same(qp.solve_cholesky(l, b, n), O.solve_cholesky(l, b, n))
    assert same(qp.matvec(s, b, n), O.matvec(s, b, n))
    assert qp.cond_estimate(l, n) == O.cond_estimate(l, n)


def test_config0_overlap_blocked_cholesky_solve_bitexact(qp):
    """Config 0: N=512 lattice Gaussians (seed 0, alpha in [0.01, 10]), cholesky_blocked,
    then solve_cholesky with b = ones -- every bit equal to the oracle's."""
    n = 512
    s, c, al = overlap_matrix(qp, n, seed=0)
    ref_s = O.overlap_s_dd(c, al)
    assert same(s, ref_s)
    l = qp.cholesky_blocked(s, n, 64)
    ref_l = O.cholesky_blocked(ref_s, n, 64)
    assert same(l, ref_l)
    ones = dd_of([1.0] * n)
    assert same(qp.solve_cholesky(l, ones, n), O.solve_cholesky(ref_l, ones, n))


def hilbert(n):
    return np.array([1.0 / (i + j + 1) for i in range(n) for j in range(n)])


def check_solution(sol, ref):
    rs, rx, rxf = ref
    assert sol.strategy == rs.strategy
    assert sol.refinement_iters == rs.refinement_iters
    assert sol.kappa == rs.kappa
    assert sol.residual_norm == rs.residual_norm
    assert same(sol.x, rx) and same(sol.x_f64, rxf)


def test_solve_spd_strategies_bitexact(qp):  # solver/src/lib.rs:271-340
    n = 10
    a = (hilbert(n).reshape(n, n) + np.eye(n)).ravel()
    sol = qp.solve_spd(a, np.ones(n), n)
    assert sol.strategy == qp.Strategy.Float64 and sol.residual_norm < 1e-12
    check_solution(sol, O.solve_spd(a, np.ones(n), n))
    n = 13
    sol = qp.solve_spd(hilbert(n), np.ones(n), n)
    assert sol.strategy in (qp.Strategy.Float128, qp.Strategy.MixedPrecision) and sol.kappa > 1e12
    check_solution(sol, O.solve_spd(hilbert(n), np.ones(n), n))
    n = 12
    a = hilbert(n)
    b = a.reshape(n, n).sum(axis=1)
    b = np.array([sum(a[i * n + j] for j in range(n)) for i in range(n)])
    sol = qp.solve_spd(a, b, n)
    assert sol.strategy == qp.Strategy.MixedPrecision and sol.residual_norm < 1e-25
    check_solution(sol, O.solve_spd(a, b, n))


def test_solve_spd_f128_hilbert_readme(qp):  # demo/src/main.rs:54-102, README.md:189-194
    for n in (12, 15, 20):
        one = np.array([1.0, 0.0])
        a = np.array([O.dd_binop("div", one, np.array([float(i + j + 1), 0.0])) for i in range(n) for j in range(n)])
        b = O.matvec(a, np.tile(one, (n, 1)), n)
        sol = qp.solve_spd_f128(a, b, n)
        assert sol.strategy == qp.Strategy.Float128
        check_solution(sol, O.solve_spd_f128(a, b, n))
        assert sol.residual_norm < 2e-31


def test_solve_spd_overlap_and_not_spd(qp):
    n = 150
    s, _, _ = overlap_matrix(qp, n, seed=9, lo=0.05, hi=10.0)
    a = s[:, 0] + s[:, 1]
    b = qp.gen_dd_normal(2, n)[:, 0]
    sol = qp.solve_spd(a, b, n)
    check_solution(sol, O.solve_spd(a, b, n))
    bad = np.array([1.0, 2, 0, 2, 1, 0, 0, 0, 1])
    with pytest.raises(qp.SolveError) as e:
        qp.solve_spd(bad, np.ones(3), 3)
    assert e.value.pivot == 1
