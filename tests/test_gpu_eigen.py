"""Symmetric dd eigensolver, canonical orthogonalization and F C = S C eps on the GPU
(hp_jacobi_eigen, hp_canonical_orthogonalize, hp_solve_gen_eigen), against the oracle's
restatements of crates/quad/src/linalg.rs:299-362 and crates/solver/src/ortho.rs:61-248.

The GPU rotates n/2 disjoint pairs at a time, the reference one pair at a time, so the two
agree to dd rounding rather than bit for bit.  The bars, with u = 2^-104:
  * eigenvalues: |lambda - lambda_ref| <= 4 n u ||A||_F (both are backward stable);
  * eigenvectors: ||A v - lambda v|| <= 8 n u ||A||_F and |V^T V - I| <= 8 n u;
  * canonical orthogonalization: the same retained count, X^T S X = I within
    8 n u ||S||_F / lambda_min(kept), and X's columns equal the oracle's up to sign within
    8 n u ||S||_F / (gap lambda_min) for well separated eigenvalues;
  * the pencil: eigenvalues within 8 n u ||F'|| of the oracle's, F c = eps S c and C^T S C = I.
The reference's own tests (linalg.rs:495-541, ortho.rs:252-399) run with their bars."""
import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu
U = 2.0 ** -104


@pytest.fixture(scope="module")
def qp():
    import quadprec_amd
    quadprec_amd.lib()
    return quadprec_amd


@pytest.fixture(params=[0, 1], ids=["pair_steps", "block_jacobi"])
def path(qp, request):
    """the scalar round-robin form (the default below n = 768) or the block form, forced"""
    qp.eigen_set_block_min(2 ** 62 if request.param == 0 else 1)
    yield request.param
    qp.eigen_set_block_min(0)


def dd(x):
    x = np.asarray(x, dtype=np.float64).ravel()
    return np.stack([x, np.zeros_like(x)], axis=1)


def val(v):
    return v[..., 0] + v[..., 1]


def dd_diff(a, b):
    return np.abs(val(O.dd_binop("sub", a, b)))


def random_sym(n, seed):
    rng = np.random.default_rng(seed)
    b = rng.standard_normal((n, n))
    a = np.zeros((n, n, 2))
    a[:, :, 0] = (b + b.T) / 2
    a[:, :, 1] = a[:, :, 0] * 2.0 ** -53 * rng.uniform(-0.5, 0.5, (n, n))
    a = 0.5 * (a + a.transpose(1, 0, 2))
    return a.reshape(n * n, 2)


def frob(a, n):
    return float(np.sqrt((val(a) ** 2).sum()))


def check_eigenpairs(a, n, ev, vecs):
    """vecs: (n, n, 2), vecs[i] = eigenvector i.  Returns (max |A v - lambda v|,
    max |V V^T - I|), both in dd."""
    worst = 0.0
    for i in range(n):
        av = O.matvec(a, np.ascontiguousarray(vecs[i]), n)
        lv = O.dd_binop("mul", np.repeat(ev[i:i + 1], n, axis=0), vecs[i])
        worst = max(worst, dd_diff(av, lv).max())
    rows = vecs.reshape(n * n, 2)
    vvt = val(O.gemm(rows, transpose_dd(rows, n), n, n, n)).reshape(n, n)
    return worst, np.abs(vvt - np.eye(n)).max()


def transpose_dd(x, n):
    return np.ascontiguousarray(x.reshape(n, n, 2).transpose(1, 0, 2)).reshape(n * n, 2)


# ---------------------------------------------------------------- reference tests
def test_jacobi_2x2(qp):  # linalg.rs:495-509
    ev, vec = qp.jacobi_eigen(dd([2, 1, 1, 2]), 2, 100)
    e = np.sort(val(ev))
    assert abs(e[0] - 1.0) < 1e-28 and abs(e[1] - 3.0) < 1e-28
    d = O.dd_binop("add", O.dd_binop("mul", vec[0:1], vec[2:3]), O.dd_binop("mul", vec[1:2], vec[3:4]))
    assert abs(val(d)[0]) < 1e-28


def test_jacobi_3x3(qp, path):  # linalg.rs:511-530
    a = dd([4, 1, 0, 1, 3, 1, 0, 1, 2])
    ev, vec = qp.jacobi_eigen(a, 3, 200)
    for i in range(3):
        vi = vec[3 * i:3 * i + 3]
        av = O.matvec(a, vi, 3)
        lv = O.dd_binop("mul", np.repeat(ev[i:i + 1], 3, axis=0), vi)
        assert dd_diff(av, lv).max() < 1e-27
    assert abs(val(ev).sum() - 9.0) < 1e-25


def test_jacobi_identity(qp):  # linalg.rs:532-540
    ev, _ = qp.jacobi_eigen(dd(np.eye(4)), 4, 10)
    assert np.all(np.abs(val(ev) - 1.0) < 1e-30)


def test_orthogonalize_identity(qp):  # ortho.rs:252-270
    x, m = qp.canonical_orthogonalize(dd(np.eye(3)), 3, 1e-8)
    assert m == 3
    xtx = val(O.gemm_atb(x, x, 3, 3, 3)).reshape(3, 3)
    assert np.abs(xtx - np.eye(3)).max() < 1e-28


def test_orthogonalize_2x2_overlap(qp):  # ortho.rs:272-293
    s = dd([1.0, 0.5, 0.5, 1.0])
    x, m = qp.canonical_orthogonalize(s, 2, 1e-8)
    assert m == 2
    xtsx = val(O.gemm_atb(x, O.gemm(s, x, 2, m, 2), m, m, 2)).reshape(m, m)
    assert np.abs(xtsx - np.eye(m)).max() < 1e-26


def test_orthogonalize_removes_linear_dependency(qp):  # ortho.rs:295-302
    _, m = qp.canonical_orthogonalize(dd([1.0, 1.0, 1.0, 1.0]), 2, 1e-8)
    assert m == 1


def check_pencil(f, s, n, ev, c, m, bar):
    worst = 0.0
    cm = c.reshape(n, m, 2)
    for k in range(m):
        ck = np.ascontiguousarray(cm[:, k])
        fc = O.matvec(f, ck, n)
        esc = O.dd_binop("mul", np.repeat(ev[k:k + 1], n, axis=0), O.matvec(s, ck, n))
        worst = max(worst, dd_diff(fc, esc).max())
    assert worst < bar, worst
    return worst


def test_solve_gen_eigen_2x2(qp, path):  # ortho.rs:304-344
    f, s = dd([2.0, 1.0, 1.0, 3.0]), dd([1.0, 0.2, 0.2, 1.0])
    ev, c, m = qp.solve_gen_eigen(f, s, 2, 1e-8)
    assert m == 2
    check_pencil(f, s, 2, ev, c, m, 1e-25)
    assert val(ev)[0] < val(ev)[1]


def test_solve_gen_eigen_identity_overlap(qp):  # ortho.rs:346-398
    f = dd([3.0, 1.0, 0.0, 1.0, 2.0, 1.0, 0.0, 1.0, 4.0])
    ev, c, m = qp.solve_gen_eigen(f, dd(np.eye(3)), 3, 1e-8)
    assert m == 3
    check_pencil(f, dd(np.eye(3)), 3, ev, c, m, 1e-25)
    assert np.all(np.diff(val(ev)) >= 0)


# ---------------------------------------------------------------- against the oracle
@pytest.mark.parametrize("n,seed", [(1, 0), (2, 1), (7, 2), (40, 3), (65, 4), (128, 5)])
def test_eigen_matches_oracle(qp, path, n, seed):
    a = random_sym(n, seed)
    ev, vec = qp.jacobi_eigen(a, n, 60)
    ref, _ = O.jacobi_eigen(a, n, 60)
    nrm = frob(a, n)
    e_g = np.sort(val(ev))
    e_r = np.sort(val(ref))
    # compare in dd after sorting by value
    og, orf = np.argsort(val(ev)), np.argsort(val(ref))
    err = dd_diff(ev[og], ref[orf]).max()
    assert err <= 4 * n * U * nrm, (err, 4 * n * U * nrm)
    assert np.allclose(e_g, e_r, rtol=0, atol=1e-25 * nrm)
    worst, orth = check_eigenpairs(a, n, ev, vec.reshape(n, n, 2))
    print(f"n={n}: eigenvalues {err:.1e}, residual {worst:.1e}, orthogonality {orth:.1e} (||A||_F {nrm:.1f})")
    assert worst <= 8 * n * U * nrm
    assert orth <= 8 * n * U


def test_eigen_degenerate_and_diagonal(qp, path):
    n = 9
    rng = np.random.default_rng(7)
    q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    lam = np.array([1.0, 1.0, 1.0, 2.0, 2.0, 3.0, -1.0, 0.0, 0.0])
    a = dd(((q * lam) @ q.T + ((q * lam) @ q.T).T) / 2)
    ev, vec = qp.jacobi_eigen(a, n, 60)
    assert np.abs(np.sort(val(ev)) - np.sort(lam)).max() < 1e-13  # the input itself is only f64-exact
    worst, orth = check_eigenpairs(a, n, ev, vec.reshape(n, n, 2))
    assert worst <= 8 * n * U * frob(a, n) and orth <= 8 * n * U
    d = dd(np.diag(np.arange(5.0)))
    ev, vec = qp.jacobi_eigen(d, 5, 1)  # diagonal: converged before the first sweep
    assert np.array_equal(val(ev), np.arange(5.0)) and np.array_equal(val(vec).reshape(5, 5), np.eye(5))


def test_eigen_reports_no_convergence(qp, path):
    a = random_sym(6, 8)
    with pytest.raises(qp.NotConverged):
        qp.jacobi_eigen(a, 6, 1)  # one check, one sweep, no further check (linalg.rs:305-359)
    with pytest.raises(O.NotConverged):
        O.jacobi_eigen(a, 6, 1)


def overlap(qp, n, seed, lo, spacing=1.4):
    c, al = qp.gen_overlap_basis(seed, n, lo, 10.0, spacing)
    return O.overlap_s_dd(c, al)


# the last case (0.4 bohr lattice) has eigenvalues 1.4e-5, 1.6e-5 and 4.4e-5 below the
# threshold 5e-5: three directions are dropped
@pytest.mark.parametrize("n,seed,lo,thr,spacing,dropped", [(64, 0, 0.01, 1e-8, 1.4, 0), (125, 1, 0.05, 1e-6, 1.4, 0),
                                                           (33, 2, 0.3, 1e-8, 1.4, 0), (64, 0, 0.01, 5e-5, 0.4, 3)])
def test_canonical_orthogonalize_overlap(qp, path, n, seed, lo, thr, spacing, dropped):
    s = overlap(qp, n, seed, lo, spacing)
    x, m = qp.canonical_orthogonalize(s, n, thr)
    xr, mr = O.canonical_orthogonalize(s, n, thr)
    assert m == mr == n - dropped
    ev_all, _ = O.jacobi_eigen_maxpivot(s, n)
    lam = np.sort(val(ev_all))
    kept = lam[lam > thr]
    nrm = frob(s, n)
    xtsx = val(O.gemm_atb(x, O.gemm(s, x, n, m, n), m, m, n)).reshape(m, m)
    e_orth = np.abs(xtsx - np.eye(m)).max()
    bar = 8 * n * U * nrm / kept[0]
    print(f"n={n}: kept {m} of {n}, lambda_min kept {kept[0]:.2e}, |X^T S X - I| {e_orth:.1e} (bar {bar:.1e})")
    assert e_orth <= bar
    # columns against the oracle's, up to sign, where the eigenvalue is isolated: a unit
    # eigenvector moves by at most (backward error) / gap, and X's column is it / sqrt(lambda)
    xg, xo = x.reshape(n, m, 2), xr.reshape(n, m, 2)
    checked = 0
    for k in range(m):
        gap = min([abs(kept[k] - kept[j]) for j in range(m) if j != k] + [abs(kept[k] - thr)])
        if gap < 1e-3 * kept[k]:
            continue
        sign = 1.0 if val(xg[:, k]) @ val(xo[:, k]) > 0 else -1.0
        d = dd_diff(np.ascontiguousarray(xg[:, k]), O.dd_binop("mul_f64", np.ascontiguousarray(xo[:, k]),
                                                               np.array([sign, 0.0])))
        assert d.max() <= 16 * n * U * nrm / (gap * kept[k] ** 0.5), (k, d.max())
        checked += 1
    assert checked > m // 2


@pytest.mark.parametrize("n,seed", [(24, 3), (48, 4)])
def test_solve_gen_eigen_matches_oracle(qp, path, n, seed):
    s = overlap(qp, n, seed, 0.1)
    f = random_sym(n, 100 + seed)
    thr = 1e-8
    ev, c, m = qp.solve_gen_eigen(f, s, n, thr)
    evr, cr, mr = O.solve_gen_eigen(f, s, n, thr)
    assert m == mr
    x, _ = O.canonical_orthogonalize(s, n, thr)
    fp = O.gemm_atb(x, O.gemm(f, x, n, m, n), m, m, n)
    fn = frob(fp, m)
    err = dd_diff(ev, evr).max()
    print(f"n={n}: kept {m}, eigenvalues vs oracle {err:.1e} (||F'|| {fn:.1e})")
    assert err <= 8 * n * U * fn
    ev_all, _ = O.jacobi_eigen_maxpivot(s, n)
    lam_min = np.sort(val(ev_all))
    lam_min = lam_min[lam_min > thr][0]
    check_pencil(f, s, n, ev, c, m, 64 * n * U * fn * frob(s, n) / lam_min ** 0.5 + 1e-300)
    ctsc = val(O.gemm_atb(c, O.gemm(s, c, n, m, n), m, m, n)).reshape(m, m)
    assert np.abs(ctsc - np.eye(m)).max() <= 16 * n * U * frob(s, n) / lam_min
