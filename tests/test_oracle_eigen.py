"""The oracle's eigensolvers: jacobi_eigen (crates/quad/src/linalg.rs:299-362, cyclic),
the solver crate's max-pivot Jacobi, canonical_orthogonalize and solve_gen_eigen
(crates/solver/src/ortho.rs:61-248), pinned three ways:

  * the reference's own tests (linalg.rs:495-541, ortho.rs:252-399), ported with their
    bars;
  * mpmath at 60 digits: eigenvalues and eigenvectors by mp.eigsy, and the pencil (F, S)
    by mp.eigsy of L^-1 F L^-T with S = L L^T from mp.cholesky (no formula of the oracle
    reused);
  * planted mistakes in the oracle's rotation (the sine's sign in the eigenvector update,
    tau from a_pp - a_qq) fail those checks by many orders of magnitude."""
import numpy as np
import pytest

import oracle_lib as O

mp = pytest.importorskip("mpmath")


def dd(x):
    x = np.asarray(x, dtype=np.float64).ravel()
    return np.stack([x, np.zeros_like(x)], axis=1)


def val(v):
    return v[..., 0] + v[..., 1]


def dd_mp(v):
    return mp.mpf(v[0]) + mp.mpf(v[1])


# ---------------------------------------------------------------- reference tests
def test_jacobi_2x2():  # linalg.rs:495-509
    ev, vec = O.jacobi_eigen(dd([2, 1, 1, 2]), 2, 100)
    e = np.sort(val(ev))
    assert abs(e[0] - 1.0) < 1e-28 and abs(e[1] - 3.0) < 1e-28
    d = O.dd_binop("add", O.dd_binop("mul", vec[0:1], vec[2:3]), O.dd_binop("mul", vec[1:2], vec[3:4]))
    assert abs(val(d)[0]) < 1e-28


def test_jacobi_3x3():  # linalg.rs:511-530
    a = dd([4, 1, 0, 1, 3, 1, 0, 1, 2])
    ev, vec = O.jacobi_eigen(a, 3, 200)
    for i in range(3):
        vi = vec[3 * i:3 * i + 3]
        av = O.matvec(a, vi, 3)
        lv = O.dd_binop("mul", np.repeat(ev[i:i + 1], 3, axis=0), vi)
        err = np.abs(val(O.dd_binop("sub", av, lv)))
        assert err.max() < 1e-27
    assert abs(val(ev).sum() - 9.0) < 1e-25


def test_jacobi_identity():  # linalg.rs:532-540
    n = 4
    ev, _ = O.jacobi_eigen(dd(np.eye(n)), n, 10)
    assert np.all(np.abs(val(ev) - 1.0) < 1e-30)


def _gemm_atb_val(x, y, n, m, p):
    return val(O.gemm_atb(x, y, m, p, n)).reshape(m, p)


def test_orthogonalize_identity():  # ortho.rs:252-270
    n = 3
    x, m = O.canonical_orthogonalize(dd(np.eye(n)), n, 1e-8)
    assert m == n
    xtx = _gemm_atb_val(x, x, n, n, n)
    assert np.abs(xtx - np.eye(n)).max() < 1e-28


def test_orthogonalize_2x2_overlap():  # ortho.rs:272-293
    n = 2
    s = dd([1.0, 0.5, 0.5, 1.0])
    x, m = O.canonical_orthogonalize(s, n, 1e-8)
    assert m == 2
    tmp = O.gemm(s, x, n, m, n)
    xtsx = _gemm_atb_val(x, tmp, n, m, m)
    assert np.abs(xtsx - np.eye(m)).max() < 1e-26


def test_orthogonalize_removes_linear_dependency():  # ortho.rs:295-302
    _, m = O.canonical_orthogonalize(dd([1.0, 1.0, 1.0, 1.0]), 2, 1e-8)
    assert m == 1


def _check_pencil(f, s, n, ev, c, m, bar):
    for k in range(m):
        ck = c.reshape(n, m, 2)[:, k]
        fc = O.matvec(f, ck, n)
        sc = O.matvec(s, ck, n)
        esc = O.dd_binop("mul", np.repeat(ev[k:k + 1], n, axis=0), sc)
        assert np.abs(val(O.dd_binop("sub", fc, esc))).max() < bar


def test_solve_gen_eigen_2x2():  # ortho.rs:304-344
    f, s = dd([2.0, 1.0, 1.0, 3.0]), dd([1.0, 0.2, 0.2, 1.0])
    ev, c, m = O.solve_gen_eigen(f, s, 2, 1e-8)
    assert m == 2
    _check_pencil(f, s, 2, ev, c, m, 1e-25)
    assert val(ev)[0] < val(ev)[1]


def test_solve_gen_eigen_identity_overlap():  # ortho.rs:346-398
    n = 3
    f = dd([3.0, 1.0, 0.0, 1.0, 2.0, 1.0, 0.0, 1.0, 4.0])
    ev, c, m = O.solve_gen_eigen(f, dd(np.eye(n)), n, 1e-8)
    assert m == 3
    _check_pencil(f, dd(np.eye(n)), n, ev, c, m, 1e-25)
    assert np.all(np.diff(val(ev)) >= 0)


# ---------------------------------------------------------------- mpmath
def random_sym(n, seed, spread=1.0):
    rng = np.random.default_rng(seed)
    q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    lam = np.linspace(-spread, 2 * spread, n) + 0.01 * rng.standard_normal(n)
    a = (q * lam) @ q.T
    a = 0.5 * (a + a.T)
    out = np.zeros((n * n, 2))
    out[:, 0] = a.ravel()
    out[:, 1] = a.ravel() * 2.0 ** -53 * rng.uniform(-0.5, 0.5, n * n)
    out = out.reshape(n, n, 2)
    out = 0.5 * (out + out.transpose(1, 0, 2))  # exact symmetry of both limbs
    return out.reshape(n * n, 2)


def mp_eig(a, n):
    A = mp.matrix(n, n)
    for i in range(n):
        for j in range(n):
            A[i, j] = dd_mp(a[i * n + j])
    E, Q = mp.eigsy(A)
    return [E[i] for i in range(n)], Q


@pytest.mark.parametrize("solver", ["cyclic", "maxpivot"])
@pytest.mark.parametrize("n,seed", [(6, 1), (11, 2)])
def test_eigenpairs_match_mpmath(solver, n, seed):
    mp.mp.dps = 60
    a = random_sym(n, seed)
    if solver == "cyclic":
        ev, vec = O.jacobi_eigen(a, n, 100)
        vecs = vec.reshape(n, n, 2)                  # row i = eigenvector i
    else:
        ev, vec = O.jacobi_eigen_maxpivot(a, n)
        vecs = vec.reshape(n, n, 2).transpose(1, 0, 2)  # column j = eigenvector j
    E, Q = mp_eig(a, n)
    scale = max(abs(e) for e in E)
    order = np.argsort(val(ev))
    for k, i in enumerate(order):
        assert abs(dd_mp(ev[i]) - E[k]) <= 1e-30 * scale
        v = [dd_mp(vecs[i, r]) for r in range(n)]
        dot = sum(v[r] * Q[r, k] for r in range(n))
        sign = 1 if dot > 0 else -1
        err = max(abs(v[r] - sign * Q[r, k]) for r in range(n))
        assert err <= 1e-29, (k, float(err))


def test_pencil_matches_mpmath():
    mp.mp.dps = 60
    n = 9
    f = random_sym(n, 5)
    rng = np.random.default_rng(6)
    b = rng.standard_normal((n, n))
    s = dd(b @ b.T / n + np.eye(n))
    ev, c, m = O.solve_gen_eigen(f, s, n, 1e-8)
    assert m == n
    F = mp.matrix(n, n)
    S = mp.matrix(n, n)
    for i in range(n):
        for j in range(n):
            F[i, j] = dd_mp(f[i * n + j])
            S[i, j] = dd_mp(s[i * n + j])
    L = mp.cholesky(S)
    Li = mp.inverse(L)
    E, _ = mp.eigsy(Li * F * Li.T)
    E = sorted(E[i] for i in range(n))
    scale = max(abs(e) for e in E)
    for k in range(n):
        assert abs(dd_mp(ev[k]) - E[k]) <= 1e-29 * scale, k
    _check_pencil(f, s, n, ev, c, m, 1e-28)
    # C^T S C = I
    sc = O.gemm(s, c, n, m, n)
    ctsc = _gemm_atb_val(c, sc, n, m, m)
    assert np.abs(ctsc - np.eye(m)).max() < 1e-28


# ---------------------------------------------------------------- planted mistakes
def test_planted_mistakes_are_caught():
    n = 8
    a = random_sym(n, 3)
    mp.mp.dps = 60
    E, _ = mp_eig(a, n)
    ev, vec = O.jacobi_eigen(a, n, 100, mistake=1)   # eigenvalues fine, eigenvectors not
    vecs = vec.reshape(n, n, 2)
    worst = 0.0
    for i in range(n):
        av = O.matvec(a, vecs[i], n)
        lv = O.dd_binop("mul", np.repeat(ev[i:i + 1], n, axis=0), vecs[i])
        worst = max(worst, np.abs(val(O.dd_binop("sub", av, lv))).max())
    assert worst > 1e-3
    # tau of the wrong sign rotates the wrong way; a_pq is still set to zero, so the sweeps
    # end, with eigenvalues far from the true ones
    try:
        ev2, _ = O.jacobi_eigen(a, n, 100, mistake=2)
    except O.NotConverged:
        return
    e2 = sorted(dd_mp(v) for v in ev2)
    assert max(abs(e2[k] - E[k]) for k in range(n)) > 1e-6
