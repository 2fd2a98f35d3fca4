"""The oracle's adjoint of the overlap build: gradients of log det S and b^T S^-1 b over
the 3n centre coordinates and n exponents (oracle hpo_overlap_adjoint_dd, assembled from
the reference's dd operations and linalg), pinned three independent ways:

  * mpmath at 60 digits: S from the closed form, log det and the quadratic form by
    mpmath's own LU, derivatives by mpmath.diff (no formula of the adjoint reused);
  * the oracle's f256 central differences (rank-2 determinant / Woodbury updates of the
    MultiFloat<4> factor), whose step error is bounded by comparing steps h and 2h;
  * the committed N = 2048 fixture: its dd reference gradient against its own f256
    differences.
Mistakes planted in the oracle's pair factors (an exponent swapped, the 3/4 term dropped,
a sign flipped) must fail the f256 check by six orders of magnitude."""
import numpy as np
import pytest

import golden_util as G
import oracle_lib as O

mp = pytest.importorskip("mpmath")
U = 2.0 ** -104


def mp_f(cent, alp, b):
    n = len(alp)
    S = mp.matrix(n, n)
    for i in range(n):
        for j in range(n):
            a, c = alp[i], alp[j]
            g = a + c
            r2 = sum((cent[i][k] - cent[j][k]) ** 2 for k in range(3))
            S[i, j] = (4 * a * c / g ** 2) ** mp.mpf(0.75) * mp.exp(-a * c / g * r2)
    x = mp.lu_solve(S, b)
    return mp.log(mp.det(S)), sum(b[i] * x[i] for i in range(n))


def dd_val(v):
    return mp.mpf(v[0]) + mp.mpf(v[1])


def test_gradients_match_mpmath_n4():
    mp.mp.dps = 60
    rng = np.random.default_rng(3)
    n = 4
    centers = rng.uniform(-1.2, 1.2, (n, 3)).ravel()
    alpha = 10 ** rng.uniform(-1, 0.5, n)
    b = np.zeros((n, 2))
    b[:, 0] = rng.standard_normal(n)
    ld, q, gl, gq, _ = O.overlap_adjoint_dd(centers, alpha, b)
    cent = [[mp.mpf(centers[3 * i + c]) for c in range(3)] for i in range(n)]
    alp = [mp.mpf(v) for v in alpha]
    bm = [mp.mpf(v) for v in b[:, 0]]
    L0, Q0 = mp_f(cent, alp, bm)
    assert abs(dd_val(ld) - L0) <= 1e-30 * abs(L0)
    assert abs(dd_val(q) - Q0) <= 1e-30 * abs(Q0)
    ex_l, ex_q = [], []
    for k in range(4 * n):
        def f(t, which):
            c2 = [row[:] for row in cent]
            a2 = alp[:]
            if k < 3 * n:
                c2[k // 3][k % 3] = t
            else:
                a2[k - 3 * n] = t
            return mp_f(c2, a2, bm)[which]
        t0 = cent[k // 3][k % 3] if k < 3 * n else alp[k - 3 * n]
        ex_l.append(mp.diff(lambda t: f(t, 0), t0))
        ex_q.append(mp.diff(lambda t: f(t, 1), t0))
    for got, ex in ((gl, ex_l), (gq, ex_q)):
        scale = max(abs(v) for v in ex)
        err = max(abs(dd_val(got[k]) - ex[k]) for k in range(4 * n))
        assert err <= 1e-29 * scale, float(err / scale)


def fd_value(x):
    return sum(mp.mpf(v) for v in x)


def fd_tolerance(fd2, kappa, gmax):
    """|g_fd(h) - g| bound: the step error (~ |g(2h) - g(h)| / 3, taken x2) plus what the dd
    gradient itself may carry (4 kappa 2^-104 of the gradient scale, floor 1e-30)."""
    return 2 * abs(fd2[0] - fd2[1]) + max(4 * kappa * U, 1e-30) * gmax


@pytest.mark.parametrize("n,lo", [(24, 0.01), (40, 0.1)])
def test_f256_differences_match_dd_gradient(n, lo):
    import quadprec_amd as qp
    mp.mp.dps = 80
    c, a = qp.gen_overlap_basis(9, n, lo, 10.0)
    b = qp.gen_dd_normal(19, n)
    ld, q, gl, gq, _ = O.overlap_adjoint_dd(c, a, b)
    coords = np.arange(4 * n)
    fl, fq = O.overlap_fd_f256(c, a, b, coords)
    s = O.overlap_s_dd(c, a)
    kap = np.linalg.cond(s[:, 0].reshape(n, n))
    for got, fd in ((gl, fl), (gq, fq)):
        gmax = float(np.max(np.abs(got[:, 0])))
        for k in coords:
            f2 = (fd_value(fd[k, 0]), fd_value(fd[k, 1]))
            err = abs(dd_val(got[k]) - f2[0])
            assert err <= fd_tolerance(f2, kap, gmax), (k, float(err))


@pytest.mark.parametrize("mistake", [1, 2, 3])
def test_f256_differences_reject_planted_mistakes(mistake):
    """Each planted error in the oracle's pair factors (a_u for a_v in the |R|^2 term, the
    3/4 term dropped, dS/dR's sign flipped) moves some gradient far outside the tolerance
    the correct gradient meets."""
    import quadprec_amd as qp
    mp.mp.dps = 80
    n = 24
    c, a = qp.gen_overlap_basis(9, n, 0.01, 10.0)
    b = qp.gen_dd_normal(19, n)
    _, _, gl, gq, _ = O.overlap_adjoint_dd(c, a, b)
    _, _, bl, bq, _ = O.overlap_adjoint_dd(c, a, b, mistake=mistake)
    coords = np.arange(4 * n)
    fl, fq = O.overlap_fd_f256(c, a, b, coords)
    s = O.overlap_s_dd(c, a)
    kap = np.linalg.cond(s[:, 0].reshape(n, n))
    caught = 0
    for good, bad, fd in ((gl, bl, fl), (gq, bq, fq)):
        gmax = float(np.max(np.abs(good[:, 0])))
        for k in coords:
            f2 = (fd_value(fd[k, 0]), fd_value(fd[k, 1]))
            tol = fd_tolerance(f2, kap, gmax)
            assert abs(dd_val(good[k]) - f2[0]) <= tol
            caught += abs(dd_val(bad[k]) - f2[0]) > 1e6 * tol
    assert caught >= 8, caught


def test_fixture_reference_gradient_matches_its_f256_differences():
    fx = G.load("adjoint_n2048")
    mp.mp.dps = 80
    kap = fx["kappa"]
    for which in ("ld", "q"):
        ref = {k: [G.fx(h) for h in v] for k, v in zip(fx["ref_coords"], fx["ref_" + which])}
        gmax = max(abs(v[0]) for v in ref.values())
        worst = 0.0
        for k, fd in zip(fx["fd_coords"], fx["fd_" + which]):
            limbs = [G.fx(h) for h in fd]
            f2 = (fd_value(limbs[:4]), fd_value(limbs[4:]))
            err = abs(dd_val(ref[k]) - f2[0])
            assert err <= fd_tolerance(f2, kap, gmax), (which, k, float(err))
            worst = max(worst, float(err / gmax))
        print(f"{which}: max |ref - fd| / max|g| = {worst:.3e} (kappa {kap:.2e})")
