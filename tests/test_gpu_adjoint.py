"""Adjoint of the overlap build and Cholesky factor on the GPU: gradients of log det S and
b^T S^-1 b over the 3n centre coordinates and n exponents, in dd, through the C ABI
(hp_overlap_adjoint, hp_dev_overlap_forward / _gradients / _backward).

Checked against the oracle's gradient assembled from the reference's dd operations
(S^-1 by solve_cholesky against unit vectors; oracle/hpo.cpp hpo_overlap_adjoint_dd) and,
at N = 2048 (seed 4, kappa ~ 1e14), against the committed f256 central differences.  The
two dd computations differ in algorithm (L^-1 by recursive doubling vs solves, reordered
sums, a refined x), so their difference is bounded by what dd arithmetic allows at the
condition number: |g - g_ref| <= 16 kappa 2^-104 max|g| (floor 1e-30) at small n and
4 kappa 2^-104 max|g| at N = 2048; against the f256 differences the dd term is
kappa 2^-104 max|g| on top of the step error.
The vector-Jacobian product of the backward is checked exactly against the same dd
combination of the two gradients."""
import numpy as np
import pytest

import golden_util as G
import oracle_lib as O

pytestmark = pytest.mark.gpu
U = 2.0 ** -104


@pytest.fixture(scope="module")
def qp():
    import quadprec_amd
    quadprec_amd.lib()
    return quadprec_amd


def kappa_of(s, n):
    w = np.linalg.eigvalsh(s[:, 0].reshape(n, n))
    return float(w[-1] / max(w[0], w[-1] * 1e-17))


def dd_err(g, r):
    d = O.dd_binop("sub", g, r)
    return np.abs(d[:, 0] + d[:, 1])


CASES = [(1, 0.01), (2, 0.5), (37, 0.01), (64, 0.5), (65, 0.01), (200, 0.01), (333, 0.5), (333, 0.02)]


@pytest.mark.parametrize("n,lo", CASES)
def test_gradients_match_oracle(qp, n, lo):
    c, a = qp.gen_overlap_basis(4, n, lo, 10.0)
    b = qp.gen_dd_normal(44, n)
    ld, q, gl, gq, x = O.overlap_adjoint_dd(c, a, b)
    kap = kappa_of(O.overlap_s_dd(c, a), n)
    modes = [(qp.HP_FACTOR_DD, qp.HP_GEMM_PLAIN), (qp.HP_FACTOR_DD, qp.HP_GEMM_SPLIT_I8),
             (qp.HP_FACTOR_DD_SPLIT, qp.HP_GEMM_SPLIT_I8)]
    for fk, gm in modes:
        g1, g2, info = qp.overlap_adjoint(c, a, b, fk, gm)
        tol = max(16 * kap * U, 1e-30)
        e_ld = abs((info.logdet[0] - ld[0]) + (info.logdet[1] - ld[1]))
        e_q = abs((info.quad[0] - q[0]) + (info.quad[1] - q[1]))
        print(f"n={n} lo={lo} kappa={kap:.2e} factor={fk} gemm={gm}: logdet {e_ld:.1e} quad {e_q / abs(q[0]):.1e} "
              f"grad_ld {dd_err(g1, gl).max() / max(np.abs(gl[:, 0]).max(), 1e-300):.1e} "
              f"grad_q {dd_err(g2, gq).max() / max(np.abs(gq[:, 0]).max(), 1e-300):.1e}")
        assert e_ld <= tol * max(1.0, abs(ld[0]))
        assert e_q <= tol * abs(q[0])
        if n > 1:
            assert dd_err(g1, gl).max() <= tol * np.abs(gl[:, 0]).max()
            assert dd_err(g2, gq).max() <= tol * np.abs(gq[:, 0]).max()
        else:
            assert not g1.any() and not g2.any()


def test_backward_is_the_dd_vjp_of_the_gradients(qp):
    """backward(c1, c2) = c1 dF1 + c2 dF2 and grad_b = 2 c2 x, bit for bit the dd operations
    applied to the full gradients; with c1 = 0 no S^-1 is formed and the result is c2 dF2."""
    import torch
    n = 150
    c, a = qp.gen_overlap_basis(7, n, 0.05, 10.0)
    b = qp.gen_dd_normal(8, n)
    st = qp.dev.OverlapAdjoint(torch.from_numpy(c.ravel()).cuda(), torch.from_numpy(a).cuda(),
                               qp.dev.to_planes(b, 2))
    g1, g2 = (qp.dev.from_planes(g) for g in st.gradients())
    x = qp.dev.from_planes(st.backward((0.0, 0.0), (0.5, 0.0))[2])  # 2 * 0.5 * x is x exactly
    for c1, c2 in (((0.75, 1e-18), (-1.25, 3e-17)), ((0.0, 0.0), (2.0, 0.0)), ((1.0, 0.0), (0.0, 0.0))):
        gc, ga, gb = st.backward(c1, c2)
        got = np.concatenate([qp.dev.from_planes(gc), qp.dev.from_planes(ga)])
        want = O.dd_binop("mul", np.array(c2), g2)
        if c1 != (0.0, 0.0):
            want = O.dd_binop("add", O.dd_binop("mul", np.array(c1), g1), want)
        assert np.array_equal(got.view(np.uint64), want.view(np.uint64)), (c1, c2)
        want_b = O.dd_binop("mul_f64", O.dd_binop("mul", np.array(c2), x), np.array([2.0, 0.0]))
        assert np.array_equal(qp.dev.from_planes(gb).view(np.uint64), want_b.view(np.uint64)), (c1, c2)
    st.close()


def test_autograd_function_matches_dd_gradients(qp):
    import torch
    n = 96
    c, a = qp.gen_overlap_basis(5, n, 0.1, 10.0)
    bv = np.random.default_rng(5).standard_normal(n)
    b = np.stack([bv, np.zeros(n)], axis=1)
    g1, g2, info = qp.overlap_adjoint(c, a, b)
    cen = torch.tensor(c.reshape(n, 3), device="cuda", requires_grad=True)
    alp = torch.tensor(a, device="cuda", requires_grad=True)
    bt = torch.tensor(bv, device="cuda", requires_grad=True)
    ld, q = qp.dev.overlap_logdet_quad(cen, alp, bt)
    assert float(ld.detach()) == info.logdet[0] + info.logdet[1]
    assert float(q.detach()) == info.quad[0] + info.quad[1]
    (0.5 * ld + 0.25 * q).backward()
    want = O.dd_binop("add", O.dd_binop("mul", np.tile([0.5, 0.0], (4 * n, 1)), g1),
                      O.dd_binop("mul", np.tile([0.25, 0.0], (4 * n, 1)), g2))
    want64 = want[:, 0] + want[:, 1]
    got = np.concatenate([cen.grad.cpu().numpy().ravel(), alp.grad.cpu().numpy()])
    assert np.array_equal(got, want64)
    st = qp.dev.OverlapAdjoint(torch.from_numpy(c.ravel()).cuda(), torch.from_numpy(a).cuda(), qp.dev.to_planes(b, 2))
    gb = qp.dev.from_planes(st.backward((0.0, 0.0), (0.25, 0.0))[2])
    assert np.array_equal(bt.grad.cpu().numpy(), gb[:, 0] + gb[:, 1])


def test_forward_reports_not_positive_definite(qp):
    c = np.array([[0.0, 0.0, 0.0], [0.0, 0.0, 0.0]])
    a = np.array([0.7, 0.7])  # two identical Gaussians: S = [[1, 1], [1, 1]]
    b = qp.gen_dd_normal(1, 2)
    with pytest.raises(qp.SolveError) as e:
        qp.overlap_adjoint(c, a, b)
    assert e.value.pivot == 1


@pytest.fixture(scope="module")
def config4(qp):
    fx = G.load("adjoint_n2048")
    n = fx["n"]
    c, a = qp.gen_overlap_basis(fx["seed"], n, fx["alpha_lo"], fx["alpha_hi"], fx["spacing"], fx["jitter"])
    b = qp.gen_dd_normal(fx["b_seed"], n)
    return fx, c, a, b


@pytest.mark.parametrize("mode", ["dd_plain", "split_split"])
def test_config4_n2048_against_f256_differences(qp, config4, mode):
    """Config 4: N = 2048, seed 4.  Every one of the 64 f256 central differences (h = 1e-15)
    agrees with the GPU gradient within 2 |fd(h) - fd(2h)| + kappa 2^-104 max|g|, and the
    256 sampled oracle dd gradient entries within 4 kappa 2^-104 max|g|."""
    fx, c, a, b = config4
    fk, gm = ((qp.HP_FACTOR_DD, qp.HP_GEMM_PLAIN) if mode == "dd_plain"
              else (qp.HP_FACTOR_DD_SPLIT, qp.HP_GEMM_SPLIT_I8))
    g1, g2, info = qp.overlap_adjoint(c, a, b, fk, gm)
    kap = fx["kappa"]
    ld = [G.fx(h) for h in fx["logdet"]]
    assert abs((info.logdet[0] - ld[0]) + (info.logdet[1] - ld[1])) <= 4 * kap * U * abs(ld[0])
    for which, g in (("ld", g1), ("q", g2)):
        gmax = np.abs(g[:, 0]).max()
        worst_fd = worst_ref = 0.0
        for k, fd in zip(fx["fd_coords"], fx["fd_" + which]):
            limbs = [G.fx(h) for h in fd]
            f_h = G.exact_sum(*limbs[:4])
            f_2h = G.exact_sum(*limbs[4:])
            err = abs(G.exact_sum(g[k, 0], g[k, 1]) - f_h)
            tol = 2 * abs(f_h - f_2h) + G.exact_sum(kap * U * gmax)
            assert err <= tol, (which, k, float(err))
            worst_fd = max(worst_fd, float(err) / gmax)
        for k, r in zip(fx["ref_coords"], fx["ref_" + which]):
            rv = [G.fx(h) for h in r]
            err = abs((g[k, 0] - rv[0]) + (g[k, 1] - rv[1]))
            assert err <= 4 * kap * U * gmax, (which, k, err)
            worst_ref = max(worst_ref, err / gmax)
        print(f"{mode} grad {which}: vs f256 fd {worst_fd:.2e}, vs oracle dd {worst_ref:.2e} (kappa {kap:.2e}, "
              f"kappa u = {kap * U:.1e})")
