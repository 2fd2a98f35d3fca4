"""Mixed-precision SPD solves with quad-double residual refinement (hp_solve_spd_refined),
checked against an independent CPU solution x*:

  x* = the oracle's dd cholesky_blocked (crates/quad/src/linalg.rs:238-293) and
  solve_cholesky, refined with residuals b - S x evaluated in MultiFloat<4>
  (crates/quad/src/multifloat/linalg.rs:17-36 mf_gemv) until the correction stagnates.

S is the Gaussian overlap matrix of config 0/2 (lattice Gaussians, exponents log-uniform),
b has dd N(0,1) entries.  The bar is ||x - x*||_inf / ||x*||_inf <= 1e-28 (kappa ~ 1e20)
and <= 1e-30 (kappa ~ 1e8 .. 1e10)."""
import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def qp():
    import quadprec_amd
    quadprec_amd.lib()
    return quadprec_amd


def to_mf4(x):
    out = np.zeros((x.shape[0], 4))
    out[:, :2] = x
    return out


def cpu_reference(s, b, n, iters=8):
    l = O.cholesky_blocked(s, n, 64)
    x = O.solve_cholesky(l, b, n)
    s4 = to_mf4(s)
    for _ in range(iters):
        ax = O.mf4_gemv(s4, to_mf4(x), n, n)
        r4 = O.mf4_binop("sub", to_mf4(b), ax)
        r = np.ascontiguousarray(r4[:, :2])
        dx = O.solve_cholesky(l, r, n)
        x = O.dd_binop("add", x, dx)
        if np.max(np.abs(dx[:, 0])) <= 1e-34 * np.max(np.abs(x[:, 0])):
            break
    return x


def rel_err(x, xs):
    d = O.dd_binop("sub", x, xs)
    return float(np.max(np.abs(d[:, 0] + d[:, 1])) / np.max(np.abs(xs[:, 0])))


@pytest.mark.parametrize("lo,spacing,kinds,bar", [
    (0.5, 1.4, ("f64", "dd", "split"), 1e-30),   # kappa ~ 1e2: every factor converges
    (0.01, 1.4, ("dd", "split"), 1e-28),          # config 0's exponent range (kappa ~ 2e9 here)
    (0.001, 0.7, ("dd", "split"), 1e-28),         # tight lattice, diffuse exponents: kappa far beyond 1e16
])
def test_refined_solve_vs_cpu_reference(qp, lo, spacing, kinds, bar):
    n = 512
    c, a = qp.gen_overlap_basis(0, n, lo, 10.0, spacing)
    s = qp.overlap_s(c, a)
    b = qp.gen_dd_normal(77, n)
    xs = cpu_reference(s, b, n)
    kind_id = {"f64": qp.HP_FACTOR_F64, "dd": qp.HP_FACTOR_DD, "split": qp.HP_FACTOR_DD_SPLIT}
    for kind in kinds:
        x, st = qp.solve_spd_refined(s, b, n, kind_id[kind], tol=bar * 1e-2, max_iter=30)
        e = rel_err(x, xs)
        print(f"alpha_lo={lo} spacing={spacing} factor={kind}: iters={st.iterations} err={e:.3e} "
              f"last={st.last_correction:.2e} "
              f"factor {st.factor_ms:.1f} ms refine {st.refine_ms:.1f} ms")
        assert e <= bar, (kind, e)


@pytest.mark.parametrize("n", [300, 700])
def test_refined_solve_ragged_vs_cpu_reference(qp, n):
    """n not a multiple of 64: every factor's last diagonal block and the block-inverse
    solves' last step are partial; each factor's refined solve against the oracle's x*."""
    c, a = qp.gen_overlap_basis(2, n, 0.05, 10.0, 1.4)
    s = qp.overlap_s(c, a)
    b = qp.gen_dd_normal(78, n)
    xs = cpu_reference(s, b, n)
    for kind in (qp.HP_FACTOR_F64, qp.HP_FACTOR_DD, qp.HP_FACTOR_DD_SPLIT):
        x, st = qp.solve_spd_refined(s, b, n, kind, tol=1e-30, max_iter=30)
        e = rel_err(x, xs)
        print(f"n={n} factor={kind}: iters={st.iterations} err={e:.3e}")
        assert e <= 1e-28, (kind, e)


@pytest.mark.parametrize("n,piv", [(64, 10), (520, 300), (520, 517)])
def test_refined_solve_split_factor_fails_cleanly_on_indefinite(qp, n, piv):
    """A negative pivot in the first block, in a later panel's first sub-block, and in the
    ragged last sub-block: HP_ENOTSPD with that pivot."""
    s = np.zeros((n * n, 2))
    for i in range(n):
        s[i * n + i, 0] = 1.0
    s[piv * n + piv, 0] = -1.0
    with pytest.raises(qp.SolveError) as e:
        qp.solve_spd_refined(s, qp.gen_dd_normal(1, n), n, qp.HP_FACTOR_DD_SPLIT)
    assert e.value.pivot == piv


def test_split_cholesky_factor_matches_plain_to_dd_accuracy(qp):
    """The split factor differs from the reference-order one only by rounding: its
    residual against S after one refinement-free solve is at the plain factor's level."""
    n = 300
    c, a = qp.gen_overlap_basis(3, n, 0.5, 10.0)
    s = qp.overlap_s(c, a)
    b = qp.gen_dd_normal(5, n)
    x1, st1 = qp.solve_spd_refined(s, b, n, qp.HP_FACTOR_DD, tol=0.0, max_iter=1)
    x2, st2 = qp.solve_spd_refined(s, b, n, qp.HP_FACTOR_DD_SPLIT, tol=0.0, max_iter=1)
    assert rel_err(x2, x1) < 1e-20


@pytest.mark.parametrize("nb", [64, 128, 256, 512])
@pytest.mark.parametrize("n", [300, 700])
def test_split_factor_panel_widths(qp, nb, n):
    """The split-dd factorisation at every panel width (three-stream schedule, ragged last
    panel): one refinement-free solve agrees with the reference-order dd factor's to dd
    accuracy at kappa ~ 1e3..1e6, and the refined solve still meets 1e-30."""
    c, a = qp.gen_overlap_basis(11, n, 0.3, 10.0)
    s = qp.overlap_s(c, a)
    b = qp.gen_dd_normal(12, n)
    x1, _ = qp.solve_spd_refined(s, b, n, qp.HP_FACTOR_DD, tol=0.0, max_iter=1)
    qp.set_split_factor_panel(nb)
    try:
        x2, _ = qp.solve_spd_refined(s, b, n, qp.HP_FACTOR_DD_SPLIT, tol=0.0, max_iter=1)
        x3, st = qp.solve_spd_refined(s, b, n, qp.HP_FACTOR_DD_SPLIT, tol=1e-32, max_iter=10)
        x4, _ = qp.solve_spd_refined(s, b, n, qp.HP_FACTOR_DD, tol=1e-32, max_iter=10)
    finally:
        qp.set_split_factor_panel(256)
    assert rel_err(x2, x1) < 1e-22, rel_err(x2, x1)
    assert rel_err(x3, x4) < 1e-30, rel_err(x3, x4)


def test_split_factor_pivots_across_binades(qp):
    """The split factor's pivots (the scaled dd reciprocal square root of its diagonal-block
    kernel) over d spanning 2^-600 .. 2^600 with random dd low parts.  S = diag(d): one
    correction from x = 0 (max_iter=1) is x = L^-T L^-1 b with L = diag(sqrt(d)), so its
    error is the pivots' (plus the block inverses' dd divisions).  The bar, 2^-99 per
    component against the oracle's dd 1/d, would fail for a pivot accurate only to binary64
    or to half of dd."""
    rng = np.random.default_rng(11)
    n = 300
    hi = np.ldexp(1.0 + rng.random(n), rng.integers(-600, 600, n))
    lo = hi * rng.uniform(-1.0, 1.0, n) * 2.0 ** -53
    s_ = hi + lo
    d = np.stack([s_, lo - (s_ - hi)], axis=1)
    smat = np.zeros((n * n, 2))
    smat[np.arange(n) * (n + 1)] = d
    b = np.zeros((n, 2))
    b[:, 0] = 1.0
    x, st = qp.solve_spd_refined(smat, b, n, qp.HP_FACTOR_DD_SPLIT, tol=0.0, max_iter=1)
    ref = O.dd_binop("div", b, d)
    diff = O.dd_binop("sub", x, ref)
    err = np.abs(diff[:, 0] + diff[:, 1]) / np.abs(ref[:, 0])
    print(f"worst relative error {err.max():.3e} (2^{np.log2(max(err.max(), 1e-300)):.1f})")
    assert err.max() <= 2.0 ** -99, err.max()
