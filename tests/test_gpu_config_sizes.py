"""The workload configurations at their full sizes, on the GPU, against the oracle.

Config 1  mf_gemm at 1024^3 (crates/quad/src/multifloat/linalg.rs:39-61): sampled rows
          bit-identical to the oracle; gemm_atb at 4096^3 (crates/quad/src/linalg.rs:168-182):
          the plain kernel's sampled rows bit-identical, the split kernel's sampled entries
          within 2^-100 of (|A^T||B|)_ij against the f256 oracle.
Config 2  N = 4096 overlap SPD solves (the three bench cases: kappa ~1e2, ~1e16 and ~1e25)
          with every factor kind the bench times, against x* = the oracle's
          cholesky_blocked (linalg.rs:238-293) + solve_cholesky refined with MultiFloat<4>
          residuals (multifloat/linalg.rs:17-36) until the correction is below 1e-34.  The bar
          is the configuration's: ||x - x*||_inf / ||x*||_inf <= 1e-30 (case a) and <= 1e-28
          (cases b and extreme), and the refinement must report convergence.
Config 3  a batch drawn like the bench's (n up to 1024, exponent floors down to 1e-3,
          spacing down to 0.7 bohr): every factor bit-identical to the oracle's
          cholesky_blocked(nb = 64); every solution's normwise backward error in f256 below
          n 2^-104; the first right-hand side within 4x the oracle's own solve_cholesky error
          (+2^-100) of x*, a bar that scales with each instance's conditioning.
"""
import time

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


def rel_err(x, xs):
    d = O.dd_binop("sub", x, xs)
    return float(np.max(np.abs(d[:, 0] + d[:, 1])) / np.max(np.abs(xs[:, 0])))


def refined_reference(s, b, n, l=None, iters=12):
    """x* for dd S and b: the oracle's dd factor and solves, corrected with MF4 residuals
    b - S x until the correction is below 1e-34 of x, or until it stops shrinking at the
    dd rounding level of x (below 1e-31 of x: x* is then exact to its dd storage)."""
    if l is None:
        l = O.cholesky_blocked(s, n, 64)
    x = O.solve_cholesky(l, b, n)
    s4, b4 = to_mf4(s), to_mf4(b)
    prev = np.inf
    for _ in range(iters):
        r4 = O.mf4_binop("sub", b4, O.mf4_gemv(s4, to_mf4(x), n, n))
        dx = O.solve_cholesky(l, np.ascontiguousarray(r4[:, :2]), n)
        x = O.dd_binop("add", x, dx)
        rel = np.max(np.abs(dx[:, 0])) / np.max(np.abs(x[:, 0]))
        if rel <= 1e-34 or (rel <= 1e-31 and rel > 0.25 * prev):
            return x
        prev = rel
    raise AssertionError(f"the oracle's MF4-residual refinement did not converge (last correction {rel:.2e})")


# ---------------------------------------------------------------- config 1

def test_mf_gemm_1024_sampled_rows_bitexact(qp):
    import torch
    n = 1024
    a = qp.gen_mf4_normal(1, n * n)
    b = qp.gen_mf4_normal(1 + 0x5EED, n * n)
    da, db = qp.dev.to_planes(a, 4), qp.dev.to_planes(b, 4)
    dc = torch.empty((4, n * n), dtype=torch.float64, device="cuda")
    qp.dev.mf_gemm(da, db, dc, n, n, n)
    torch.cuda.synchronize()
    c = qp.dev.from_planes(dc).reshape(n, n, 4)
    for row in (0, 1, 255, 511, 512, 777, 1022, 1023):
        ref = O.mf4_gemm_rows(a, b, n, n, n, row, 1).reshape(n, n, 4)[row]
        assert np.array_equal(c[row].view(np.uint64), ref.view(np.uint64)), row


@pytest.fixture(scope="module")
def atb_4096(qp):
    m = n = k = 4096
    a = qp.gen_dd_normal(1, k * m)   # A stored k x m
    b = qp.gen_dd_normal(1 + 0x5EED, k * n)
    return m, n, k, a, b


def dev_gemm_atb(qp, a, b, m, n, k, mode):
    import torch
    da, db = qp.dev.to_planes(a, 2), qp.dev.to_planes(b, 2)
    dc = torch.empty((2, m * n), dtype=torch.float64, device="cuda")
    qp.dev.gemm(da, db, dc, m, n, k, trans_a=True, mode=mode)
    torch.cuda.synchronize()
    return qp.dev.from_planes(dc).reshape(m, n, 2)


def test_gemm_atb_4096_plain_sampled_rows_bitexact(qp, atb_4096):
    m, n, k, a, b = atb_4096
    c = dev_gemm_atb(qp, a, b, m, n, k, qp.HP_GEMM_PLAIN)
    for row in (0, 1, 2047, 3001, 4095):
        ref = O.gemm_atb_rows(a, b, m, n, k, row, 1).reshape(m, n, 2)[row]
        assert np.array_equal(c[row].view(np.uint64), ref.view(np.uint64)), row


def test_gemm_atb_4096_split_sampled_vs_f256(qp, atb_4096):
    m, n, k, a, b = atb_4096
    c = dev_gemm_atb(qp, a, b, m, n, k, qp.HP_GEMM_SPLIT_I8)
    rows = np.array([0, 511, 1024, 2047, 2048, 3000, 4000, 4095])
    cols = np.arange(3, 4096, 16)
    at_s = np.ascontiguousarray(a.reshape(k, m, 2)[:, rows].transpose(1, 0, 2)).reshape(-1, 2)
    b_s = np.ascontiguousarray(b.reshape(k, n, 2)[:, cols]).reshape(-1, 2)
    ref = O.mf4_gemm(to_mf4(at_s), to_mf4(b_s), len(rows), len(cols), k)
    got = np.ascontiguousarray(c[rows][:, cols]).reshape(-1, 2)
    d = ((got[:, 0] - ref[:, 0]) - ref[:, 1] + got[:, 1]) - ref[:, 2] - ref[:, 3]
    ab = (np.abs(at_s[:, 0]).reshape(len(rows), k) @ np.abs(b_s[:, 0]).reshape(k, len(cols))).ravel()
    err = float(np.max(np.abs(d) / ab))
    print(f"split gemm_atb 4096^3: componentwise error {err:.3e} of (|A^T||B|)_ij")
    assert err <= 2.0 ** -100


# ---------------------------------------------------------------- config 2

SOLVE_N = 4096
# (name, alpha_lo, spacing, bar, factor kinds) -- the bench's N = 4096 cases, seed 2,
# right-hand side gen_dd_normal(102)
SOLVE_CASES = [
    ("a_kappa_small", 0.5, 1.4, 1e-30, ("f64", "dd")),
    ("b_kappa_1e16", 0.01, 1.4, 1e-28, ("split", "dd")),
    ("b_extreme_kappa_1e25", 0.001, 0.7, 1e-28, ("split", "dd")),
]


@pytest.mark.parametrize("name,lo,spacing,bar,kinds", SOLVE_CASES, ids=[c[0] for c in SOLVE_CASES])
def test_config2_solve_n4096_vs_refined_oracle(qp, name, lo, spacing, bar, kinds):
    import torch
    n, seed = SOLVE_N, 2
    c, a = qp.gen_overlap_basis(seed, n, lo, 10.0, spacing)
    b = qp.gen_dd_normal(seed + 100, n)
    t0 = time.perf_counter()
    s = O.overlap_s_dd(c, a)
    xs = refined_reference(s, b, n)
    t_ref = time.perf_counter() - t0
    # the device builds its own S; it must be the oracle's, bit for bit
    S = torch.empty((2, n * n), dtype=torch.float64, device="cuda")
    qp.dev.overlap_s(torch.from_numpy(c.ravel()).cuda(), torch.from_numpy(a).cuda(), S, n)
    torch.cuda.synchronize()
    assert np.array_equal(qp.dev.from_planes(S).view(np.uint64), s.view(np.uint64))
    B = qp.dev.to_planes(b, 2)
    X = torch.empty((2, n), dtype=torch.float64, device="cuda")
    kind_id = {"f64": qp.HP_FACTOR_F64, "dd": qp.HP_FACTOR_DD, "split": qp.HP_FACTOR_DD_SPLIT}
    for kind in kinds:
        st = qp.dev.solve_spd_refined(S, B, X, n, kind_id[kind], tol=bar * 1e-2, max_iter=30)
        torch.cuda.synchronize()
        e = rel_err(qp.dev.from_planes(X), xs)
        print(f"{name} factor={kind}: iters={st.iterations} converged={st.converged} "
              f"last={st.last_correction:.2e} rel_err={e:.3e} (oracle x* {t_ref:.1f} s)")
        assert st.converged != qp.HP_NOT_CONVERGED, (kind, st.iterations, st.last_correction)
        assert e <= bar, (kind, e)


# ---------------------------------------------------------------- config 3

def batch_case(qp, count=24, nrhs=4, seed=3):
    """Instances drawn as the bench draws them (exponent floor log-uniform in [1e-3, 0.5],
    spacing U[0.7, 1.4]) with sizes spread over [64, 1024], plus the two extremes."""
    rng = np.random.default_rng(seed)
    sizes = np.linspace(64, 1024, count).astype(np.int32)
    sizes[1::3] += 7                       # ragged last blocks
    sizes = np.minimum(sizes, 1024)
    lo = 10 ** rng.uniform(np.log10(1e-3), np.log10(0.5), count)
    sp = rng.uniform(0.7, 1.4, count)
    lo[-1], sp[-1] = 1e-3, 0.7             # the largest instance at the tightest settings
    lo[-2], sp[-2] = 0.5, 1.4
    mats, rhs = [], []
    for i, n in enumerate(sizes):
        c, a = qp.gen_overlap_basis(seed * 100000 + i, int(n), float(lo[i]), 10.0, float(sp[i]))
        mats.append(O.overlap_s_dd(c, a))
        rhs.append(qp.gen_dd_normal(seed * 1000 + i, nrhs * int(n)))
    return [int(v) for v in sizes], nrhs, mats, rhs


@pytest.mark.parametrize("schedule", [1, 2], ids=["per_instance", "panel_steps"])
def test_config3_batch_at_size(qp, schedule):
    import torch
    sizes, nrhs, mats, rhs = batch_case(qp)
    qp.dev.set_batch_schedule(schedule)
    S = qp.dev.to_planes(np.concatenate(mats), 2)
    B = qp.dev.to_planes(np.concatenate(rhs), 2)
    L = torch.zeros_like(S)
    X = torch.zeros_like(B)
    st = torch.full((len(sizes),), -7, dtype=torch.int32, device="cuda")
    try:
        qp.dev.batch_cholesky_solve(sizes, nrhs, S, L, B, X, st)
        torch.cuda.synchronize()
    finally:
        qp.dev.set_batch_schedule(0)
    Lh, Xh, sth = qp.dev.from_planes(L), qp.dev.from_planes(X), st.cpu().numpy()
    mo = ro = 0
    worst_bwd = worst_ratio = 0.0
    for i, n in enumerate(sizes):
        assert sth[i] == 0, (i, n, sth[i])
        ref_l = O.cholesky_blocked(mats[i], n, 64)
        lower = np.tril(np.ones((n, n), dtype=bool)).ravel()
        assert np.array_equal(Lh[mo:mo + n * n][lower].view(np.uint64), ref_l[lower].view(np.uint64)), (i, n)
        s4 = to_mf4(mats[i])
        snorm = float(np.max(np.abs(mats[i][:, 0]).reshape(n, n).sum(axis=1)))
        for r in range(nrhs):
            b = rhs[i][r * n:(r + 1) * n]
            x = Xh[ro + r * n:ro + (r + 1) * n]
            res = O.mf4_binop("sub", to_mf4(b), O.mf4_gemv(s4, to_mf4(x), n, n))
            bwd = float(np.max(np.abs(res[:, 0])) / (snorm * np.max(np.abs(x[:, 0])) + np.max(np.abs(b[:, 0]))))
            worst_bwd = max(worst_bwd, bwd / (n * 2.0 ** -104))
            assert bwd <= n * 2.0 ** -104, (i, n, r, bwd)
            if r == 0:
                xs = refined_reference(mats[i], b, n, l=ref_l)
                e_ref = rel_err(O.solve_cholesky(ref_l, b, n), xs)
                e = rel_err(x, xs)
                worst_ratio = max(worst_ratio, e / (e_ref + 2.0 ** -100))
                assert e <= 4 * e_ref + 2.0 ** -100, (i, n, e, e_ref)
        mo += n * n
        ro += nrhs * n
    print(f"config 3 batch of {len(sizes)} (n {min(sizes)}..{max(sizes)}): backward error <= "
          f"{worst_bwd:.2e} of n 2^-104, forward error <= {worst_ratio:.2f} x the oracle's own solve")
