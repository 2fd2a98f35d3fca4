"""Batched uneven factor + solve in both schedules (one workgroup per instance; panel steps
across the chip), eager and as a replayed hipGraph.

Each instance's factor must be bit-identical to the oracle's cholesky_blocked(nb=64)
(crates/quad/src/linalg.rs:238-293) -- both schedules keep the per-element order -- with
the upper triangle left as S, and each solution must agree with the oracle's
solve_cholesky (linalg.rs:123-126) to the dd accuracy the conditioning allows (the batched
backward solve is reordered)."""
import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def qp():
    import quadprec_amd
    quadprec_amd.lib()
    return quadprec_amd


@pytest.fixture(params=[1, 2], ids=["per_instance", "panel_steps"])
def schedule(qp, request):
    qp.dev.set_batch_schedule(request.param)
    yield request.param
    qp.dev.set_batch_schedule(0)


def pack(qp, sizes, nrhs, seed0=100, lo=0.5):
    mats, rhs = [], []
    for i, n in enumerate(sizes):
        c, a = qp.gen_overlap_basis(seed0 + i, n, lo, 10.0)
        mats.append(qp.overlap_s(c, a))
        rhs.append(qp.gen_dd_normal(seed0 + 1000 + i, nrhs * n))
    return mats, rhs


def run_batch(qp, sizes, nrhs, mats, rhs, graph=False, replays=1):
    import torch
    S = qp.dev.to_planes(np.concatenate(mats), 2)
    B = qp.dev.to_planes(np.concatenate(rhs), 2)
    L = torch.zeros_like(S)
    X = torch.zeros_like(B)
    st = torch.full((len(sizes),), -7, dtype=torch.int32, device="cuda")
    if graph:
        g = qp.dev.batch_cholesky_solve(sizes, nrhs, S, L, B, X, st, graph=True)
        outs = []
        for _ in range(replays):
            L.zero_()
            X.zero_()
            g.launch()
            torch.cuda.synchronize()
            outs.append((qp.dev.from_planes(L), qp.dev.from_planes(X), st.cpu().numpy().copy()))
        return outs
    qp.dev.batch_cholesky_solve(sizes, nrhs, S, L, B, X, st)
    torch.cuda.synchronize()
    return qp.dev.from_planes(L), qp.dev.from_planes(X), st.cpu().numpy()


def check(sizes, nrhs, mats, rhs, L, X, st):
    mo = ro = 0
    for i, n in enumerate(sizes):
        assert st[i] == 0
        li = L[mo:mo + n * n]
        ref_l = O.cholesky_blocked(mats[i], n, 64)
        lower = np.tril(np.ones((n, n), dtype=bool)).ravel()
        assert np.array_equal(li[lower].view(np.uint64), ref_l[lower].view(np.uint64)), i
        assert np.array_equal(li[~lower].view(np.uint64), mats[i][~lower].view(np.uint64)), i  # upper keeps S
        for r in range(nrhs):
            b = rhs[i][r * n:(r + 1) * n]
            x = X[ro + r * n:ro + (r + 1) * n]
            xr = O.solve_cholesky(ref_l, b, n)
            d = O.dd_binop("sub", x, xr)
            err = np.max(np.abs(d[:, 0] + d[:, 1])) / np.max(np.abs(xr[:, 0]))
            assert err < 1e-27, (i, r, err)
        mo += n * n
        ro += nrhs * n


def test_batch_factor_bitexact_and_solutions(qp, schedule):
    sizes = [64, 65, 100, 130, 200, 37, 1, 128, 300, 257, 90, 150]
    nrhs = 4
    mats, rhs = pack(qp, sizes, nrhs)
    L, X, st = run_batch(qp, sizes, nrhs, mats, rhs)
    check(sizes, nrhs, mats, rhs, L, X, st)


def test_batch_graph_replays_match_eager(qp, schedule):
    sizes = [70, 140, 33, 256, 190]
    nrhs = 3
    mats, rhs = pack(qp, sizes, nrhs, seed0=500)
    L0, X0, st0 = run_batch(qp, sizes, nrhs, mats, rhs)
    outs = run_batch(qp, sizes, nrhs, mats, rhs, graph=True, replays=3)
    check(sizes, nrhs, mats, rhs, L0, X0, st0)
    for L, X, st in outs:
        assert np.array_equal(L.view(np.uint64), L0.view(np.uint64))
        assert np.array_equal(X.view(np.uint64), X0.view(np.uint64))
        assert np.array_equal(st, st0)


def test_batch_reports_breakdown_per_instance(qp, schedule):
    sizes = [50, 80, 40]
    nrhs = 1
    mats, rhs = pack(qp, sizes, nrhs, seed0=900)
    bad = mats[1].copy()
    bad[33 * 80 + 33] = [-1.0, 0.0]
    mats[1] = bad
    L, X, st = run_batch(qp, sizes, nrhs, mats, rhs)
    ref = None
    try:
        O.cholesky_blocked(bad, 80, 64)
    except O.CholeskyError as e:
        ref = e.pivot
    assert st[0] == 0 and st[2] == 0
    assert st[1] == ref + 1


def test_autotuned_schedule_is_one_of_the_two(qp):
    """With no pin, the first batch of a class times both schedules and the graph captures
    the faster; either way the results are the reference's."""
    import torch
    qp.dev.set_batch_schedule(0)
    sizes = [200, 64, 130, 90, 257]
    nrhs = 2
    mats, rhs = pack(qp, sizes, nrhs, seed0=700)
    S = qp.dev.to_planes(np.concatenate(mats), 2)
    B = qp.dev.to_planes(np.concatenate(rhs), 2)
    L, X = torch.zeros_like(S), torch.zeros_like(B)
    st = torch.full((len(sizes),), -7, dtype=torch.int32, device="cuda")
    g = qp.dev.batch_cholesky_solve(sizes, nrhs, S, L, B, X, st, graph=True)
    assert g.schedule in (1, 2)
    L.zero_()
    X.zero_()
    g.launch()
    torch.cuda.synchronize()
    check(sizes, nrhs, mats, rhs, qp.dev.from_planes(L), qp.dev.from_planes(X), st.cpu().numpy())


def test_schedules_give_identical_solutions(qp):
    """The panel-step schedule solves in block steps across the batch with the per-instance
    solve's operations in its order, so X (and L) agree bit for bit between the schedules,
    ragged sizes and eight right-hand sides included."""
    sizes = [64, 65, 1, 129, 300, 191, 257, 40]
    nrhs = 8
    mats, rhs = pack(qp, sizes, nrhs, seed0=1300)
    out = {}
    for sch in (1, 2):
        qp.dev.set_batch_schedule(sch)
        try:
            out[sch] = run_batch(qp, sizes, nrhs, mats, rhs)
        finally:
            qp.dev.set_batch_schedule(0)
    (L1, X1, s1), (L2, X2, s2) = out[1], out[2]
    assert np.array_equal(s1, s2) and not s1.any()
    assert np.array_equal(L1.view(np.uint64), L2.view(np.uint64))
    assert np.array_equal(X1.view(np.uint64), X2.view(np.uint64))
    check(sizes, nrhs, mats, rhs, L2, X2, s2)
