"""Split-precision dd GEMM (INT8 matrix cores + CRT) against the f256 oracle.

The oracle is the MultiFloat<4> GEMM restatement (crates/quad/src/multifloat/linalg.rs:39-61
on dd inputs promoted with from_f128, multifloat/conv.rs:120-131), accurate to ~2^-200.
The bar is componentwise: |C - C*|_ij <= tol * (|A||B|)_ij with tol = 2^-100, and the
split result must be at least as accurate as the plain dd GEMM on the same inputs up to a
factor 4 (both are measured, the worst ratios are printed)."""
import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu
TOL = 2.0 ** -100


@pytest.fixture(scope="module")
def qp():
    import quadprec_amd
    quadprec_amd.lib()
    return quadprec_amd


def to_mf4(x):
    out = np.zeros((x.shape[0], 4))
    out[:, :2] = x
    return out


def dev_gemm(qp, a, b, m, n, k, mode, trans_a=False, slices=0):
    import torch
    da, db = qp.dev.to_planes(a, 2), qp.dev.to_planes(b, 2)
    dc = torch.empty((2, m * n), dtype=torch.float64, device="cuda")
    qp.dev.gemm(da, db, dc, m, n, k, trans_a=trans_a, mode=mode, slices=slices)
    torch.cuda.synchronize()
    return qp.dev.from_planes(dc)


def comp_err(c, ref4, absab):
    """max_ij |c - ref| / (|A||B|)_ij, with ref in 4 limbs."""
    d = np.zeros(c.shape[0])
    # (c.hi - r0) + (c.lo - r1) - r2 - r3, each difference is exact or nearly so
    d = ((c[:, 0] - ref4[:, 0]) - ref4[:, 1] + c[:, 1]) - ref4[:, 2] - ref4[:, 3]
    return float(np.max(np.abs(d) / np.maximum(absab, 1e-300)))


def absprod(a, b, m, n, k, trans_a=False):
    A = np.abs(a[:, 0]).reshape((k, m) if trans_a else (m, k))
    if trans_a:
        A = A.T
    return (A @ np.abs(b[:, 0]).reshape(k, n)).ravel()


@pytest.mark.parametrize("m,n,k", [(64, 64, 64), (100, 37, 513), (200, 130, 300), (1, 300, 17), (257, 5, 4)])
def test_split_i8_vs_f256(qp, m, n, k):
    a = qp.gen_dd_normal(31 + m, m * k)
    b = qp.gen_dd_normal(32 + n, k * n)
    ref = O.mf4_gemm(to_mf4(a), to_mf4(b), m, n, k)
    ab = absprod(a, b, m, n, k)
    c_split = dev_gemm(qp, a, b, m, n, k, qp.HP_GEMM_SPLIT_I8)
    c_plain = dev_gemm(qp, a, b, m, n, k, qp.HP_GEMM_PLAIN)
    e_split, e_plain = comp_err(c_split, ref, ab), comp_err(c_plain, ref, ab)
    print(f"m={m} n={n} k={k}: split {e_split:.3e}  plain {e_plain:.3e}")
    assert e_split <= TOL
    assert e_split <= max(4 * e_plain, 2.0 ** -104)


@pytest.mark.parametrize("variant", [32, 64, 128, 3064, 4064, 5064, 6064, 7064])
def test_split_i8_every_k_stage_variant(qp, variant):
    m, n, k = 300, 270, 576  # 9 stages of 64: whole three-stage trips for 6064
    a = qp.gen_dd_normal(41, m * k)
    b = qp.gen_dd_normal(42, k * n)
    qp.split_gemm_set_variant(m, n, k, variant)
    try:
        c = dev_gemm(qp, a, b, m, n, k, qp.HP_GEMM_SPLIT_I8)
        assert qp.split_gemm_variant(m, n, k) == variant
    finally:
        qp.split_gemm_set_variant(m, n, k, 0)
    rows = np.arange(0, m, 37)
    a_s = a.reshape(m, k, 2)[rows].reshape(-1, 2)
    ref = O.mf4_gemm(to_mf4(a_s), to_mf4(b), len(rows), n, k)
    got = np.ascontiguousarray(c.reshape(m, n, 2)[rows]).reshape(-1, 2)
    assert comp_err(got, ref, absprod(a_s, b, len(rows), n, k)) <= TOL


def test_split_i8_gemm_atb_and_dynamic_range(qp):
    m, n, k = 90, 70, 150
    a = qp.gen_dd_normal(5, k * m)   # stored k x m
    b = qp.gen_dd_normal(6, k * n)
    a[::7] *= 1e-150                 # rows of wildly different magnitude
    b[3::11] *= 1e120
    a[k * 0:k * 0 + m] = 0.0         # one zero row of the stored matrix
    at = a.reshape(k, m, 2).transpose(1, 0, 2).reshape(m * k, 2)
    ref = O.mf4_gemm(to_mf4(at), to_mf4(b), m, n, k)
    ab = absprod(a, b, m, n, k, trans_a=True)
    c = dev_gemm(qp, a, b, m, n, k, qp.HP_GEMM_SPLIT_I8, trans_a=True)
    assert comp_err(c, ref, ab) <= TOL


def test_split_i8_exact_on_small_integers(qp):
    # integer inputs well inside W bits: the split product is exact
    m, n, k = 33, 47, 129
    rng = np.random.default_rng(1)
    a = np.zeros((m * k, 2))
    b = np.zeros((k * n, 2))
    a[:, 0] = rng.integers(-1000, 1000, m * k)
    b[:, 0] = rng.integers(-1000, 1000, k * n)
    c = dev_gemm(qp, a, b, m, n, k, qp.HP_GEMM_SPLIT_I8)
    ex = (a[:, 0].reshape(m, k).astype(np.int64) @ b[:, 0].reshape(k, n).astype(np.int64)).ravel()
    assert np.array_equal(c[:, 0], ex.astype(np.float64)) and not c[:, 1].any()


@pytest.mark.parametrize("variant,k", [(128, 70000), (3064, 70000), (4064, 70000), (5064, 70000), (6064, 69888),
                                       (7064, 70000)])
def test_split_i8_exact_beyond_one_reduction_chunk(qp, variant, k):
    """k = 70000: the int32 sums are reduced mod m once after 65536 products; K is padded to
    70016 with zero residues and N = 47 leaves a partial dword of CRT digits per row
    (5064 has no reduction and runs 4064 there; 6064 takes whole three-stage trips,
    k = 69888 = 364 * 192)."""
    m, n = 33, 47
    rng = np.random.default_rng(2)
    a = np.zeros((m * k, 2))
    b = np.zeros((k * n, 2))
    a[:, 0] = rng.integers(-1000, 1000, m * k)
    b[:, 0] = rng.integers(-1000, 1000, k * n)
    qp.split_gemm_set_variant(m, n, k, variant)
    try:
        c = dev_gemm(qp, a, b, m, n, k, qp.HP_GEMM_SPLIT_I8)
    finally:
        qp.split_gemm_set_variant(m, n, k, 0)
    ex = (a[:, 0].reshape(m, k).astype(np.int64) @ b[:, 0].reshape(k, n).astype(np.int64)).ravel()
    assert np.array_equal(c[:, 0], ex.astype(np.float64)) and not c[:, 1].any()


@pytest.mark.parametrize("m,n,k", [(1100, 1030, 64), (1100, 1030, 128), (1100, 1030, 192), (1100, 1030, 256),
                                   (1100, 1030, 1000), (2048, 2048, 256)])
def test_split_i8_persistent_matches_tiled(qp, m, n, k):
    """The persistent int8 kernel (7064: one block per CU walking tiles, the next tile's first
    K stages loaded during the digit epilogue) forms the same digits as the one-tile-per-block
    kernel 4064, hence the same dd product bit for bit: full tiles (their stores issued between
    the next tile's loads and its waits), ragged edge tiles (stores first), 1 to 16 K stages,
    several tiles per block, and the 4 x 8 patch order (2048^2)."""
    a = qp.gen_dd_normal(61, m * k)
    b = qp.gen_dd_normal(62, k * n)
    out = {}
    for v in (4064, 7064):
        qp.split_gemm_set_variant(m, n, k, v)
        try:
            out[v] = dev_gemm(qp, a, b, m, n, k, qp.HP_GEMM_SPLIT_I8)
        finally:
            qp.split_gemm_set_variant(m, n, k, 0)
    assert np.array_equal(out[4064], out[7064])


@pytest.mark.parametrize("slices", [24, 28, 32])
def test_split_i8_slices_tradeoff(qp, slices):
    m, n, k = 64, 64, 256
    a = qp.gen_dd_normal(7, m * k)
    b = qp.gen_dd_normal(8, k * n)
    ref = O.mf4_gemm(to_mf4(a), to_mf4(b), m, n, k)
    c = dev_gemm(qp, a, b, m, n, k, qp.HP_GEMM_SPLIT_I8, slices=slices)
    e = comp_err(c, ref, absprod(a, b, m, n, k))
    # W = floor((log2 M - 12 - 4) / 2) bits per input: 24 -> ~82, 28 -> ~96, 32 -> ~112
    W = {24: 82, 28: 96, 32: 112}[slices]
    assert e <= 2.0 ** (4 - W)


def test_split_i8_full_size_sampled(qp):
    """4096^3 split GEMM: a sample of 8 rows x 256 columns against the f256 oracle."""
    m = n = k = 4096
    a = qp.gen_dd_normal(1, m * k)
    b = qp.gen_dd_normal(1 + 0x5EED, k * n)
    c = dev_gemm(qp, a, b, m, n, k, qp.HP_GEMM_SPLIT_I8).reshape(m, n, 2)
    rows = np.array([0, 511, 1024, 2047, 2048, 3000, 4000, 4095])
    cols = np.arange(0, 4096, 16)
    a_s = a.reshape(m, k, 2)[rows].reshape(-1, 2)
    b_s = np.ascontiguousarray(b.reshape(k, n, 2)[:, cols]).reshape(-1, 2)
    ref = O.mf4_gemm(to_mf4(a_s), to_mf4(b_s), len(rows), len(cols), k)
    got = np.ascontiguousarray(c[rows][:, cols]).reshape(-1, 2)
    ab = absprod(a_s, b_s, len(rows), len(cols), k)
    assert comp_err(got, ref, ab) <= TOL


def split_norm_bound(a, b, m, n, k):
    """max_p|A_ip| sum_p|B_pj| + sum_p|A_ip| max_p|B_pj| per element (the split form's norm)."""
    A = np.abs(a[:, 0]).reshape(m, k)
    B = np.abs(b[:, 0]).reshape(k, n)
    return (np.outer(A.max(axis=1), B.sum(axis=0)) + np.outer(A.sum(axis=1), B.max(axis=0))).ravel()


def abs_err(c, ref4):
    return np.abs(((c[:, 0] - ref4[:, 0]) - ref4[:, 1] + c[:, 1]) - ref4[:, 2] - ref4[:, 3])


@pytest.mark.parametrize("bits", [30, 53, 64, 80, 96, 104, 110])
def test_auto_mode_meets_the_requested_bits(qp, bits):
    """HP_GEMM_AUTO: the plan takes the fewest moduli whose W reaches bits + 1 (or the plain
    kernel when its (k+1) 2^-102 bound meets bits and it is faster), and the result is within
    2^-bits of the split norm against the f256 oracle."""
    m, n, k = 130, 90, 1000
    a = qp.gen_dd_normal(70 + bits, m * k)
    b = qp.gen_dd_normal(71 + bits, k * n)
    c = dev_gemm(qp, a, b, m, n, k, qp.HP_GEMM_AUTO, slices=bits)
    plan = qp.gemm_plan(m, n, k, bits)
    assert plan is not None
    r = qp.split_slices_for_bits(k, bits)
    assert r >= 8 and r <= 32
    if plan.mode == qp.HP_GEMM_SPLIT_I8:
        assert plan.slices == r and plan.w_bits >= bits + 1
    else:
        assert plan.mode == qp.HP_GEMM_PLAIN and bits <= 102 - 10
    ref = O.mf4_gemm(to_mf4(a), to_mf4(b), m, n, k)
    bound = 2.0 ** -bits * split_norm_bound(a, b, m, n, k) + 2.0 ** -105 * np.abs(ref[:, 0])
    err = abs_err(c, ref)
    print(f"bits={bits}: mode={plan.mode} slices={plan.slices} W={plan.w_bits} plain {plan.ms_plain:.3f} ms "
          f"split {plan.ms_split:.3f} ms; worst error / bound {np.max(err / bound):.3e}")
    assert np.all(err <= bound)


def test_auto_mode_rejects_unreachable_accuracy(qp):
    import torch
    a = qp.dev.to_planes(qp.gen_dd_normal(1, 64), 2)
    c = torch.empty((2, 64), dtype=torch.float64, device="cuda")
    with pytest.raises(qp.HyperprecError):
        qp.dev.gemm(a, a, c, 8, 8, 8, mode=qp.HP_GEMM_AUTO, slices=113)


def test_split_nonfinite_rows_and_columns_give_nan(qp):
    """inf / NaN in a row of A or a column of B: that row / column of C is NaN, exactly where
    the reference's dd arithmetic (the bit-exact oracle) puts NaNs; the other entries stay
    within the split bound."""
    m, n, k = 70, 50, 130
    a = qp.gen_dd_normal(81, m * k)
    b = qp.gen_dd_normal(82, k * n)
    a[3 * k + 5] = [np.inf, 0.0]
    a[10 * k + 7] = [np.nan, np.nan]
    b[8 * n + 20] = [-np.inf, 0.0]
    b[2 * n + 33] = [np.nan, 0.0]
    c = dev_gemm(qp, a, b, m, n, k, qp.HP_GEMM_SPLIT_I8)
    ref = O.gemm(a, b, m, n, k)
    assert np.array_equal(np.isnan(c[:, 0]), np.isnan(ref[:, 0]))
    assert np.isnan(c[:, 0]).reshape(m, n)[[3, 10]].all() and np.isnan(c[:, 0]).reshape(m, n)[:, [20, 33]].all()
    fin = ~np.isnan(ref[:, 0])
    a_f, b_f = a.copy(), b.copy()
    a_f[~np.isfinite(a_f)] = 0.0
    b_f[~np.isfinite(b_f)] = 0.0
    ref4 = O.mf4_gemm(to_mf4(a_f), to_mf4(b_f), m, n, k)
    bound = 2.0 ** -100 * split_norm_bound(a_f, b_f, m, n, k)
    assert np.all(abs_err(c, ref4)[fin] <= bound[fin])


def test_split_bound_is_normwise_not_componentwise(qp):
    """A row [1, e] times a column [e, 1] with e = 1e-20: C = 2e-20.  The split result meets its
    normwise bound (2^(1-W) of ~2) but keeps only ~42 bits of C (e is truncated at 2^-108 of
    its row's maximum); the plain kernel keeps dd accuracy componentwise.  This is the
    documented contract of HP_GEMM_SPLIT_I8."""
    m, n, k = 4, 4, 2
    e = np.array([1e-20, 0.0])
    a = np.zeros((m * k, 2))
    b = np.zeros((k * n, 2))
    for i in range(m):
        a[i * k] = [1.0, 0.0]
        a[i * k + 1] = e
    for j in range(n):
        b[j] = e
        b[n + j] = [1.0, 0.0]
    c_split = dev_gemm(qp, a, b, m, n, k, qp.HP_GEMM_SPLIT_I8)
    c_plain = dev_gemm(qp, a, b, m, n, k, qp.HP_GEMM_PLAIN)
    ref = O.mf4_gemm(to_mf4(a), to_mf4(b), m, n, k)
    norm = split_norm_bound(a, b, m, n, k)  # 2^(1-W) of it is the bound; W > 107 at k = 2
    es, ep = abs_err(c_split, ref), abs_err(c_plain, ref)
    assert np.all(es <= 2.0 ** -106 * norm + 2.0 ** -106 * np.abs(ref[:, 0]))
    rel_s, rel_p = np.max(es / np.abs(ref[:, 0])), np.max(ep / np.abs(ref[:, 0]))
    print(f"graded: split componentwise {rel_s:.2e}, plain {rel_p:.2e}")
    assert rel_p <= 2.0 ** -100
    assert rel_s > 2.0 ** -60  # the componentwise loss the normwise bound allows
