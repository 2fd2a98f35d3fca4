"""The oracle's assemble_overlap_mixed over contracted s/p/d/f shells (crates/overlap/src:
screen.rs:57-101, integral.rs:65-346, assemble.rs:35-85), pinned by:

  * the reference's own tests (integral.rs:367-436, assemble.rs:116-164, screen.rs:138-165);
  * mpmath at 50 digits with an independent closed form: each Cartesian primitive pair as
    exp(-mu R^2) times a product of 1D moments, the polynomial factors expanded about the
    Gaussian product centre (binomials, Gamma((k+1)/2) / gamma^((k+1)/2)), not the
    Obara-Saika recurrence;
  * the single-primitive s-shell restatement (hpo_assemble_overlap_s), bit for bit."""
import math

import numpy as np
import pytest

import oracle_lib as O

mp = pytest.importorskip("mpmath")


def norm(alpha):
    return (2.0 * alpha / math.pi) ** 0.75


def cart(l):
    return [(lx, ly, l - lx - ly) for lx in range(l, -1, -1) for ly in range(l - lx, -1, -1)]


def mp_prim(ai, A, la, aj, B, lb, absolute=False):
    """Cartesian primitive overlap <(x-A)^la exp(-ai|r-A|^2) | (x-B)^lb exp(-aj|r-B|^2)>;
    absolute: the same sum with every term taken positive (the scale of the recurrence's
    intermediate terms, which bounds its rounding error)."""
    ai, aj = mp.mpf(ai), mp.mpf(aj)
    g = ai + aj
    mu = ai * aj / g
    A = [mp.mpf(x) for x in A]
    B = [mp.mpf(x) for x in B]
    r2 = sum((A[k] - B[k]) ** 2 for k in range(3))
    out = mp.exp(-mu * r2)
    for d in range(3):
        P = (ai * A[d] + aj * B[d]) / g
        pa, pb = P - A[d], P - B[d]
        s = mp.mpf(0)
        for k in range(la[d] + 1):
            for m in range(lb[d] + 1):
                n = k + m
                if n % 2:
                    continue
                t = (mp.binomial(la[d], k) * mp.binomial(lb[d], m) * pa ** (la[d] - k) * pb ** (lb[d] - m)
                     * mp.gamma(mp.mpf(n + 1) / 2) / g ** (mp.mpf(n + 1) / 2))
                s += abs(t) if absolute else t
        out *= s
    return out


def mp_block(shell_a, shell_b):
    (A, la, pa), (B, lb, pb) = shell_a, shell_b
    ca, cb = cart(la), cart(lb)
    blk = np.zeros((len(ca), len(cb)))
    absb = np.zeros((len(ca), len(cb)))
    envb = np.zeros((len(ca), len(cb)))
    for u, lu in enumerate(ca):
        for v, lv in enumerate(cb):
            tot = mp.mpf(0)
            ab = mp.mpf(0)
            env = mp.mpf(0)
            for ai, ci in pa:
                for aj, cj in pb:
                    term = mp.mpf(ci) * mp.mpf(cj) * mp_prim(ai, A, lu, aj, B, lv)
                    tot += term
                    env += abs(term)
                    ab += abs(mp.mpf(ci) * mp.mpf(cj)) * mp_prim(ai, A, lu, aj, B, lv, absolute=True)
            blk[u, v] = float(tot)
            absb[u, v] = float(ab)
            envb[u, v] = float(env)
    return blk, absb, envb


# ---------------------------------------------------------------- reference tests
def test_self_overlap_normalized():  # integral.rs:367-383
    s, _ = O.assemble_overlap_mixed([((0, 0, 0), 0, [(1.0, norm(1.0))])])
    assert abs(s[0, 0] - 1.0) < 1e-14


def test_f64_f128_agree_tight():  # integral.rs:385-407, via the tier the screen picks plus a forced-diffuse run
    sa = ((0, 0, 0), 0, [(10.0, 0.5), (5.0, 0.5)])
    sb = ((1.0, 0, 0), 0, [(8.0, 0.6), (3.0, 0.4)])
    s64, c64 = O.assemble_overlap_mixed([sa, sb])
    s128, c128 = O.assemble_overlap_mixed([sa, sb], (100.0, 0.0, 1e-14, 12.0))  # every pair diffuse -> F128
    assert c64[1] == 0 and c128[1] == 3
    assert abs(s64[1, 0] - s128[1, 0]) < 1e-14


def test_f128_diffuse_finite_positive():  # integral.rs:409-425
    sh = ((0, 0, 0), 0, [(100.0, 0.001), (0.03, 0.999)])
    s, c = O.assemble_overlap_mixed([sh])
    assert c[1] == 1 and np.isfinite(s[0, 0]) and s[0, 0] > 0


def test_p_shell_block_size():  # integral.rs:427-436
    s, _ = O.assemble_overlap_mixed([((0, 0, 0), 1, [(1.0, 1.0)])])
    assert s.shape == (3, 3)


def h2():
    prim = [(3.42525, 0.15433 * norm(3.42525)), (0.62391, 0.53533 * norm(0.62391)), (0.16886, 0.44463 * norm(0.16886))]
    return [((0, 0, 0), 0, prim), ((1.4, 0, 0), 0, prim)]


def test_h2_overlap_symmetric():  # assemble.rs:116-129
    s, _ = O.assemble_overlap_mixed(h2())
    assert abs(s[0, 1] - s[1, 0]) < 1e-15
    assert abs(s[0, 0] - 1.0) < 0.02 and abs(s[1, 1] - 1.0) < 0.02
    assert 0.5 < s[0, 1] < 0.8


def test_all_tight_uses_f64():  # assemble.rs:131-137
    _, c = O.assemble_overlap_mixed(h2())
    assert c[1] == 0 and c[0] > 0


def test_diffuse_triggers_f128():  # assemble.rs:139-163
    prim = [(50.0, 0.3 * norm(50.0)), (0.04, 0.7 * norm(0.04))]
    _, c = O.assemble_overlap_mixed([((0, 0, 0), 0, prim), ((2.0, 0, 0), 0, prim)])
    assert c[1] > 0


@pytest.mark.parametrize("b_center,kind,expect", [((1.0, 0, 0), "tight", {0}), ((2.0, 0, 0), "diffuse", {1}),
                                                  ((20.0, 0, 0), "diffuse", {0, 2}), ((50.0, 0, 0), "tight", {2})])
def test_screen_tiers(b_center, kind, expect):  # screen.rs:138-165
    prim = [(100.0, 0.5), (10.0, 0.5)] if kind == "tight" else [(50.0, 0.3), (0.04, 0.7)]
    _, c = O.assemble_overlap_mixed([((0, 0, 0), 0, prim), (b_center, 0, prim)])
    # pairs: (0,0), (1,0), (1,1); the cross pair's tier is the one that differs from the self pairs
    self_pairs = O.assemble_overlap_mixed([((0, 0, 0), 0, prim)])[1]
    cross = [k for k in range(3) if c[k] - 2 * self_pairs[k] == 1]
    assert len(cross) == 1 and cross[0] in expect


# ---------------------------------------------------------------- mpmath
def random_shells(seed, count, lmax, diffuse=False):
    rng = np.random.default_rng(seed)
    shells = []
    for i in range(count):
        l = int(rng.integers(0, lmax + 1))
        npr = int(rng.integers(1, 4))
        alphas = 10 ** rng.uniform(-1.5 if diffuse else -0.7, 1.5, npr)
        prims = [(float(a), float(rng.uniform(0.2, 1.0) * norm(a))) for a in alphas]
        shells.append((tuple(rng.uniform(-1.5, 1.5, 3)), l, prims))
    return shells


ALL_DD = (1e300, 0.0, 1e-300, 1e300)  # every shell "diffuse" and near: every pair in the F128 tier


@pytest.mark.parametrize("seed,diffuse", [(1, False), (2, True), (3, True)])
def test_blocks_match_mpmath(seed, diffuse):
    """The dd tier (forced for every pair) to 64 ulps of the term scale; the screened run
    (mostly the f64 tier) to 1e-12 of it: the f64 path forms P - A = a_j (B - A) / gamma by
    a difference that loses up to log2(|A| / |P - A|) bits when one exponent dominates, and
    the recurrence carries that through, which is the reference's own rounding."""
    mp.mp.dps = 50
    shells = random_shells(seed, 6, 3, diffuse)
    s_dd, c_dd = O.assemble_overlap_mixed(shells, ALL_DD)
    s64, counts = O.assemble_overlap_mixed(shells)
    n = len(shells)
    assert c_dd[1] == n * (n + 1) // 2
    if diffuse:
        assert counts[1] > 0
    env_dd = O.assemble_overlap_envelope(shells, ALL_DD)
    env64 = O.assemble_overlap_envelope(shells)
    offs = np.cumsum([0] + [(l + 1) * (l + 2) // 2 for _, l, _ in shells])
    for i in range(n):
        for j in range(i + 1):
            blk, absb, envb = mp_block(shells[i], shells[j])
            floor = absb.max() if i == j else 0.0  # same centre: odd components are rounding noise
            for s, env, ulps in ((s_dd, env_dd, 64 * 2.0 ** -53), (s64, env64, 1e-12)):
                got = s[offs[i]:offs[i + 1], offs[j]:offs[j + 1]]
                bar = ulps * (absb + floor) + 1e-300
                assert np.all(np.abs(got - blk) <= bar), (i, j, ulps, float((np.abs(got - blk) / bar).max()))
                # the envelope: the sum of the magnitudes of the contraction's terms
                e = env[offs[i]:offs[i + 1], offs[j]:offs[j + 1]]
                assert np.all(np.abs(e - envb) <= bar), (i, j, float((np.abs(e - envb) / bar).max()))


def test_single_primitive_s_equals_s_restatement():
    rng = np.random.default_rng(4)
    n = 30
    centers = rng.uniform(-3, 3, (n, 3))
    alpha = 10 ** rng.uniform(-2, 1, n)
    coeff = np.array([norm(a) for a in alpha])
    shells = [(tuple(centers[i]), 0, [(float(alpha[i]), float(coeff[i]))]) for i in range(n)]
    s_full, c_full = O.assemble_overlap_mixed(shells)
    s_s, c_s = O.assemble_overlap_s(centers.ravel(), alpha, coeff)
    assert np.array_equal(s_full.ravel(), np.asarray(s_s).ravel()) and tuple(c_full) == tuple(c_s)


# ---------------------------------------------------------------- per-entry bound
U = 2.0 ** -53
MOLECULES = [(1, 6, False, O.DEFAULT_SCREEN), (2, 10, True, O.DEFAULT_SCREEN), (3, 14, True, (0.3, 50.0, 1e-10, 6.0)),
             (4, 8, True, ALL_DD)]


@pytest.mark.parametrize("seed,atoms,diffuse,params", MOLECULES)
def test_exp_differences_stay_within_the_entry_bound(seed, atoms, diffuse, params):
    """The device assembly differs from this one only through libm's exp (each faithful, so
    they differ by at most 2 ulps).  Moving every exp here by up to 2 ulps changes each
    entry by at most 16 ulps of its envelope (the sum of the magnitudes of its terms): the
    per-entry bar of the GPU test.  The envelope bounds |S| and is zero exactly where S is."""
    from test_gpu_overlap_shells import molecule
    shells = molecule(seed, atoms, diffuse)
    s, _ = O.assemble_overlap_mixed(shells, params)
    env = O.assemble_overlap_envelope(shells, params)
    assert np.all(env * (1 + 8 * U) >= np.abs(s))
    assert np.array_equal(env == 0, s == 0) or np.all(s[env == 0] == 0)
    try:
        O.set_exp_ulp_shift(2)
        s2, _ = O.assemble_overlap_mixed(shells, params)
    finally:
        O.set_exp_ulp_shift(0)
    d = np.abs(s2 - s)
    assert (d > 0).mean() > 0.3  # the shift reaches most entries
    assert np.all(d <= 16 * U * env), float((d / (16 * U * env + 1e-300)).max())


def test_entry_bound_catches_small_entry_errors():
    """A small entry wrong in its fifth digit passes a bar of 8 ulps of max |S| and fails the
    per-entry bar."""
    from test_gpu_overlap_shells import molecule
    shells = molecule(2, 10, True)
    s, _ = O.assemble_overlap_mixed(shells)
    env = O.assemble_overlap_envelope(shells)
    small = np.argwhere((np.abs(s) > 0) & (np.abs(s) < 1e-8 * np.abs(s).max()) & (env < 2 * np.abs(s)))
    assert len(small) > 0
    u, v = small[0]
    bad = s.copy()
    bad[u, v] *= 1 + 1e-5
    d = np.abs(bad - s)
    assert d.max() <= 8 * U * np.abs(s).max()
    assert not np.all(d <= 16 * U * env)
